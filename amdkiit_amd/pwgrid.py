"""Cell, FFT grid and plane-wave spheres (host side, FP64).

Host restatement of the reference's setup rules, which are not on the hot
path but fix every size the kernels see:

* grid size:  n = NINT(L/pi * sqrt(E_rho) + 0.5), then the smallest n' >= n
  that is even and has only the factors 2, 3, 5
  (set_cell.f90:177-212 set_r_grids, fft_support.f90:50-118 optimal_fft_order);
* cutoffs:    E_rho = 4 E_cut (Ry), G^2 [bohr^-2] < E (Ry), strict inequality
  as in the packing loop (init_module.f90:38,54-55, gvectors.f90:229-236);
* half sphere: m1 > 0, or (m1 = 0, m2 > 0), or (m1 = m2 = 0, m3 >= 0)
  (gvectors.f90:27-41), ordered by |G|^2 (gvectors.f90:248,514-564).

The plane-wave order produced here is the *caller* order handed to the C ABI;
the library builds its own device order from the Miller indices.
"""
import numpy as np

ANG2BOHR = 1.0 / 0.529177210859  # constants.f90:17


def _permitted(n):
    m = n
    for f in (2, 3, 5):
        while m % f == 0:
            m //= f
    return m == 1


def optimal_fft_order(n, order=2):
    """Smallest n' >= n that is a multiple of `order` with factors 2,3,5 only."""
    while not (_permitted(n) and n % order == 0):
        n += 1
    return n


def grid_size(length_bohr, ecut_rho_ry):
    n = int(np.floor(length_bohr / np.pi * np.sqrt(ecut_rho_ry) + 0.5 + 0.5))
    return optimal_fft_order(n, 2)


class Cell:
    """Lattice vectors as columns of `a` (bohr).  `symmetry` is the
    reference's lattice code (set_cell.f90:17-42: 1 cubic, 2 FCC, 3 BCC, ...),
    which selects the minimum-image rule of the ion-ion term
    (apply_pbc.f90); 0 means "not given": cubic when the cell is, otherwise
    the nearest image in fractional coordinates."""

    def __init__(self, a, symmetry=0):
        self.a = np.array(a, dtype=np.float64).reshape(3, 3)
        self.omega = abs(np.linalg.det(self.a))
        # g_j = 2 pi * row j of inv(A)
        self.g = 2.0 * np.pi * np.linalg.inv(self.a)  # rows are g1, g2, g3
        self.symmetry = int(symmetry)

    @classmethod
    def cubic(cls, length):
        return cls(np.eye(3) * length, symmetry=1)

    @classmethod
    def fcc(cls, a):
        """Face-centred cubic with conventional constant a (bohr): the
        primitive vectors of set_vec (set_cell.f90:49-52) as columns."""
        h = a / 2.0
        return cls(np.array([[-h, 0.0, h], [0.0, h, h], [-h, h, 0.0]]).T, symmetry=2)

    @property
    def is_cubic(self):
        L = np.diag(self.a)
        return bool(np.allclose(self.a, np.diag(L), rtol=0, atol=1e-12) and np.allclose(L, L[0], rtol=0, atol=1e-12))

    def pbc_mode(self):
        """Minimum-image rule: 1 cubic box, 2 FCC Wigner-Seitz, 3 fractional."""
        if self.symmetry == 2:
            return 2
        if self.symmetry in (0, 1) and self.is_cubic:
            return 1
        return 3

    def gvec(self, mill):
        """Cartesian G (bohr^-1) for Miller indices (N,3)."""
        return np.asarray(mill, dtype=np.float64) @ self.g


def half_sphere(cell, n, ecut_ry):
    """Half sphere |G|^2 < ecut_ry (bohr^-2), reference convention and order."""
    n1, n2, n3 = n
    # bounding box
    gl = np.linalg.norm(cell.g, axis=1)
    mx = [int(np.ceil(np.sqrt(ecut_ry) / np.min(gl) * 1.0)) + 1] * 3
    mx = [min(mx[0], n1 - 1), min(mx[1], n2 - 1), min(mx[2], n3 - 1)]
    r1 = np.arange(0, mx[0] + 1)
    r2 = np.arange(-mx[1], mx[1] + 1)
    r3 = np.arange(-mx[2], mx[2] + 1)
    M = np.stack(np.meshgrid(r1, r2, r3, indexing="ij"), -1).reshape(-1, 3)
    keep = (M[:, 0] > 0) | ((M[:, 0] == 0) & (M[:, 1] > 0)) | ((M[:, 0] == 0) & (M[:, 1] == 0) & (M[:, 2] >= 0))
    M = M[keep]
    G = cell.gvec(M)
    g2 = np.einsum("ij,ij->i", G, G)
    M = M[g2 < ecut_ry]
    g2 = g2[g2 < ecut_ry]
    order = np.lexsort((M[:, 2], M[:, 1], M[:, 0], np.round(g2, 10)))
    return np.ascontiguousarray(M[order].astype(np.int32))


class PWSetup:
    """Everything the C ABI needs about the basis: grid, spheres, |G|^2."""

    def __init__(self, cell, ecut_ry, n=None):
        self.cell = cell
        self.ecut = float(ecut_ry)
        self.ecut_rho = 4.0 * self.ecut
        if n is None:
            L = np.linalg.norm(cell.a, axis=0)
            n = tuple(grid_size(L[i], self.ecut_rho) for i in range(3))
        self.n = tuple(int(x) for x in n)
        self.mill_rho = half_sphere(cell, self.n, self.ecut_rho)
        g2 = self.g2_of(self.mill_rho)
        self.mill_pw = np.ascontiguousarray(self.mill_rho[g2 < self.ecut])
        self.ngpw = len(self.mill_pw)
        self.ngrho = len(self.mill_rho)
        self.nr = self.n[0] * self.n[1] * self.n[2]

    def g2_of(self, mill):
        G = self.cell.gvec(mill)
        return np.einsum("ij,ij->i", G, G)

    @property
    def g2_pw(self):
        return self.g2_of(self.mill_pw)

    @property
    def g2_rho(self):
        return self.g2_of(self.mill_rho)

    def rays(self, mill):
        """Number of distinct (m2, m3) x-rays (full sphere), as the reference counts."""
        full = np.concatenate([mill, -mill])
        return len(np.unique(full[:, 1:], axis=0))
