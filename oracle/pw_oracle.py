"""FP64 CPU oracle for the Kohn-Sham wavefunction hot path (numpy).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module; the product path
(libamdk and amdkiit_amd/) never does.

Each function restates one reference routine with the same arithmetic in
FP64, citing the routine it follows (paths relative to the AMDKIIT source
tree).  numpy's FFT stands in for FFTW: it is checked against a brute-force
DFT on small grids (tests/test_oracle.py), which pins sign, scaling and the
centring convention of fft_interface.f90.

Arrays: coefficients are complex128 (nb, ngpw) in caller order with Miller
indices `mill` (ngpw, 3); real-space grids are (n3, n2, n1), x fastest.
"""
import numpy as np


# ----------------------------------------------------------------------------
# FFT conventions: fft_interface.f90:97-144 (do_fft) and :515-532 (phasen).
# backward (G->r): f(r_j) = sum_G F(G) exp(+i G.r_j), unscaled
# forward  (r->G): F(G)   = (1/N) sum_j f(r_j) exp(-i G.r_j)
def grid_index(mill, n):
    n1, n2, n3 = n
    return (np.mod(mill[:, 2], n3), np.mod(mill[:, 1], n2), np.mod(mill[:, 0], n1))


def fft_backward(grid):
    return np.fft.ifftn(grid) * grid.size


def fft_forward(grid):
    return np.fft.fftn(grid) / grid.size


def pack_pair(c1, c2, mill, n):
    """Two real bands into one complex grid: psi1 + i psi2
    (wfn_initialize.f90:300-307, density.f90:32-48)."""
    g = np.zeros((n[2], n[1], n[0]), np.complex128)
    ip = grid_index(mill, n)
    im = grid_index(-mill, n)
    c2 = np.zeros_like(c1) if c2 is None else c2
    g[im] = np.conj(c1) + 1j * np.conj(c2)
    g[ip] = c1 + 1j * c2  # G = 0 written last, as the reference does
    return g


def gsq(mill, gmat):
    G = np.asarray(mill, np.float64) @ gmat
    return np.einsum("ij,ij->i", G, G)


def vpsi(c0, occ, v_r, mill, n, gmat):
    """Local + kinetic H|psi>: C2 = -(f/2)(|G|^2 c + 2 (V psi)(G)), f = 0 -> 1.
    Follows VPSI (wfn_initialize.f90:268-365)."""
    nb, ng = c0.shape
    c2 = np.zeros_like(c0)
    g2 = gsq(mill, gmat)
    ip = grid_index(mill, n)
    im = grid_index(-mill, n)
    for ia in range(0, nb, 2):
        has2 = ia + 1 < nb
        psy = pack_pair(c0[ia], c0[ia + 1] if has2 else None, mill, n)
        psy = fft_backward(psy)
        psy = v_r * psy
        psy = fft_forward(psy)
        fp = psy[ip] + psy[im]
        fm = psy[ip] - psy[im]
        fi = occ[ia] * 0.5
        fi = 1.0 if fi == 0.0 else fi
        c2[ia] = -fi * (g2 * c0[ia] + (fp.real + 1j * fm.imag))
        if has2:
            fi2 = occ[ia + 1] * 0.5
            fi2 = 1.0 if fi2 == 0.0 else fi2
            c2[ia + 1] = -fi2 * (g2 * c0[ia + 1] + (fp.imag - 1j * fm.real))
    return c2


def eval_density(c0, occ, mill, n, omega):
    """rho(r) = sum_i f_i |psi_i(r)|^2 / Omega (density.f90:5-98, closed shell)."""
    nb = c0.shape[0]
    rho = np.zeros((n[2], n[1], n[0]))
    for ia in range(0, nb, 2):
        has2 = ia + 1 < nb
        psy = fft_backward(pack_pair(c0[ia], c0[ia + 1] if has2 else None, mill, n))
        rho += occ[ia] / omega * psy.real ** 2
        if has2:
            rho += occ[ia + 1] / omega * psy.imag ** 2
    return rho


def dotp(ca, cb):
    """<a|b> over the full sphere from half-sphere coefficients, G=0 first
    (DOTP, wfn_initialize.f90:382-426)."""
    return 2.0 * np.real(np.vdot(ca, cb)) - np.real(ca[0]) * np.real(cb[0])
