"""Pseudopotential tables (host setup, FP64).

Produces, per species, everything the device path consumes:
  V_s(G)   local form factor on the density sphere (already / Omega)
  rho_s(G) Gaussian ion charge, -Z exp(-G^2 r_c^2 / 4) / Omega
  twnl(G)  Kleinman-Bylander projector Y_lm(G^) * w(|G|) / sqrt(Omega)
  D        projector coupling matrix (diagonal KB weights for the CPMD
           numeric format; GTH h_ij blocks otherwise)

Three sources:
* UPF v2 files (UpfSpecies), restating readupf.f90, form_factor_upf
  (exp_igr.f90:142-209) and upf_nlppgrid/upf_nlpro
  (pseudopotential.f90:234-354).
* CPMD "NORMCONSERVING NUMERIC" files (the reference's own test inputs),
  restating readpsp.f90:11-67, exp_igr.f90:218-288 (form_factor),
  pseudopotential.f90:434-529 (nlppgrid), :108-162 (nlpro) with the same
  quadrature (SIMPSN, math.f90:319-343), natural cubic spline on 5000
  points (spline.f90, spline_inter.f90, interpolgrid.f90), spherical Bessel
  functions (bessel_func.f90:3-53) and real Y_lm order
  (spherical_func.f90:1-50).
* Analytic GTH/HGH parameters (Goedecker-Teter-Hutter 1996, Hartwigsen-
  Goedecker-Hutter 1998, Krack 2005) for the synthetic configurations; the
  reference has no reader for these, so they are tabulated here in closed form.
"""
import math

import numpy as np

NSPLN = 5000      # max_parameter_pp.f90 nspln
RAGGIO = 1.2      # system_data_types.f90:72
FOURPI = 4.0 * math.pi


# ----------------------------------------------------------------------------
# numerics shared with the reference
def simpsn_weights(n):
    """Weights of SIMPSN (math.f90:319-343): end-corrected trapezoid whose
    upper correction sits on points N-4..N-1 and which drops point N."""
    w = np.ones(n)
    for k, c in ((0, 109 / 48), (1, -5 / 48), (2, 63 / 48), (3, 49 / 48)):
        w[k] *= c
    for k, c in ((n - 2, 109 / 48), (n - 3, -5 / 48), (n - 4, 63 / 48), (n - 5, 49 / 48)):
        w[k] *= c
    w[n - 1] = 0.0
    return w


def simpsn(f):
    return float(np.dot(simpsn_weights(len(f)), f))


class NaturalSpline:
    """Natural cubic spline (spline.f90: spline_params/spline_thomas/interpolate)."""

    def __init__(self, x, y):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        n = len(x)
        h = np.diff(x)
        b = np.diff(y) / h
        m = n - 2
        d = 2.0 * (h[:-1] + h[1:])
        u = 6.0 * (b[1:] - b[:-1])
        a = h[1:].copy()
        c = h[1:].copy()
        d = d.copy()
        u = u.copy()
        for i in range(1, m):
            s = a[i - 1] / d[i - 1]
            d[i] -= s * c[i - 1]
            u[i] -= s * u[i - 1]
        xs = np.zeros(m)
        xs[m - 1] = u[m - 1] / d[m - 1]
        for i in range(m - 2, -1, -1):
            xs[i] = (u[i] - c[i] * xs[i + 1]) / d[i]
        self.x, self.y, self.h = x, y, h
        self.sdd = np.concatenate([[0.0], xs, [0.0]])

    def __call__(self, xq):
        xq = np.asarray(xq, np.float64)
        x, y, h, s = self.x, self.y, self.h, self.sdd
        # spline_hunt: first i with xq < x(i) -> interval i-1
        ig = np.searchsorted(x, xq, side="right") - 1
        ig = np.clip(ig, 0, len(x) - 2)
        hi = h[ig]
        x0 = x[ig]
        x1 = x[ig + 1]
        return (s[ig + 1] / 6.0 / hi * (xq - x0) ** 3 + s[ig] / 6.0 / hi * (x1 - xq) ** 3
                + (y[ig + 1] / hi - s[ig + 1] / 6.0 * hi) * (xq - x0)
                + (y[ig] / hi - hi / 6.0 * s[ig]) * (x1 - xq))


def sph_bessel(l, x):
    """j_l(x) (bessel_func.f90:3-53): series below 1, upward recurrence above."""
    x = np.atleast_1d(np.asarray(x, np.float64))
    out = np.zeros_like(x)
    small = np.abs(x) < 1e-15
    out[small] = 1.0 if l == 0 else 0.0
    lo = (~small) & (x < 1.0)
    if lo.any():
        xx = x[lo]
        fact = np.ones_like(xx)
        for i in range(1, l + 1):
            fact *= xx / (2 * i + 1)
        jl = np.ones_like(xx)
        jr = np.ones_like(xx)
        for k in range(1, 41):
            jr = -jr * (0.5 * xx * xx) / (k * (2 * (l + k) + 1))
            jl = jl + jr
        out[lo] = jl * fact
    hi = (~small) & (x >= 1.0)
    if hi.any():
        xx = x[hi]
        jl = np.sin(xx) / xx
        jlp = (-np.cos(xx) + jl) / xx
        for i in range(2, l + 2):
            jr = -jl + (2 * i - 1) * jlp / xx
            jl, jlp = jlp, jr
        out[hi] = jl
    return out


def ylm(L, G):
    """Real spherical harmonic number L (1-based, order s, px, pz, py, dxy, dxz,
    dz2, dyz, dx2-y2, f) at directions G (N,3); spherical_func.f90:1-50."""
    G = np.asarray(G, np.float64)
    g2 = np.einsum("ij,ij->i", G, G)
    ok = g2 >= 1e-12
    g2s = np.where(ok, g2, 1.0)
    gx, gy, gz = G[:, 0], G[:, 1], G[:, 2]
    r = np.sqrt(g2s)
    table = {
        1: np.full(len(G), math.sqrt(1 / FOURPI)),
        2: math.sqrt(3 / FOURPI) * gx / r,
        3: math.sqrt(3 / FOURPI) * gz / r,
        4: math.sqrt(3 / FOURPI) * gy / r,
        5: math.sqrt(15 / FOURPI) * gx * gy / g2s,
        6: math.sqrt(15 / FOURPI) * gx * gz / g2s,
        7: math.sqrt(5 / FOURPI / 4) * (3 * gz ** 2 / g2s - 1),
        8: math.sqrt(15 / FOURPI) * gy * gz / g2s,
        9: math.sqrt(15 / FOURPI / 4) * (gx ** 2 - gy ** 2) / g2s,
    }
    y = table[L]
    return np.where(ok, y, math.sqrt(1 / FOURPI) if L == 1 else 0.0)


def erf(x):
    return np.vectorize(math.erf)(x)


# ----------------------------------------------------------------------------
class Species:
    """A pseudopotential ready for tabulation.

    local(g2)      -> V_s(G) * Omega (short-range local part, see module doc)
    projectors     -> list of (l, lm_index, radial(g2) -> 4 pi int r^2 p j_l)
    D              -> coupling matrix over the projector list
    precon_weights -> WSG(is, iv) as the PCG preconditioner reads it
    """

    zv = 0.0
    symbol = "X"

    def local(self, g2):
        raise NotImplementedError

    def projectors(self):
        return []

    def dmat(self):
        return np.zeros((0, 0))

    def precon_weights(self):
        """Per-projector weights of the nonlocal term of hess (cg.f90:20-40):
        WSG(is, iv), the Kleinman-Bylander weights of the CPMD format, i.e. the
        diagonal of D."""
        return np.diag(self.dmat()).copy()


def read_cpmd_psp(path, lmax_read):
    """CPMD numeric pseudopotential (readpsp.f90:11-67): Z, ZV, the
    &POTENTIAL table (r, V_l) and the &WAVEFUNCTION table (r, u_l)."""
    with open(path) as f:
        lines = f.read().split("\n")

    def block(name):
        for i, ln in enumerate(lines):
            if "&" + name in ln:
                return i
        raise ValueError(f"no &{name} block in {path}")

    def key(name, start):
        for ln in lines[start:]:
            if name in ln and "=" in ln:
                return ln.split("=", 1)[1].split()
            if "&END" in ln:
                break
        raise ValueError(f"no {name} in {path}")

    ia = block("ATOM")
    z = float(key("Z ", ia)[0])
    zv = float(key("ZV", ia)[0])
    ip = block("POTENTIAL")
    m = int(lines[ip + 1].split()[0])
    pot = np.array([[float(t) for t in lines[ip + 2 + k].split()[:1 + lmax_read]] for k in range(m)])
    iw = block("WAVEFUNCTION")
    mw = int(lines[iw + 1].split()[0])
    wf = np.array([[float(t) for t in lines[iw + 2 + k].split()[:1 + lmax_read]] for k in range(mw)])
    if mw != m:
        raise ValueError("incompatible meshes")
    return dict(z=z, zv=zv, rr=pot[:, 0], vr=pot[:, 1:], rps=wf[:, 1:], mesh=m)


class CpmdNumericSpecies(Species):
    """Kleinman-Bylander species from a CPMD numeric file.

    lmax: highest channel in the file used (yaml LMAX); kb: PP_SCHEME
    KLEINMAN-BYLANDER.  As in yamlread.f90:300-314 the local channel is the
    highest one for KB species and the only one otherwise.
    """

    def __init__(self, path, symbol, lmax, kb, g2max_pw, g2max_rho):
        self.symbol = symbol
        nl = lmax + 1  # channels read (LMAX+1)
        d = read_cpmd_psp(path, nl)
        self.zv = d["zv"]
        rr = d["rr"]
        m = d["mesh"]
        # logarithmic grid rw (exp_igr.f90:227-231) and interpolation of V_l
        dlog = math.log(rr[-1] / rr[0]) / (m - 1)
        rw = np.empty(m)
        rw[0] = rr[0]
        fac = math.exp(dlog)
        for k in range(1, m):
            rw[k] = rw[k - 1] * fac
        vr = np.stack([NaturalSpline(rr, d["vr"][:, l])(rw) for l in range(nl)], 1)
        l_local = nl - 1 if kb else 0  # 0-based channel index
        vloc = vr[:, l_local]
        zv = self.zv
        ef = erf(rw / RAGGIO)
        fint = rw ** 3 * vloc + zv * ef * rw ** 2
        check = vloc + zv * ef / rw
        fint = np.where(np.abs(check) < 1e-8, 0.0, fint)
        self._rw, self._dlog, self._fint = rw, dlog, fint
        # local form factor on the rho spline grid (exp_igr.f90:245-265)
        ggnh = np.arange(NSPLN) * (g2max_rho / (NSPLN - 1))
        sw = simpsn_weights(m)
        xg = np.sqrt(ggnh)
        xrg = np.outer(xg, rw)
        big = xg > 1e-6
        kern = np.ones_like(xrg)
        kern[big] = np.sin(xrg[big]) / xrg[big]
        voo = dlog * FOURPI * (kern * (fint * sw)).sum(1)
        self._loc = NaturalSpline(ggnh, voo)
        # projectors (pseudopotential.f90:434-529): only the s wavefunction is
        # interpolated onto rw, as the reference does (il = 1,1 loop)
        rps = d["rps"].copy()
        rps[:, 0] = NaturalSpline(rr, d["rps"][:, 0])(rw)
        self._proj = []
        wsg = []
        ggng = np.arange(NSPLN) * (g2max_pw / (NSPLN - 1))
        if kb:
            for l in range(nl):
                if l == l_local:
                    continue
                gnl = vr[:, l] - vloc
                gnl = np.where(np.abs(gnl) < 1e-10, 0.0, gnl)
                w = simpsn(rps[:, l] ** 2 * gnl * rw)
                wk = 1.0 / (dlog * w)
                jl = sph_bessel(l, np.outer(np.sqrt(ggng), rw).reshape(-1)).reshape(NSPLN, m)
                tab = dlog * FOURPI * (jl * (rw ** 2 * rps[:, l] * gnl * sw)).sum(1)
                spl = NaturalSpline(ggng, tab)
                base = 0 if l == 0 else l * l
                for mm in range(2 * l + 1):
                    self._proj.append((l, base + mm + 1, spl))
                    wsg.append(wk)
        self._wsg = np.array(wsg)

    def local(self, g2):
        return self._loc(g2)

    def projectors(self):
        return self._proj

    def dmat(self):
        return np.diag(self._wsg)


class GTHSpecies(Species):
    """Analytic GTH/HGH pseudopotential (separable, dual-space Gaussian).

    params: zion, rloc, [C1..C4], and per l: (r_l, h matrix (n x n)).
    """

    def __init__(self, symbol, zion, rloc, cs, nonlocal_):
        self.symbol = symbol
        self.zv = float(zion)
        self.rloc = float(rloc)
        self.cs = list(cs) + [0.0] * (4 - len(cs))
        self.nonlocal_ = nonlocal_  # list over l of (r_l, h)
        self._proj = []
        blocks = []
        for l, (rl, h) in enumerate(nonlocal_):
            h = np.atleast_2d(np.asarray(h, np.float64))
            if not np.any(h):
                continue
            nproj = h.shape[0]
            base = 0 if l == 0 else l * l
            for mm in range(2 * l + 1):
                for i in range(nproj):
                    self._proj.append((l, base + mm + 1, self._radial(l, i + 1, rl)))
            blocks.append((l, h))
        # D over the projector list: index = (l-block, m, i)
        n = len(self._proj)
        D = np.zeros((n, n))
        k = 0
        for l, h in blocks:
            nproj = h.shape[0]
            for mm in range(2 * l + 1):
                D[k:k + nproj, k:k + nproj] = h
                k += nproj
        self._D = D

    def local(self, g2):
        g2 = np.asarray(g2, np.float64)
        rl = self.rloc
        z = self.zv
        x = g2 * rl * rl
        c1, c2, c3, c4 = self.cs
        gauss = (2 * math.pi) ** 1.5 * rl ** 3 * np.exp(-0.5 * x) * (
            c1 + c2 * (3 - x) + c3 * (15 - 10 * x + x * x) + c4 * (105 - 105 * x + 21 * x * x - x ** 3))
        # -Z erf(r/(sqrt2 rl))/r + Z erf(r/raggio)/r  ->  4 pi Z (e^{-G^2 R^2/4} - e^{-G^2 rl^2/2}) / G^2
        safe = np.where(g2 > 1e-12, g2, 1.0)
        lr = np.where(g2 > 1e-12,
                      FOURPI * z * (np.exp(-0.25 * g2 * RAGGIO ** 2) - np.exp(-0.5 * x)) / safe,
                      FOURPI * z * (0.5 * rl * rl - 0.25 * RAGGIO ** 2))
        return gauss + lr

    @staticmethod
    def _radial(l, i, rl):
        """w(G) = 4 pi int r^2 p_i^l(r) j_l(G r) dr for the normalised GTH
        projector p_i^l = sqrt2 r^(l+2(i-1)) e^{-r^2/2rl^2} / (rl^(l+(4i-1)/2) sqrt(Gamma(l+(4i-1)/2)))."""
        r = np.linspace(0.0, 12.0 * rl + 2.0, 6001)[1:]
        dr = r[1] - r[0]
        p = (math.sqrt(2.0) * r ** (l + 2 * (i - 1)) * np.exp(-0.5 * (r / rl) ** 2)
             / (rl ** (l + (4 * i - 1) / 2.0) * math.sqrt(math.gamma(l + (4 * i - 1) / 2.0))))

        def f(g2):
            g2 = np.asarray(g2, np.float64)
            grid = np.arange(NSPLN) * (max(float(g2.max()), 1e-8) * 1.0001 / (NSPLN - 1))
            tab = np.empty(NSPLN)
            for s0 in range(0, NSPLN, 500):
                gg = np.sqrt(grid[s0:s0 + 500])
                tab[s0:s0 + 500] = FOURPI * (_sph_jl_vec(l, np.outer(gg, r)) * (r * r * p)).sum(1) * dr
            return NaturalSpline(grid, tab)(g2)

        return f

    def projectors(self):
        return self._proj

    def dmat(self):
        return self._D


def read_upf(path):
    """UPF v2 file (the fields readupf.f90:19-425 reads): PP_HEADER
    attributes, PP_R, PP_RAB, PP_LOCAL (Ry), PP_BETA.k with their
    angular_momentum (r * beta(r), Ry^(1/2) units), PP_DIJ (Ry).  Parsed as
    text, as the reference does, not as XML (PP_INFO is free text)."""
    import re
    with open(path) as f:
        txt = f.read()

    def attrs(tag_text):
        return {k: v.strip() for k, v in re.findall(r'(\w+)\s*=\s*"([^"]*)"', tag_text)}

    m = re.search(r"<PP_HEADER(.*?)/>", txt, re.S)
    if m is None:
        raise ValueError(f"{path}: no PP_HEADER (only UPF v2 is read)")
    head = attrs(m.group(1))

    def array(tag):
        mm = re.search(r"<" + re.escape(tag) + r"\b([^>]*)>(.*?)</" + re.escape(tag) + r">", txt, re.S)
        if mm is None:
            raise ValueError(f"{path}: no {tag}")
        return attrs(mm.group(1)), np.array(mm.group(2).split(), np.float64)

    mesh = int(head["mesh_size"])
    r = array("PP_R")[1][:mesh]
    rab = array("PP_RAB")[1][:mesh]
    vloc = array("PP_LOCAL")[1][:mesh]
    nbeta = int(head["number_of_proj"])
    betas = []
    for k in range(1, nbeta + 1):
        a, v = array(f"PP_BETA.{k}")
        betas.append((int(a["angular_momentum"]), v[:mesh]))
    dij = array("PP_DIJ")[1][:nbeta * nbeta].reshape(nbeta, nbeta) if nbeta else np.zeros((0, 0))
    return dict(zv=float(head["z_valence"]), lmax=int(head["l_max"]), mesh=mesh, r=r, rab=rab, vloc=vloc,
                betas=betas, dij=dij)


class UpfSpecies(Species):
    """Norm-conserving UPF v2 species as the reference tabulates it.

    * Local part (exp_igr.f90:142-209, form_factor_upf): V_loc = PP_LOCAL / 2
      (Ry -> Ha, readupf.f90:420), 4 pi SIMPSN of (r^2 V + Z erf(r / 1.2) r)
      sin(Gr)/(Gr) with the PP_RAB weights, on the 5000-point G^2 grid.
    * Projector list (readupf.f90:296-309): iv runs over l, then m, then the
      projectors of that l; lp is the Y_lm index, lf the projector number.
    * Radial tables (pseudopotential.f90:234-291, upf_nlppgrid): for each new
      l the table is 4 pi SIMPSN(rab r beta j_l(Gr)) computed from the file's
      beta function number l + 1 (1-based, in file order), and every further
      projector of the same l copies the table of the first.  For files with
      several projectors per l this is not the projector's own radial
      function; it is restated as written because the recorded energies
      come from it.
    * D (readupf.f90:319-347): hlsg = PP_DIJ / 2 (Ry -> Ha) on the diagonal,
      coupling projectors of the same (l, m); off-diagonal PP_DIJ must vanish.
    """

    def __init__(self, path, symbol, g2max_pw, g2max_rho):
        d = read_upf(path)
        self.symbol = symbol
        self.zv = d["zv"]
        r, rab, m = d["r"], d["rab"], d["mesh"]
        sw = simpsn_weights(m)
        vr = 0.5 * d["vloc"]
        fint = r * r * vr + self.zv * erf(r / RAGGIO) * r
        ggnh = np.arange(NSPLN) * (g2max_rho / (NSPLN - 1))
        xg = np.sqrt(ggnh)
        xrg = np.outer(xg, r)
        big = xg > 1e-6
        kern = np.ones_like(xrg)
        kern[big] = np.sin(xrg[big]) / xrg[big]
        voo = FOURPI * (kern * (fint * rab * sw)).sum(1)
        self._loc = NaturalSpline(ggnh, voo)
        # projector bookkeeping
        nbeta = len(d["betas"])
        npro = {}
        pplist = []  # (l, index within l), 1-based index as the reference
        for l, _ in d["betas"]:
            npro[l] = npro.get(l, 0) + 1
            if npro[l] > 4:
                raise ValueError(f"{path}: more than 4 projectors for l = {l}")
            pplist.append((l, npro[l]))
        dij = d["dij"]
        for i in range(nbeta):
            for j in range(nbeta):
                if i != j and dij[i, j] != 0.0:
                    raise ValueError(f"{path}: non-diagonal PP_DIJ is not supported (readupf.f90:335-340)")
        hlsg = {}
        for i, (l, k) in enumerate(pplist):
            hlsg[(l, k, k)] = 0.5 * dij[i, i]
        lmax1 = d["lmax"] + 1
        ivs = []  # (l, lp, lf)
        lp = 0
        for L in range(1, lmax1 + 1):
            for _ in range(2 * L - 1):
                lp += 1
                for k in range(1, npro.get(L - 1, 0) + 1):
                    ivs.append((L - 1, lp, k))
        ggng = np.arange(NSPLN) * (g2max_pw / (NSPLN - 1))
        xg = np.sqrt(ggng)
        self._proj = []
        lold, spl = -1, None
        for (l, lpv, k) in ivs:
            if l != lold:
                if l >= nbeta:
                    raise ValueError(f"{path}: no beta function number {l + 1} for l = {l}")
                gnl = d["betas"][l][1]  # file's beta number l + 1 (see class note)
                jl = sph_bessel(l, np.outer(xg, r).reshape(-1)).reshape(NSPLN, m)
                tab = FOURPI * (jl * (rab * r * gnl * sw)).sum(1)
                spl = NaturalSpline(ggng, tab)
            lold = l
            self._proj.append((l, lpv, spl))
        n = len(ivs)
        D = np.zeros((n, n))
        for a, (l, lpa, ka) in enumerate(ivs):
            for b, (l2, lpb, kb) in enumerate(ivs):
                if l == l2 and lpa == lpb:
                    D[a, b] = hlsg.get((l, ka, kb), 0.0) if ka == kb else 0.0
        self._D = D
        self.ivs = ivs

    def local(self, g2):
        return self._loc(g2)

    def projectors(self):
        return self._proj

    def dmat(self):
        return self._D

    def precon_weights(self):
        """Zero: hess (cg.f90:20-28) reads WSG(is, iv) for UPF species too,
        but only the CPMD-format reader sets WSG (pseudopotential.f90:480-489),
        so for a UPF species the preconditioner has no nonlocal term (WSG is a
        static module array, system_data_types.f90:78)."""
        return np.zeros(len(self._proj))


def _sph_jl_vec(l, x):
    x = np.asarray(x, np.float64)
    small = np.abs(x) < 1e-8
    xs = np.where(small, 1.0, x)
    if l == 0:
        v = np.sin(xs) / xs
        return np.where(small, 1.0, v)
    if l == 1:
        v = np.sin(xs) / xs ** 2 - np.cos(xs) / xs
        return np.where(small, x / 3.0, v)
    if l == 2:
        v = (3 / xs ** 3 - 1 / xs) * np.sin(xs) - 3 * np.cos(xs) / xs ** 2
        return np.where(small, x * x / 15.0, v)
    raise ValueError("l > 2 not tabulated")


# Published parameter sets (a.u.).  LDA: Goedecker-Teter-Hutter 1996 / HGH
# 1998; PBE: Krack 2005.  h12 for Si s follows the HGH relation
# h12 = -1/2 sqrt(3/5) h22.
def gth(symbol, functional):
    f = functional.upper()
    if symbol == "H":
        if f == "PBE":
            return GTHSpecies("H", 1, 0.2, [-4.17890044, 0.72446331], [])
        return GTHSpecies("H", 1, 0.2, [-4.180237, 0.725075], [])
    if symbol == "O":
        if f == "PBE":
            return GTHSpecies("O", 6, 0.24455430, [-16.66721480, 2.48731132],
                              [(0.22095592, [[18.33745811]]), (0.21133247, [[0.0]])])
        return GTHSpecies("O", 6, 0.247621, [-16.580318, 2.395701],
                          [(0.221786, [[18.266917]]), (0.256829, [[0.0]])])
    if symbol == "Si":
        h11, h22 = 5.906928, -1.261894
        h12 = -0.5 * math.sqrt(3.0 / 5.0) * h22
        return GTHSpecies("Si", 4, 0.44, [-7.336103],
                          [(0.422738, [[h11, h12], [h12, h22]]), (0.484278, [[2.727013]])])
    raise ValueError(f"no GTH parameters for {symbol}")


# ----------------------------------------------------------------------------
class Tables:
    """Per-species G tables on the caller's sphere order (what the C ABI takes)."""

    def __init__(self, pw, species):
        self.species = species
        self.nsp = len(species)
        omega = pw.cell.omega
        g2r = pw.g2_rho
        g2p = pw.g2_pw
        Gp = pw.cell.gvec(pw.mill_pw)
        self.zv = np.array([s.zv for s in species])
        self.vps = np.stack([s.local(g2r) / omega for s in species])
        self.rhos = np.stack([-s.zv * np.exp(-0.25 * RAGGIO ** 2 * g2r) / omega for s in species])
        self.maxproj = max([len(s.projectors()) for s in species] + [1])
        self.nproj = np.array([len(s.projectors()) for s in species], np.int32)
        self.twnl = np.zeros((self.nsp, self.maxproj, pw.ngpw))
        self.proj_l = np.zeros((self.nsp, self.maxproj), np.int32)
        self.dmat = np.zeros((self.nsp, self.maxproj, self.maxproj))
        for isp, s in enumerate(species):
            for k, (l, lm, rad) in enumerate(s.projectors()):
                self.twnl[isp, k] = ylm(lm, Gp) * rad(g2p) / math.sqrt(omega)
                self.proj_l[isp, k] = l
            D = s.dmat()
            self.dmat[isp, :D.shape[0], :D.shape[1]] = D
        # PCG preconditioner weights WSG(is, iv) (hess, cg.f90:20-40)
        self.wsg = np.zeros((self.nsp, self.maxproj))
        for isp, s in enumerate(species):
            w = s.precon_weights()
            self.wsg[isp, :len(w)] = w

