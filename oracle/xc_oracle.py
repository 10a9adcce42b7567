"""FP64 exchange-correlation oracle (TEST INFRASTRUCTURE ONLY).

Energy densities f(rho, sigma) = rho * eps_xc are written once from the
published formulas; potentials v_rho = df/drho and v_sigma = df/dsigma are
taken by complex-step differentiation, which is exact to rounding for these
analytic expressions and independent of the hand-derived derivatives in the
HIP kernels.  The low-density smoothing is the reference's (xc.f90:100-133,
gc_cutoff = 1e-8 from constants.f90:10).

Functionals (libxc names used by the reference's input):
  XC_LDA_XC_TETER93  Goedecker-Teter-Hutter 1996 Pade fit (unpolarised)
  XC_LDA_X + XC_LDA_C_PZ   Slater exchange + Perdew-Zunger 1981 (CA data)
  XC_GGA_X_PBE + XC_GGA_C_PBE  Perdew-Burke-Ernzerhof 1996, PW92 correlation
    with libxc's PW_MOD constants.
"""
import numpy as np

GC_CUTOFF = 1.0e-8


def _rs(rho):
    return (3.0 / (4.0 * np.pi * rho)) ** (1.0 / 3.0)


def teter93(rho, sigma=None):
    a = (0.4581652932831429, 2.217058676663745, 0.7405551735357053, 0.01968227878617998)
    b = (1.0, 4.504130959426697, 1.110667363742916, 0.02359291751427506)
    rs = _rs(rho)
    num = a[0] + rs * (a[1] + rs * (a[2] + rs * a[3]))
    den = rs * (b[0] + rs * (b[1] + rs * (b[2] + rs * b[3])))
    return rho * (-num / den)


def lda_x(rho):
    return rho * (-0.75 * (3.0 / np.pi) ** (1.0 / 3.0) * rho ** (1.0 / 3.0))


def pz81_c(rho):
    rs = _rs(rho)
    g, b1, b2 = -0.1423, 1.0529, 0.3334
    A, B, C, D = 0.0311, -0.048, 0.0020, -0.0116
    hi = g / (1.0 + b1 * np.sqrt(rs) + b2 * rs)
    lo = A * np.log(rs) + B + C * rs * np.log(rs) + D * rs
    return rho * np.where(np.real(rs) >= 1.0, hi, lo)


def pz81(rho, sigma=None):
    return lda_x(rho) + pz81_c(rho)


# Perdew-Wang 1992 unpolarised correlation, G(rs; A, alpha1, beta1..4), p = 1
PW92 = dict(A=0.0310907, a1=0.21370, b1=7.5957, b2=3.5876, b3=1.6382, b4=0.49294)
# PBE 1996
PBE_KAPPA = 0.804
PBE_BETA = 0.06672455060314922
PBE_MU = 0.2195149727645171
PBE_GAMMA = (1.0 - np.log(2.0)) / np.pi ** 2


def pw92_c(rs):
    A, a1, b1, b2, b3, b4 = (PW92[k] for k in ("A", "a1", "b1", "b2", "b3", "b4"))
    den = 2.0 * A * (b1 * np.sqrt(rs) + b2 * rs + b3 * rs ** 1.5 + b4 * rs * rs)
    return -2.0 * A * (1.0 + a1 * rs) * np.log(1.0 + 1.0 / den)


def pbe_fx(s2):
    """PBE exchange enhancement F_x(s^2)."""
    return 1.0 + PBE_KAPPA - PBE_KAPPA / (1.0 + PBE_MU * s2 / PBE_KAPPA)


def pbe_h(ec, t2):
    """PBE gradient correction H(rs, t^2) of the correlation (phi = 1), given
    eps_c^PW92(rs)."""
    beta, gamma = PBE_BETA, PBE_GAMMA
    Ac = (beta / gamma) / (np.exp(-ec / gamma) - 1.0)
    at2 = Ac * t2
    return gamma * np.log(1.0 + (beta / gamma) * t2 * (1.0 + at2) / (1.0 + at2 + at2 * at2))


def pbe(rho, sigma):
    kf = (3.0 * np.pi ** 2 * rho) ** (1.0 / 3.0)
    ex_lda = -0.75 * (3.0 / np.pi) ** (1.0 / 3.0) * rho ** (1.0 / 3.0)
    s2 = sigma / (4.0 * kf * kf * rho * rho)
    rs = _rs(rho)
    ec = pw92_c(rs)
    ks2 = 4.0 * kf / np.pi
    t2 = sigma / (4.0 * ks2 * rho * rho)
    return rho * (ex_lda * pbe_fx(s2) + ec + pbe_h(ec, t2))


FUNCTIONALS = {"TETER93": (teter93, False), "PZ": (pz81, False), "PBE": (pbe, True)}


def evaluate(name, rho, sigma):
    """f, v_rho, v_sigma on the grid (complex step, h = 1e-20 relative)."""
    f, gga = FUNCTIONALS[name]
    rho = np.asarray(rho, np.float64)
    r = np.where(rho > 1e-30, rho, 1e-30)
    s = np.where(sigma > 1e-60, sigma, 1e-60) if gga else None
    hr = 1e-20 * r
    if gga:
        fv = np.real(f(r + 0j, s + 0j))
        vr = np.imag(f(r + 1j * hr, s + 0j)) / hr
        hs = 1e-20 * s
        vs = np.imag(f(r + 0j, s + 1j * hs)) / hs
    else:
        fv = np.real(f(r + 0j))
        vr = np.imag(f(r + 1j * hr)) / hr
        vs = np.zeros_like(r)
    return fv, vr, vs


def smooth(rho):
    """smoo, dsmoo of xc.f90:100-118 (rho clipped at 0)."""
    r = np.maximum(rho, 0.0)
    smoo = np.zeros_like(r)
    dsmoo = np.zeros_like(r)
    act = r > 0.1 * GC_CUTOFF
    full = r > 4.0 * GC_CUTOFF
    smoo[full] = 1.0
    mid = act & ~full
    texp = np.exp(3.0 * (1.0 - r[mid] / GC_CUTOFF))
    sm = 1.0 / (1.0 + texp)
    smoo[mid] = sm
    dsmoo[mid] = 3.0 / GC_CUTOFF * texp * sm * sm
    return smoo, dsmoo, act
