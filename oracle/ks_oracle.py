"""FP64 restatement of one Kohn-Sham SCF step of the reference (numpy).

TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline leg).  Follows the call order of wfn_opt (do_opt.f90:80-129)
and each routine it calls; every method cites the routine it restates.

Conventions: coefficient arrays are complex128 (nb, ngpw) in the caller
order of PWSetup (G = 0 first); density-sphere arrays are complex128
(ngrho,); grids are (n3, n2, n1).  Forces are -dE/dR (Ha/bohr).
"""
import math

import numpy as np

from . import pw_oracle as P
from . import xc_oracle as X

RAGGIO = 1.2


def min_image(d, cell):
    """Minimum-image difference vector of apply_pbc (apply_pbc.f90:1-60),
    m = 1: case 1 (simple cubic) folds each component into (-A/2, A/2];
    case 2 (FCC, A the conventional constant) rotates by 45 degrees about z
    into the tetragonal frame of the FCC lattice, folds into its box and
    moves by the body-centre vector when outside |x| + |y| + sqrt2 |z| <
    2 m_a; cases 3 and above take the nearest image in fractional
    coordinates (s -= NINT(s))."""
    mode = cell.pbc_mode()
    d = np.array(d, np.float64)
    if mode == 1:
        A = cell.a[0, 0]
        d = np.fmod(d, A)
        return d - A * np.trunc(d / (A / 2.0))
    if mode == 2:
        A = math.sqrt(2.0) * np.linalg.norm(cell.a[:, 0])
        s2 = math.sqrt(2.0)
        ma = A * s2 / 4.0
        x, y, z = (d[0] + d[1]) / s2, (-d[0] + d[1]) / s2, d[2]
        x, y, z = math.fmod(x, 2 * ma), math.fmod(y, 2 * ma), math.fmod(z, 2 * ma * s2)
        x -= 2 * ma * math.trunc(x / ma)
        y -= 2 * ma * math.trunc(y / ma)
        z -= 2 * ma * s2 * math.trunc(z / (s2 * ma))
        if not (abs(x) + abs(y) + s2 * abs(z) < 2 * ma):
            x -= math.copysign(ma, x)
            y -= math.copysign(ma, y)
            z -= math.copysign(s2 * ma, z)
        return np.array([(x - y) / s2, (x + y) / s2, z])
    s = np.linalg.solve(cell.a, d)
    s = s - np.sign(s) * np.floor(np.abs(s) + 0.5)  # NINT: halves away from zero
    return cell.a @ s


class KSOracle:
    def __init__(self, pw, tables, sp_of_atom, tau, xc="TETER93", workers=1):
        self.pw = pw
        self.T = tables
        self.sp = np.asarray(sp_of_atom, int)
        self.xc = xc
        self.workers = workers
        self.omega = pw.cell.omega
        self.Gp = pw.cell.gvec(pw.mill_pw)
        self.Gr = pw.cell.gvec(pw.mill_rho)
        self.g2p = pw.g2_pw
        self.g2r = pw.g2_rho
        self.ip = P.grid_index(pw.mill_pw, pw.n)
        self.im = P.grid_index(-pw.mill_pw, pw.n)
        self.irp = P.grid_index(pw.mill_rho, pw.n)
        self.irm = P.grid_index(-pw.mill_rho, pw.n)
        self.set_positions(tau)
        self.hnm1 = None

    # -- FFT helpers ------------------------------------------------------------
    def _bwd(self, g):
        import scipy.fft as F
        return F.ifftn(g, workers=self.workers) * g.size

    def _fwd(self, g):
        import scipy.fft as F
        return F.fftn(g, workers=self.workers) / g.size

    def _sphere_to_grid(self, a, b=None):
        """Two real fields given on the density half sphere -> a(r) + i b(r)."""
        n = self.pw.n
        g = np.zeros((n[2], n[1], n[0]), np.complex128)
        b = np.zeros_like(a) if b is None else b
        g[self.irm] = np.conj(a) + 1j * np.conj(b)
        g[self.irp] = a + 1j * b
        return self._bwd(g)

    def _grid_to_sphere2(self, f):
        """FFT of a(r) + i b(r) -> (A(G), B(G)) on the density half sphere
        (unpacking as xc.f90:268-276)."""
        F = self._fwd(f)
        fa = F[self.irp] + F[self.irm]
        fb = F[self.irp] - F[self.irm]
        A = 0.5 * (fa.real + 1j * fb.imag)
        B = 0.5 * (fa.imag - 1j * fb.real)
        return A, B

    # -- structure factors: exp_igrxfactor (exp_igr.f90:4-140) ------------------
    def set_positions(self, tau):
        self.tau = np.asarray(tau, np.float64).reshape(-1, 3)
        self.eigr = np.exp(-1j * (self.Gr @ self.tau.T))          # (ngrho, nat)
        self.eigr_pw = self.eigr[: self.pw.ngpw]
        # rho sphere starts with the psi sphere in PWSetup order
        assert np.array_equal(self.pw.mill_rho[: self.pw.ngpw], self.pw.mill_pw)
        vps = self.T.vps[self.sp]                                  # (nat, ngrho)
        rhos = self.T.rhos[self.sp]
        self.eigrxvps = np.einsum("ga,ag->g", self.eigr, vps)
        self.eigrxrhos = np.einsum("ga,ag->g", self.eigr, rhos)
        # projectors beta_{a,k}(G) = (-i)^l twnl e^{-iGR}, G=0 entry as stored
        rows = []
        self.proj_index = []
        for a in range(len(self.tau)):
            s = self.sp[a]
            for k in range(self.T.nproj[s]):
                l = self.T.proj_l[s, k]
                rows.append(((-1j) ** l) * self.T.twnl[s, k] * self.eigr_pw[:, a])
                self.proj_index.append((a, s, k))
        self.beta = np.array(rows) if rows else np.zeros((0, self.pw.ngpw), complex)

    # -- overlaps ---------------------------------------------------------------
    @staticmethod
    def ovlap(ca, cb):
        """OVLAP_D (wfn_initialize.f90:234-263): 2 Re(A^H B) - Re a0 Re b0."""
        return 2.0 * np.real(np.conj(ca) @ cb.T) - np.outer(ca[:, 0].real, cb[:, 0].real)

    # -- projections: fnl (pseudopotential.f90:164-226) --------------------------
    def fnl(self, c0):
        return self.ovlap(self.beta, c0)  # (nproj_total, nb)

    def dcoef(self):
        """Block-diagonal D over all atom projectors."""
        n = len(self.proj_index)
        D = np.zeros((n, n))
        for i, (a, s, k) in enumerate(self.proj_index):
            for j, (b, t, q) in enumerate(self.proj_index):
                if a == b:
                    D[i, j] = self.T.dmat[s, k, q]
        return D

    def fnonloc(self, c2, nl, occ):
        """FNONLOC (wfn_initialize.f90:560-612): C2 += -f~ sum beta D nl."""
        if len(self.beta) == 0:
            return c2
        f = np.where(occ < 1e-5, 1.0, occ)
        q = self.dcoef() @ nl * (-f)[None, :]
        return c2 + (q.T @ self.beta)

    def e_nl(self, nl, occ):
        """E_nl = sum_i f_i nl_i^T D nl_i (pp_energy, pseudopotential.f90:52-97)."""
        if len(self.beta) == 0:
            return 0.0
        return float(np.einsum("pb,pq,qb,b->", nl, self.dcoef(), nl, occ))

    # -- density, potential, XC ---------------------------------------------------
    def density(self, c0, occ):
        """eval_density (density.f90:5-98)."""
        rho = np.zeros((self.pw.n[2], self.pw.n[1], self.pw.n[0]))
        for ia in range(0, len(c0), 2):
            has2 = ia + 1 < len(c0)
            psy = self._bwd(P.pack_pair(c0[ia], c0[ia + 1] if has2 else None, self.pw.mill_pw, self.pw.n))
            rho += occ[ia] / self.omega * psy.real ** 2
            if has2:
                rho += occ[ia + 1] / self.omega * psy.imag ** 2
        return rho

    def rhor2g(self, rho):
        """rhor2g (density.f90:128-150)."""
        return self._fwd(rho.astype(np.complex128))[self.irp]

    def vks(self, rho, rho_g):
        """eval_pot (potential.f90:4-46) + cal_vpot_ex (xc.f90:4-434).
        Returns V_KS(r) and E_xc."""
        g2 = self.g2r
        vg = np.zeros(self.pw.ngrho, np.complex128)
        nz = g2 > 1e-12
        vg[nz] = 4.0 * np.pi / g2[nz] * (rho_g[nz] + self.eigrxrhos[nz])
        vg += self.eigrxvps
        vg[~nz] = 0.0
        # gradient of rho from rho(G) (GRAD_CAL, xc.f90:436-497)
        G = self.Gr
        gx = self._sphere_to_grid(1j * G[:, 0] * rho_g, 1j * G[:, 1] * rho_g)
        gz = self._sphere_to_grid(1j * G[:, 2] * rho_g).real
        drx, dry, drz = gx.real, gx.imag, gz
        sigma = drx ** 2 + dry ** 2 + drz ** 2
        fv, vr, vs = X.evaluate(self.xc, rho, sigma)
        smoo, dsmoo, act = X.smooth(rho)
        v1 = np.where(act, dsmoo * fv + smoo * vr, 0.0)
        v2 = np.where(act, smoo * 2.0 * vs, 0.0)
        exc = self.omega / self.pw.nr * float(np.sum(np.where(act, smoo * fv, 0.0)))
        # V_xc(G) = v1(G) - i G . (v2 grad rho)(G)
        A, B = self._grid_to_sphere2(v1 + 1j * v2 * drx)
        vg = vg + A - 1j * G[:, 0] * B
        A, B = self._grid_to_sphere2(v2 * dry + 1j * v2 * drz)
        vg = vg - 1j * G[:, 1] * A - 1j * G[:, 2] * B
        v_r = self._sphere_to_grid(vg).real
        return v_r, exc

    # -- energies (total_energy.f90, pseudopotential.f90:3-50) --------------------
    def energies(self, c0, occ, rho_g, nl, exc):
        g2 = self.g2r
        nz = g2 > 1e-12
        rt = rho_g + self.eigrxrhos
        eh = float(np.sum(4.0 * np.pi / g2[nz] * np.abs(rt[nz]) ** 2))
        zv = self.T.zv[self.sp]
        eself = float(np.sum(zv * zv)) / RAGGIO / math.sqrt(2.0 * math.pi)
        esr, _ = self.e_sr()
        estat = eh * self.omega + esr - eself
        eloc = 2.0 * self.omega * float(np.real(0.5 * self.eigrxvps[0] * np.conj(rho_g[0])
                                                + np.sum(self.eigrxvps[1:] * np.conj(rho_g[1:]))))
        ekin = float(np.sum(occ[:, None] * self.g2p[None, :] * np.abs(c0) ** 2))
        enl = self.e_nl(nl, occ)
        etot = enl + eloc + ekin + exc + estat
        return dict(etot=etot, ekin=ekin, estat=estat, eself=eself, esr=esr, eloc=eloc, enl=enl, exc=exc, eh=eh)

    def e_sr(self):
        """Short-range ion-ion energy and forces, minimum image, nlp = 0
        (e_sr, total_energy.f90:133-239; min_image restates apply_pbc)."""
        tau = self.tau
        zv = self.T.zv[self.sp]
        cell = self.pw.cell
        rckj = math.sqrt(2.0) * RAGGIO
        esr = 0.0
        f = np.zeros_like(tau)
        nat = len(tau)
        for i in range(nat):
            for j in range(i + 1, nat):
                d = min_image(tau[i] - tau[j], cell)
                r2 = float(d @ d)
                r = math.sqrt(r2)
                arg = r / rckj
                if arg > 20.0:
                    continue
                z2 = zv[i] * zv[j]
                add = z2 * math.erfc(arg) / r
                esr += add
                pre = 2.0 * z2 / math.sqrt(math.pi) * math.exp(-arg * arg) / rckj
                rep = (add + pre) / r2
                f[i] += rep * d
                f[j] -= rep * d
        return esr, f

    # -- gradient and optimizer --------------------------------------------------
    def hpsi(self, c0, occ, v_r, nl):
        """C2 = VPSI + FNONLOC (= -f H psi)."""
        c2 = P.vpsi(c0, occ, v_r, self.pw.mill_pw, self.pw.n, self.pw.cell.g)
        return self.fnonloc(c2, nl, occ)

    def lagrange(self, c0, c2, occ):
        """forces + eval_GMAX (gradient.f90:5-61)."""
        fsc = (occ != 0.0).astype(np.float64)
        gam = self.ovlap(c2, c0) * np.outer(fsc, fsc)   # GAM(i,j) = <c2_i|c0_j>
        c2 = c2 - gam.T @ c0
        c2[:, 0] = c2[:, 0].real
        nocc = int(np.sum(occ > 1e-5))
        gemax = float(np.abs(c2[:nocc]).max())
        cnorm = sum(P.dotp(c2[i], c2[i]) for i in range(nocc))
        cnorm = math.sqrt(cnorm / (nocc * self.pw.ngpw))
        return c2, gemax, cnorm

    def hess(self):
        """Diagonal preconditioner K(G)^-1 (hess, cg.f90:3-61), HCUT = 0.5;
        the nonlocal term weighs twnl^2 by WSG(is, iv) (cg.f90:26,38), which is
        zero for UPF species (only the CPMD-format reader sets it)."""
        vpp = np.full(self.pw.ngpw, float(np.real(self.eigrxvps[0])))
        natsp = np.bincount(self.sp, minlength=self.T.nsp)
        for s in range(self.T.nsp):
            for k in range(self.T.nproj[s]):
                vpp += natsp[s] * self.T.wsg[s, k] * self.T.twnl[s, k] ** 2
        vpp = 0.5 * self.g2p + vpp
        vpp = np.maximum(vpp, 0.5)
        return 1.0 / vpp

    def pcg(self, c0, c2, occ, istep):
        """PCG (cg.f90:63-145): preconditioned fixed-step descent with momentum.
        `istep` is 1-based within one wfn_opt call."""
        fi = np.where(occ > 0.01, 1.0 / np.where(occ > 0.01, occ, 1.0), 1.0)
        p = fi[:, None] * self.hess()[None, :] * c2
        nocc = int(np.sum(occ > 1e-5))
        ggnorm = sum(P.dotp(p[i], p[i]) for i in range(nocc))
        step = 5.0 * 5.0 / 400.0
        if istep == 1:
            self.ghist = [ggnorm]
            self.hnm1 = p.copy()
        else:
            forget = 0.60 if istep <= 10 else (0.91 if istep <= 20 else 0.95)
            gamma = forget * ggnorm / self.ghist[-1]
            self.ghist.append(ggnorm)
            p = p + gamma * self.hnm1
            self.hnm1 = p.copy()
        return c0 + step * self.hnm1

    @staticmethod
    def ortho(c):
        """ortho_gs (wfn_initialize.f90:448-499): Cholesky S = U^T U, C <- C U^-1,
        then Im c(G=0) = 0 (ZCLEAN)."""
        s = KSOracle.ovlap(c, c)
        u = np.linalg.cholesky(s).T          # upper, S = U^T U
        uinv = np.linalg.inv(u)
        c = uinv.T @ c                       # rows are bands: C U^-1
        c[:, 0] = c[:, 0].real
        return c

    # -- one SCF step (do_opt.f90:80-129) ---------------------------------------
    def scf_step(self, c0, occ, istep):
        rho = self.density(c0, occ)
        rho_g = self.rhor2g(rho)
        v_r, exc = self.vks(rho, rho_g)
        nl = self.fnl(c0)
        e = self.energies(c0, occ, rho_g, nl, exc)
        c2 = self.hpsi(c0, occ, v_r, nl)
        c2, gemax, cnorm = self.lagrange(c0, c2, occ)
        c0 = self.pcg(c0, c2, occ, istep)
        c0 = self.ortho(c0)
        e.update(gemax=gemax, cnorm=cnorm, nelec=float(rho.sum() * self.omega / self.pw.nr))
        return c0, e

    # -- forces: POTFOR + NLFOR + e_sr (nuclear_grad.f90) -----------------------
    def forces(self, c0, occ, rho_g, nl):
        G = self.Gr
        g2 = self.g2r
        nz = np.arange(self.pw.ngrho) > 0
        rt = rho_g + self.eigrxrhos
        vcgs = np.zeros_like(rt)
        vcgs[nz] = 4.0 * np.pi / g2[nz] * np.conj(rt[nz])
        f = np.zeros_like(self.tau)
        for a in range(len(self.tau)):
            s = self.sp[a]
            t = (self.T.rhos[s] * vcgs + self.T.vps[s] * np.conj(rho_g)) * self.eigr[:, a]
            t[~nz] = 0.0
            f[a] = 2.0 * self.omega * np.real((1j * G * t[:, None]).sum(0))
        # nonlocal: dfnl_k = <d beta / dR_k | psi> with d beta/dR = -iG beta
        if len(self.beta):
            D = self.dcoef()
            for k in range(3):
                dbeta = -1j * self.Gp[:, k][None, :] * self.beta
                dfnl = self.ovlap(dbeta, c0)
                # F = -dE/dR,  E_nl = sum_i f_i nl^T D nl
                contrib = -2.0 * np.einsum("pb,pq,qb,b->p", dfnl, D, nl, occ)
                for i, (a, s, kk) in enumerate(self.proj_index):
                    f[a, k] += contrib[i]
        _, fsr = self.e_sr()
        return f + fsr
