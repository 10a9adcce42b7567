"""Host-side mirror of the reference's hot-path subroutines over libamdk.

Each method keeps the name and argument meaning of the reference routine it
replaces; arrays are numpy, band-major complex128 of shape (nb, ngpw)
(= Fortran COMPLEX*16 C(NGPW, NB)).  Errors from the library raise
AmdkError (the reference STOPs).
"""
import ctypes as C

import numpy as np

from . import lib as L


class KSContext:
    def __init__(self, pw, nbmax, precision="fp64"):
        self.lib = L.load()
        self.pw = pw
        self.nbmax = int(nbmax)
        self.precision = {"fp64": L.PREC_FP64, "mixed": L.PREC_MIXED, "bf16": L.PREC_BF16}[precision]
        d = L.Desc()
        d.n1, d.n2, d.n3 = pw.n
        a = np.asfortranarray(pw.cell.a)  # columns a_j ; a_cell[3*j+i] = (a_j)_i
        for j in range(3):
            for i in range(3):
                d.a_cell[3 * j + i] = a[i, j]
        self._mpw = np.ascontiguousarray(pw.mill_pw, dtype=np.int32).reshape(-1)
        self._mrho = np.ascontiguousarray(pw.mill_rho, dtype=np.int32).reshape(-1)
        d.ngpw = pw.ngpw
        d.miller_pw = L.iptr(self._mpw)
        d.ngrho = pw.ngrho
        d.miller_rho = L.iptr(self._mrho)
        d.nbmax = self.nbmax
        d.precision = self.precision
        d.symmetry = getattr(pw.cell, "symmetry", 0)
        h = C.c_void_p()
        rc = self.lib.amdk_ctx_create(C.byref(d), C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.lib.amdk_last_error(h).decode() if h else "?"
            self.lib.amdk_ctx_destroy(h)
            self.h = None
            raise L.AmdkError(rc, msg)
        self.ldc = self.lib.amdk_ldc(h)
        self.cbytes = 16 if self.precision == L.PREC_FP64 else 8

    def close(self):
        if getattr(self, "h", None):
            self.lib.amdk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            raise L.AmdkError(rc, self.lib.amdk_last_error(self.h).decode())

    @staticmethod
    def _flat(c):
        c = np.ascontiguousarray(c, dtype=np.complex128)
        return c, c.view(np.float64).reshape(-1)

    # -- reference routines ------------------------------------------------
    def set_vks(self, v_r):
        """V_POT(r) after cal_vpot_ex, dense (n3, n2, n1) or flat x-fastest."""
        v = np.ascontiguousarray(v_r, dtype=np.float64).reshape(-1)
        self._chk(self.lib.amdk_set_vks(self.h, L.dptr(v)))

    def vpsi(self, c0, occ):
        """VPSI: returns C2 = -(f/2)(|G|^2 c + 2 (V psi)(G))."""
        c0, f0 = self._flat(c0)
        nb, ng = c0.shape
        occ = np.ascontiguousarray(occ, dtype=np.float64)
        c2 = np.zeros_like(c0)
        f2 = c2.view(np.float64).reshape(-1)
        self._chk(self.lib.amdk_vpsi(self.h, L.dptr(f0), ng, L.dptr(f2), L.dptr(occ), nb))
        return c2

    def eval_density(self, c0, occ):
        """eval_density: rho(r), shape (n3, n2, n1)."""
        c0, f0 = self._flat(c0)
        nb, ng = c0.shape
        occ = np.ascontiguousarray(occ, dtype=np.float64)
        n1, n2, n3 = self.pw.n
        rho = np.zeros((n3, n2, n1))
        self._chk(self.lib.amdk_density(self.h, L.dptr(f0), ng, L.dptr(occ), nb, L.dptr(rho.reshape(-1))))
        return rho

    def set_species(self, tables):
        """Species tables (vps, rhos, twnl, D) of psp.Tables, caller order."""
        T = tables
        self._sp_keep = dict(
            zv=np.ascontiguousarray(T.zv, np.float64),
            vps=np.ascontiguousarray(T.vps, np.float64),
            rhos=np.ascontiguousarray(T.rhos, np.float64),
            nproj=np.ascontiguousarray(T.nproj, np.int32),
            proj_l=np.ascontiguousarray(T.proj_l, np.int32),
            twnl=np.ascontiguousarray(T.twnl, np.float64),
            dmat=np.ascontiguousarray(T.dmat, np.float64),
        )
        k = self._sp_keep
        d = L.Species()
        d.nsp, d.maxproj = T.nsp, T.maxproj
        d.zv, d.vps, d.rhos = L.dptr(k["zv"]), L.dptr(k["vps"]), L.dptr(k["rhos"])
        d.nproj, d.proj_l = L.iptr(k["nproj"]), L.iptr(k["proj_l"])
        d.twnl, d.dmat = L.dptr(k["twnl"]), L.dptr(k["dmat"])
        self._chk(self.lib.amdk_set_species(self.h, C.byref(d)))
        # hess's WSG weights (zero for UPF species, as in the reference)
        k["wsg"] = np.ascontiguousarray(T.wsg, np.float64)
        self._chk(self.lib.amdk_set_precon_weights(self.h, L.dptr(k["wsg"])))

    def set_positions(self, sp_of_atom, tau):
        """exp_igrxfactor + E_SR for positions tau (nat, 3) bohr."""
        sp = np.ascontiguousarray(sp_of_atom, np.int32)
        t = np.ascontiguousarray(tau, np.float64).reshape(-1)
        self._chk(self.lib.amdk_set_positions(self.h, len(sp), L.iptr(sp), L.dptr(t)))
        self.nat = len(sp)
        self.nproj = self.lib.amdk_nproj(self.h)

    def fnl(self, c0):
        c0, f0 = self._flat(c0)
        nb, ng = c0.shape
        nl = np.zeros((nb, max(self.nproj, 1)))
        self._chk(self.lib.amdk_fnl(self.h, L.dptr(f0), ng, nb, L.dptr(nl.reshape(-1))))
        return nl[:, : self.nproj].T.copy()

    def fnonloc(self, nl, c2, occ):
        c2, f2 = self._flat(c2.copy())
        nb, ng = c2.shape
        nlf = np.ascontiguousarray(np.asarray(nl, np.float64).T).reshape(-1)
        occ = np.ascontiguousarray(occ, np.float64)
        self._chk(self.lib.amdk_fnonloc(self.h, L.dptr(nlf), L.dptr(f2), ng, L.dptr(occ), nb))
        return c2

    def overlap(self, ca, cb):
        ca, fa = self._flat(ca)
        cb, fb = self._flat(cb)
        A = np.zeros((cb.shape[0], ca.shape[0]))
        self._chk(self.lib.amdk_overlap(self.h, L.dptr(fa), L.dptr(fb), ca.shape[1], ca.shape[0], cb.shape[0],
                                        L.dptr(A.reshape(-1))))
        return A.T.copy()

    def overlap_noalloc(self, ca, cb):
        """overlap through the allocation-free entry point."""
        ca, fa = self._flat(ca)
        cb, fb = self._flat(cb)
        A = np.zeros((cb.shape[0], ca.shape[0]))
        self._chk(self.lib.amdk_overlap_noalloc(self.h, L.dptr(fa), L.dptr(fb), ca.shape[1], ca.shape[0], cb.shape[0],
                                                L.dptr(A.reshape(-1))))
        return A.T.copy()

    def lagrange(self, c0, c2, occ):
        c0, f0 = self._flat(c0)
        c2, f2 = self._flat(c2.copy())
        occ = np.ascontiguousarray(occ, np.float64)
        gm, cn = C.c_double(), C.c_double()
        self._chk(self.lib.amdk_lagrange(self.h, L.dptr(f0), L.dptr(f2), c0.shape[1], L.dptr(occ), c0.shape[0],
                                         C.byref(gm), C.byref(cn)))
        return c2, gm.value, cn.value

    def ortho(self, c):
        c, f = self._flat(c.copy())
        self._chk(self.lib.amdk_ortho(self.h, L.dptr(f), c.shape[1], c.shape[0]))
        return c

    def vks(self, rho, xc):
        r = np.ascontiguousarray(rho, np.float64).reshape(-1)
        v = np.zeros_like(r)
        e = L.Energies()
        self._chk(self.lib.amdk_vks(self.h, L.dptr(r), L.XC[xc], L.dptr(v), C.byref(e)))
        n1, n2, n3 = self.pw.n
        return v.reshape(n3, n2, n1), e.as_dict()

    def forces(self, c0, occ):
        c0, f0 = self._flat(c0)
        occ = np.ascontiguousarray(occ, np.float64)
        f = np.zeros(3 * self.nat)
        self._chk(self.lib.amdk_forces(self.h, L.dptr(f0), c0.shape[1], L.dptr(occ), c0.shape[0], L.dptr(f)))
        return f.reshape(-1, 3)

    def wfn_load(self, c0, occ):
        c0, f0 = self._flat(c0)
        occ = np.ascontiguousarray(occ, np.float64)
        self._chk(self.lib.amdk_wfn_load(self.h, L.dptr(f0), c0.shape[1], L.dptr(occ), c0.shape[0]))
        self.nb_state = c0.shape[0]

    def wfn_store(self):
        out = np.zeros((self.nb_state, self.pw.ngpw), np.complex128)
        self._chk(self.lib.amdk_wfn_store(self.h, L.dptr(out.view(np.float64).reshape(-1)), self.pw.ngpw))
        return out

    def scf_step(self, istep, xc):
        e = L.Energies()
        self._chk(self.lib.amdk_scf_step(self.h, istep, L.XC[xc], C.byref(e)))
        return e.as_dict()

    def scf_step_tfor(self, istep, xc, want_forces):
        """One SCF step; with want_forces also the forces of the pre-update state."""
        e = L.Energies()
        f = np.zeros(3 * self.nat) if want_forces else None
        fp = L.dptr(f) if want_forces else None
        self._chk(self.lib.amdk_scf_step_tfor(self.h, istep, L.XC[xc], C.byref(e), fp))
        return e.as_dict(), (f.reshape(-1, 3) if want_forces else None)

    def set_ortho_method(self, method):
        """'cholesky' (ortho_gs, default) or 'lowdin' for the SCF step."""
        self._chk(self.lib.amdk_set_ortho_method(self.h, {"cholesky": 0, "lowdin": 1}[method]))

    def ortho_resident(self):
        self._chk(self.lib.amdk_wfn_ortho(self.h))

    def scf_forces(self):
        f = np.zeros(3 * self.nat)
        self._chk(self.lib.amdk_scf_forces(self.h, L.dptr(f)))
        return f.reshape(-1, 3)

    # -- device-resident helpers ---------------------------------------------
    def bands_to_dev(self, c, dst_ptr):
        c, f = self._flat(c)
        self._chk(self.lib.amdk_bands_to_dev(self.h, L.dptr(f), c.shape[1], c.shape[0], C.c_void_p(dst_ptr)))

    def bands_from_dev(self, src_ptr, nb):
        out = np.zeros((nb, self.pw.ngpw), np.complex128)
        f = out.view(np.float64).reshape(-1)
        self._chk(self.lib.amdk_bands_from_dev(self.h, C.c_void_p(src_ptr), nb, L.dptr(f), self.pw.ngpw))
        return out

    def vpsi_dev(self, c0_ptr, c2_ptr, occ, nb):
        occ = np.ascontiguousarray(occ, dtype=np.float64)
        self._chk(self.lib.amdk_vpsi_dev(self.h, C.c_void_p(c0_ptr), C.c_void_p(c2_ptr), L.dptr(occ), nb))

    def profile(self, on=True):
        self._chk(self.lib.amdk_profile_enable(self.h, 1 if on else 0))

    def profile_query(self, name):
        t = C.c_double()
        n = C.c_int()
        self._chk(self.lib.amdk_profile_query(self.h, name.encode(), C.byref(t), C.byref(n)))
        return t.value, n.value

    def sphere_geometry(self, which=0):
        out = (C.c_int * 6)()
        self._chk(self.lib.amdk_sphere_info(self.h, which, out))
        return dict(zip(("npw", "ldc", "nrow", "nslot", "nplane", "R"), list(out)))

    def profile_names(self):
        buf = C.create_string_buffer(1 << 16)
        self._chk(self.lib.amdk_profile_names(self.h, buf, len(buf)))
        return [n for n in buf.value.decode().split("\n") if n]

    def profile_reset(self):
        self._chk(self.lib.amdk_profile_reset(self.h))
