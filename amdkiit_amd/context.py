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
        self.precision = L.PREC_FP64 if precision == "fp64" else L.PREC_MIXED
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

    def profile_reset(self):
        self._chk(self.lib.amdk_profile_reset(self.h))
