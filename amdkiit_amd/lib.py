"""ctypes binding of libamdk (include/amdk.h).

This is the Python face of the same C ABI that the Fortran module binds; it
passes plain pointers and sizes.  Loading fails loudly if the library has not
been built: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libamdk.so")

OK, ERR_ARG, ERR_DEVICE, ERR_OOM, ERR_NOT_PD, ERR_NONFINITE, ERR_STATE = range(7)
PREC_FP64, PREC_MIXED, PREC_BF16 = 0, 1, 2

# Exported symbols with (restype, argtypes).  Kept in one table so the CPU
# test can check that every declaration in include/amdk.h is exported.
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i = C.c_int


class Desc(C.Structure):
    _fields_ = [
        ("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int),
        ("a_cell", C.c_double * 9),
        ("ngpw", C.c_int), ("miller_pw", _ip),
        ("ngrho", C.c_int), ("miller_rho", _ip),
        ("nbmax", C.c_int), ("precision", C.c_int), ("symmetry", C.c_int),
    ]


class Species(C.Structure):
    _fields_ = [
        ("nsp", C.c_int), ("maxproj", C.c_int),
        ("zv", _dp), ("vps", _dp), ("rhos", _dp),
        ("nproj", _ip), ("proj_l", _ip), ("twnl", _dp), ("dmat", _dp),
    ]


class Layout(C.Structure):
    _fields_ = [("grid_ld1", C.c_int), ("grid_ld2", C.c_int), ("nl_nghmax", C.c_int), ("force_na_max", C.c_int)]


FORCE_ESR, FORCE_POTFOR, FORCE_NLFOR = 1, 2, 4


class Energies(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("etot", "ekin", "estat", "eself", "esr", "eloc", "enl", "exc", "eh", "gemax", "cnorm", "nelec")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


XC = {"TETER93": 0, "PZ": 1, "PBE": 2}

SIGNATURES = {
    "amdk_ctx_create": (_i, [C.POINTER(Desc), C.POINTER(_vp)]),
    "amdk_ctx_destroy": (None, [_vp]),
    "amdk_last_error": (C.c_char_p, [_vp]),
    "amdk_ldc": (_i, [_vp]),
    "amdk_pw_order": (_i, [_vp, _ip, _ip]),
    "amdk_set_vks": (_i, [_vp, _dp]),
    "amdk_vpsi": (_i, [_vp, _dp, _i, _dp, _dp, _i]),
    "amdk_vpsi_dev": (_i, [_vp, _vp, _vp, _dp, _i]),
    "amdk_density": (_i, [_vp, _dp, _i, _dp, _i, _dp]),
    "amdk_density_dev": (_i, [_vp, _vp, _dp, _i, _vp]),
    "amdk_set_species": (_i, [_vp, C.POINTER(Species)]),
    "amdk_set_precon_weights": (_i, [_vp, _dp]),
    "amdk_set_positions": (_i, [_vp, _i, _ip, _dp]),
    "amdk_nproj": (_i, [_vp]),
    "amdk_fnl": (_i, [_vp, _dp, _i, _i, _dp]),
    "amdk_fnonloc": (_i, [_vp, _dp, _dp, _i, _dp, _i]),
    "amdk_overlap": (_i, [_vp, _dp, _dp, _i, _i, _i, _dp]),
    "amdk_overlap_noalloc": (_i, [_vp, _dp, _dp, _i, _i, _i, _dp]),
    "amdk_lagrange": (_i, [_vp, _dp, _dp, _i, _dp, _i, _dp, _dp]),
    "amdk_ortho": (_i, [_vp, _dp, _i, _i]),
    "amdk_vks": (_i, [_vp, _dp, _i, _dp, C.POINTER(Energies)]),
    "amdk_forces": (_i, [_vp, _dp, _i, _dp, _i, _dp]),
    "amdk_wfn_load": (_i, [_vp, _dp, _i, _dp, _i]),
    "amdk_wfn_store": (_i, [_vp, _dp, _i]),
    "amdk_scf_step": (_i, [_vp, _i, _i, C.POINTER(Energies)]),
    "amdk_scf_forces": (_i, [_vp, _dp]),
    "amdk_scf_step_tfor": (_i, [_vp, _i, _i, C.POINTER(Energies), _dp]),
    "amdk_wfn_ortho": (_i, [_vp]),
    "amdk_bands_to_dev": (_i, [_vp, _dp, _i, _i, _vp]),
    "amdk_bands_from_dev": (_i, [_vp, _vp, _i, _dp, _i]),
    "amdk_sphere_info": (_i, [_vp, _i, _ip]),
    "amdk_set_layout": (_i, [_vp, C.POINTER(Layout)]),
    "amdk_set_positions_atco": (_i, [_vp, _i, _ip, _i, _dp]),
    "amdk_forces_add": (_i, [_vp, _dp, _i, _dp, _i, _i, _dp]),
    "amdk_pcg": (_i, [_vp, _dp, _dp, _dp, _i, _dp, _i, _i]),
    "amdk_ekin": (_i, [_vp, _dp, _i, _dp, _i, _dp]),
    "amdk_enl": (_i, [_vp, _dp, _dp, _i, _dp]),
    "amdk_ion_energies": (_i, [_vp, _dp, _dp]),
    "amdk_set_ortho_method": (_i, [_vp, _i]),
    "amdk_profile_enable": (_i, [_vp, _i]),
    "amdk_profile_names": (_i, [_vp, C.c_char_p, _i]),
    "amdk_profile_query": (_i, [_vp, C.c_char_p, _dp, _ip]),
    "amdk_profile_reset": (_i, [_vp]),
}

_lib = None


def load(path=LIBPATH):
    """Load libamdk.so and attach signatures.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libamdk.so not built at {path}: run python -m amdkiit_amd.build")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class AmdkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"amdk error {code}: {msg}")
        self.code = code


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)
