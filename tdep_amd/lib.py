"""ctypes binding of libtc3.so.  Loading fails loudly when the library is missing:
there is no CPU fallback for any device computation."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libtc3.so")

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "tc3_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    "tc3_destroy": (None, [ctypes.c_void_p]),
    "tc3_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "tc3_set_structure": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_int_p, _c_dbl_p]),
    "tc3_set_fc2": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_int_p, _c_int_p, _c_dbl_p, _c_dbl_p]),
    "tc3_set_fc3": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_dbl_p,
                                   _c_dbl_p, _c_dbl_p]),
    "tc3_set_isotopes": (ctypes.c_int, [ctypes.c_void_p, _c_dbl_p, ctypes.c_int]),
    "tc3_natural_isotopes": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_int_p]),
    "tc3_set_qmesh": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, ctypes.c_int]),
    "tc3_set_qmesh_sym": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, ctypes.c_int, ctypes.c_int]),
    "tc3_set_thirdorder": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "tc3_mesh_info": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p, _c_int_p,
                                     ctypes.POINTER(ctypes.c_int64), _c_int_p]),
    "tc3_get_mesh": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p, _c_int_p, _c_dbl_p, _c_dbl_p]),
    "tc3_get_symmetry": (ctypes.c_int, [ctypes.c_void_p, _c_dbl_p, _c_dbl_p, _c_int_p]),
    "tc3_harmonic": (ctypes.c_int, [ctypes.c_void_p, _c_dbl_p]),
    "tc3_get_harmonic": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_dbl_p]),
    "tc3_scatter": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_dbl_p, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_int]),
    "tc3_get_scattering": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_dbl_p]),
    "tc3_get_xi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_dbl_p, _c_int_p]),
    "tc3_kappa": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, _c_dbl_p,
                                 _c_dbl_p, _c_dbl_p, _c_int_p]),
    "tc3_kappa_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, _c_dbl_p,
                                         _c_dbl_p, _c_int_p]),
    "tc3_get_kappa_log": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_dbl_p, _c_int_p]),
    "tc3_get_modes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_dbl_p, _c_dbl_p]),
    "tc3_get_timings": (ctypes.c_int, [ctypes.c_void_p, _c_dbl_p, ctypes.c_int]),
    "tc3_get_kernel_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_dbl_p, _c_int_p,
                                            ctypes.POINTER(ctypes.c_int64)]),
    "tc3_write_grid_hdf5": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p]),
    "tc3_set_fc4": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 8),
    "tc3_set_montecarlo4": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "tc3_write_cumulative_hdf5": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                 ctypes.c_char_p, ctypes.c_char_p]),
    "tc3_set_montecarlo": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, ctypes.c_int]),
    "tc3_set_boundary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "tc3_get_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_int_p]),
    "tc3_get_xi_storage": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), _c_int_p, _c_int_p]),
    "tc3_adjoint_gamma": (ctypes.c_int, [ctypes.c_void_p, _c_dbl_p, ctypes.c_int, _c_dbl_p]),
    "tc3_adjoint_kappa_rta": (ctypes.c_int, [ctypes.c_void_p, _c_dbl_p, ctypes.c_int, _c_dbl_p]),
    "tc3_jvp": (ctypes.c_int, [ctypes.c_void_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p]),
    "tc3_get_psi2": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                    _c_dbl_p, _c_int_p]),
}


class Tc3Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tc3 status {code}: {msg}")
        self.code = code


_lib = None


def load(path=None):
    """Load libtc3.so (built in-tree by tdep_amd/build.py) and declare its ABI."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIBPATH
    if not os.path.exists(p):
        raise ImportError(f"{p} is missing: build it with `python tdep_amd/build.py` "
                          "(there is no CPU fallback)")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_c_dbl_p)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_c_int_p)
