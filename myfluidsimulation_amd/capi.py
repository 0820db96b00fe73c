"""ctypes binding of libfluidsim.so (the C ABI in include/fluidsim.h).

The product path: every call goes to the HIP library.  There is no CPU
fallback -- if the library is missing or was built for another ABI version,
``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FLUIDSIM_LIB") or os.path.join(HERE, "lib", "libfluidsim.so")
ABI_VERSION = 1

FS_OK = 0
FS_ERR_INVALID = -1
FS_ERR_DEVICE = -2
FS_ERR_ALLOC = -3
FS_ERR_IO = -4
FS_ERR_STATE = -5

BGK, HOME, HYBRID = 0, 1, 2
METHOD_NAMES = {BGK: "bgk", HOME: "home", HYBRID: "hybrid"}

FLAG_SOLID = 0x01
FLAG_BAND = 0x02
FLAG_INLET = 0x04
FLAG_OUTLET = 0x08

RHO, UX, UY, UZ = 0, 1, 2, 3

BOUNCE_HALFWAY = 0
BOUNCE_REFERENCE_TABLE = 1

_p = C.c_void_p
_f = C.c_float
_i = C.c_int
_sz = C.c_size_t
_fp = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)

# name -> (restype, argtypes); every function declared in include/fluidsim.h
SIGNATURES = {
    "fs_version": (_i, []),
    "fs_last_error": (C.c_char_p, []),
    "fs_device_count": (_i, [C.POINTER(_i)]),
    "fs_hip_versions": (_i, [C.POINTER(_i), C.POINTER(_i)]),
    "fs_solver_create": (_i, [_i, _i, _i, _i, _f, _f, _f, _f, C.POINTER(_p)]),
    "fs_solver_destroy": (_i, [_p]),
    "fs_solver_dims": (_i, [_p, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_sz)]),
    "fs_solver_method": (_i, [_p, C.POINTER(_i)]),
    "fs_solver_set_solid_mask": (_i, [_p, _p]),
    "fs_solver_set_legacy_mapping": (_i, [_p, _p, _p]),
    "fs_solver_set_flags": (_i, [_p, _p]),
    "fs_solver_copy_flags": (_i, [_p, _p]),
    "fs_solver_set_inlet_profile": (_i, [_p, _p]),
    "fs_solver_set_force": (_i, [_p, _f, _f, _f]),
    "fs_solver_set_omegas": (_i, [_p, _f, _f, _f]),
    "fs_solver_set_bounce_rule": (_i, [_p, _i]),
    "fs_solver_reset": (_i, [_p]),
    "fs_solver_reinit_from_macro": (_i, [_p, _p, _p, _p, _p]),
    "fs_solver_step": (_i, [_p, _i]),
    "fs_solver_step_async": (_i, [_p, _i]),
    "fs_solver_set_stream": (_i, [_p, _p]),
    "fs_solver_get_stream": (_i, [_p, C.POINTER(_p)]),
    "fs_solver_synchronize": (_i, [_p]),
    "fs_solver_timed_steps": (_i, [_p, _i, C.POINTER(_f)]),
    "fs_solver_set_variant": (_i, [_p, _i]),
    "fs_solver_get_variant": (_i, [_p, C.POINTER(_i)]),
    "fs_solver_autotune": (_i, [_p, _i, C.POINTER(_i)]),
    "fs_solver_macro_ptr": (_i, [_p, _i, C.POINTER(_p)]),
    "fs_solver_copy_macro": (_i, [_p, _i, _p]),
    "fs_solver_state_size": (_i, [_p, C.POINTER(_sz)]),
    "fs_solver_copy_state": (_i, [_p, _p]),
    "fs_solver_load_state": (_i, [_p, _p]),
    "fs_batch_create": (_i, [_p, _i, C.POINTER(_p)]),
    "fs_batch_destroy": (_i, [_p]),
    "fs_batch_info": (_i, [_p, C.POINTER(_i), C.POINTER(_sz)]),
    "fs_batch_set_flags": (_i, [_p, _i, _p]),
    "fs_batch_set_inlet_profile": (_i, [_p, _i, _p]),
    "fs_batch_set_bounce_rule": (_i, [_p, _i, _i]),
    "fs_batch_reinit_from_macro": (_i, [_p, _i, _p, _p, _p, _p]),
    "fs_batch_step": (_i, [_p, _i]),
    "fs_batch_timed_steps": (_i, [_p, _i, C.POINTER(_f)]),
    "fs_batch_set_launch": (_i, [_p, _i, _i, _i]),
    "fs_batch_get_launch": (_i, [_p, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "fs_batch_autotune": (_i, [_p, _i, C.POINTER(_i), C.POINTER(_i)]),
    "fs_batch_copy_macro": (_i, [_p, _i, _i, _p]),
    "fs_batch_copy_state": (_i, [_p, _i, _p]),
    "fs_adjoint_create": (_i, [_p, _i, _i, _p, C.POINTER(_p)]),
    "fs_adjoint_destroy": (_i, [_p]),
    "fs_adjoint_info": (_i, [_p, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_sz)]),
    "fs_adjoint_forward": (_i, [_p] + [_p] * 10),
    "fs_adjoint_backward": (_i, [_p] + [_p] * 10),
    "fs_load_stl": (_i, [C.c_char_p, _p, _i, C.POINTER(_i), _p, _p]),
    "fs_voxelize": (_i, [_p, _i, _p, _p, _i, _i, _i, _f, _f, _f, _f, _i, _f, _p]),
    "fs_band": (_i, [_p, _i, _i, _i, _i, _p]),
    "fs_band_noalloc": (_i, [_p, _i, _i, _i, _i, _p, _p]),
    "fs_geometry_release": (_i, []),
    "fs_speed": (_i, [_p, _p, _p, _p, _sz]),
    "fs_diff": (_i, [_p, _p, _p, _sz]),
    "fs_shift": (_i, [_p, _f, _p, _sz]),
    "fs_rho_gauss": (_i, [_p, _i, _i, _i, _f, _f, _f, _f, _f]),
    "fs_rho_slab": (_i, [_p, _i, _i, _i, _i, _f, _f, _f]),
    "fs_reinit_equilibrium": (_i, [_p, _p, _p, _p, _p, _sz]),
    "fs_write_scalar_vtk": (_i, [C.c_char_p, _p, _i, _i, _i]),
    "fs_write_vector_vtk": (_i, [C.c_char_p, _p, _p, _p, _i, _i, _i]),
    "fs_write_scalar_vtk_binary": (_i, [C.c_char_p, _p, _i, _i, _i]),
    "fs_write_vector_vtk_binary": (_i, [C.c_char_p, _p, _p, _p, _i, _i, _i]),
    "fs_device_malloc": (_i, [C.POINTER(_p), _sz]),
    "fs_device_free": (_i, [_p]),
    "fs_memcpy_htod": (_i, [_p, _p, _sz]),
    "fs_memcpy_dtoh": (_i, [_p, _p, _sz]),
    "fs_memset_device": (_i, [_p, _i, _sz]),
    "fs_device_synchronize": (_i, []),
    "fs_copy_probe": (_i, [_p, _p, _sz, _i, C.POINTER(_f)]),
    "write_scalar_vtk": (None, [C.c_char_p, _p, _i, _i, _i]),
    "write_vector_vtk": (None, [C.c_char_p, _p, _p, _p, _i, _i, _i]),
    "make_rho_gauss": (None, [_p, _i, _i, _i, _f, _f, _f, _f, _f]),
    "make_rho_slab": (None, [_p, _i, _i, _i, _i, _f, _f, _f]),
    "reinit_equilibrium_from_macro": (None, [_p, _p, _p, _p, _p, _i]),
    "launch_speed": (None, [_p, _p, _p, _p, _i]),
    "launch_diff": (None, [_p, _p, _p, _i]),
    "launch_shift": (None, [_p, _f, _p, _i]),
}



class DomainDesc(C.Structure):
    """fs_domain_desc"""
    _fields_ = [("method", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("tau", C.c_float), ("fx", C.c_float), ("fy", C.c_float), ("fz", C.c_float)]


_LIB = None


class FluidSimError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load libfluidsim.so (once).  Raises if it is missing."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "or `make -C myfluidsimulation_amd`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        v = L.fs_version()
        if v != ABI_VERSION:
            raise RuntimeError(f"libfluidsim ABI {v} != {ABI_VERSION}")
        _LIB = L
    return _LIB


def check(rc: int, where: str) -> None:
    if rc != FS_OK:
        msg = lib().fs_last_error()
        raise FluidSimError(rc, where, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


def hip_versions():
    """(HIP version the library was built against, runtime version it is
    bound to), each major*10000000 + minor*100000 + patch"""
    b, r = C.c_int(0), C.c_int(0)
    call("fs_hip_versions", C.byref(b), C.byref(r))
    return b.value, r.value


def device_count() -> int:
    n = C.c_int(0)
    call("fs_device_count", C.byref(n))
    return n.value
