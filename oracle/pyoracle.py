"""ctypes binding of the CPU oracle (oracle/build/libfsoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline -- never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libfsoracle.so")

BGK, HOME, HYBRID = 0, 1, 2
F32, F64 = 0, 1

_p = C.c_void_p
_i = C.c_int
_d = C.c_double
_f = C.c_float
_LIB = None


def build() -> None:
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_version": (_i, []),
            "orc_num_threads": (_i, []),
            "orc_create": (_p, [_i, _i, _i, _i, _i, _d, _d, _d, _d]),
            "orc_destroy": (None, [_p]),
            "orc_set_flags": (_i, [_p, _p]),
            "orc_set_inlet": (_i, [_p, _p]),
            "orc_set_omegas": (_i, [_p, _d, _d, _d]),
            "orc_reset": (_i, [_p]),
            "orc_set_bounce": (_i, [_p, _i]),
            "orc_push_write_counts": (_i, [_p, _i, _i, _i, _i, _p]),
            "orc_reinit_f32": (_i, [_p, _p, _p, _p, _p]),
            "orc_step": (_i, [_p, _i]),
            "orc_get_macro_f32": (_i, [_p, _i, _p]),
            "orc_get_macro_f64": (_i, [_p, _i, _p]),
            "orc_state_size": (C.c_size_t, [_p]),
            "orc_get_state_f32": (_i, [_p, _p]),
            "orc_get_state_f64": (_i, [_p, _p]),
            "orc_set_state_f32": (_i, [_p, _p]),
            "orc_load_stl": (_i, [C.c_char_p, _p, _i, C.POINTER(_i), _p, _p]),
            "orc_voxelize": (_i, [_p, _i, _p, _p, _i, _i, _i, _f, _f, _f, _f, _i, _f, _p]),
            "orc_tri_box_overlap": (_i, [_p, _p, _p, _p, _p]),
            "orc_distance": (_i, [_p, _i, _i, _i, _i, _p]),
            "orc_band": (_i, [_p, _i, _i, _i, _i, _p]),
            "orc_rho_gauss": (_i, [_p, _i, _i, _i, _f, _f, _f, _f, _f]),
            "orc_rho_slab": (_i, [_p, _i, _i, _i, _i, _f, _f, _f]),
            "orc_speed": (_i, [_p, _p, _p, _p, C.c_size_t]),
            "orc_write_scalar_vtk": (_i, [C.c_char_p, _p, _i, _i, _i]),
            "orc_write_vector_vtk": (_i, [C.c_char_p, _p, _p, _p, _i, _i, _i]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return None if a is None else np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1))


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed ({rc})")


def num_threads() -> int:
    return lib().orc_num_threads()


class Oracle:
    """CPU restatement of one reference solver (FP32 or FP64)."""

    def __init__(self, method, nx, ny, nz, tau, force=(0.0, 0.0, 0.0), precision=F32):
        self.h = lib().orc_create(int(method), int(precision), nx, ny, nz, float(tau),
                                  float(force[0]), float(force[1]), float(force[2]))
        if not self.h:
            raise RuntimeError("orc_create failed")
        self.method = method
        self.n = nx * ny * nz
        self.nx, self.ny, self.nz = nx, ny, nz

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_flags(self, flags):
        a = np.ascontiguousarray(np.asarray(flags, np.uint8).reshape(-1))
        _chk(lib().orc_set_flags(self.h, _ptr(a)), "set_flags")

    def set_inlet(self, ux):
        a = _f32(ux)
        _chk(lib().orc_set_inlet(self.h, _ptr(a)), "set_inlet")

    def set_omegas(self, a, b, c):
        _chk(lib().orc_set_omegas(self.h, a, b, c), "set_omegas")

    def reset(self):
        _chk(lib().orc_reset(self.h), "reset")

    def set_bounce(self, rule):
        _chk(lib().orc_set_bounce(self.h, int(rule)), "set_bounce")

    def reinit(self, rho=None, ux=None, uy=None, uz=None):
        r, u, v, w = _f32(rho), _f32(ux), _f32(uy), _f32(uz)
        _chk(lib().orc_reinit_f32(self.h, _ptr(r), _ptr(u), _ptr(v), _ptr(w)), "reinit")

    def step(self, n=1):
        _chk(lib().orc_step(self.h, int(n)), "step")

    def macro(self, which, dtype=np.float32):
        out = np.empty(self.n, dtype)
        fn = lib().orc_get_macro_f32 if dtype == np.float32 else lib().orc_get_macro_f64
        _chk(fn(self.h, int(which), _ptr(out)), "macro")
        return out

    def macros(self, dtype=np.float32):
        return [self.macro(k, dtype) for k in range(4)]

    def state(self, dtype=np.float32):
        out = np.empty(lib().orc_state_size(self.h), dtype)
        fn = lib().orc_get_state_f32 if dtype == np.float32 else lib().orc_get_state_f64
        _chk(fn(self.h, _ptr(out)), "state")
        return out

    def set_state(self, a):
        a = _f32(a)
        _chk(lib().orc_set_state_f32(self.h, _ptr(a)), "set_state")


def load_stl(path):
    L = lib()
    n = C.c_int()
    bmin = np.zeros(3, np.float32)
    bmax = np.zeros(3, np.float32)
    rc = L.orc_load_stl(path.encode(), None, 0, C.byref(n), _ptr(bmin), _ptr(bmax))
    if rc != 0:
        return None
    tris = np.zeros((max(n.value, 1), 9), np.float32)
    _chk(L.orc_load_stl(path.encode(), _ptr(tris), n.value, C.byref(n), _ptr(bmin), _ptr(bmax)),
         "load_stl")
    return tris[: n.value], bmin, bmax


def voxelize(tris, bbmin, bbmax, nx, ny, nz, uniform_scale=0.5, translate=(0.5, 0.5, 0.5),
             refine=1, threshold=0.5):
    t = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 9))
    bmin = np.ascontiguousarray(np.asarray(bbmin, np.float32))
    bmax = np.ascontiguousarray(np.asarray(bbmax, np.float32))
    out = np.empty(nx * ny * nz, np.uint8)
    _chk(lib().orc_voxelize(_ptr(t), int(t.shape[0]), _ptr(bmin), _ptr(bmax), nx, ny, nz,
                            float(uniform_scale), float(translate[0]), float(translate[1]),
                            float(translate[2]), int(refine), float(threshold), _ptr(out)),
         "voxelize")
    return out


def tri_box_overlap(c, half, v0, v1, v2) -> bool:
    arrs = [np.ascontiguousarray(np.asarray(a, np.float32)) for a in (c, half, v0, v1, v2)]
    return bool(lib().orc_tri_box_overlap(*[_ptr(a) for a in arrs]))


def distance(solid, nx, ny, nz, max_dist):
    s = np.ascontiguousarray(np.asarray(solid, np.uint8).reshape(-1))
    out = np.empty(nx * ny * nz, np.int32)
    _chk(lib().orc_distance(_ptr(s), nx, ny, nz, int(max_dist), _ptr(out)), "distance")
    return out


def push_write_counts(solid, nx, ny, nz, table):
    """Writers per populations slot in one push step (19*N int32)."""
    s = np.ascontiguousarray(np.asarray(solid, np.uint8).reshape(-1))
    out = np.empty(19 * nx * ny * nz, np.int32)
    _chk(lib().orc_push_write_counts(_ptr(s), nx, ny, nz, int(table), _ptr(out)), "counts")
    return out


def band(solid, nx, ny, nz, d0):
    s = np.ascontiguousarray(np.asarray(solid, np.uint8).reshape(-1))
    out = np.empty(nx * ny * nz, np.uint8)
    _chk(lib().orc_band(_ptr(s), nx, ny, nz, int(d0), _ptr(out)), "band")
    return out


def rho_gauss(nx, ny, nz, cx, cy, cz, sigma, amp):
    out = np.empty(nx * ny * nz, np.float32)
    _chk(lib().orc_rho_gauss(_ptr(out), nx, ny, nz, cx, cy, cz, sigma, amp), "rho_gauss")
    return out


def rho_slab(nx, ny, nz, axis, amp, center, width):
    out = np.empty(nx * ny * nz, np.float32)
    _chk(lib().orc_rho_slab(_ptr(out), nx, ny, nz, axis, amp, center, width), "rho_slab")
    return out


def speed(u, v, w):
    u, v, w = _f32(u), _f32(v), _f32(w)
    out = np.empty(u.size, np.float32)
    _chk(lib().orc_speed(_ptr(u), _ptr(v), _ptr(w), _ptr(out), u.size), "speed")
    return out


def write_scalar_vtk(path, field, nx, ny, nz):
    a = _f32(field)
    _chk(lib().orc_write_scalar_vtk(path.encode(), _ptr(a), nx, ny, nz), "vtk")


def write_vector_vtk(path, a, b, c, nx, ny, nz):
    a, b, c = _f32(a), _f32(b), _f32(c)
    _chk(lib().orc_write_vector_vtk(path.encode(), _ptr(a), _ptr(b), _ptr(c), nx, ny, nz), "vtk")
