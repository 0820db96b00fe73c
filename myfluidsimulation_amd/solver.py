"""Python handle on one libfluidsim solver (the reference's Solver concept).

Mirrors LBM3D_Legacy / LBM3D_Home / LBM3D_Hybrid (src/lbm/*.hpp of the
reference): create (= init), set_solid_mask, set_legacy_mapping, reset,
reinit_from_macro, step, macro pointers / copies.  Host arrays are numpy;
device pointers may be passed as ints (e.g. torch ``tensor.data_ptr()``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import BGK, HOME, HYBRID, call, lib


def _host_ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class DeviceBuffer:
    """Device allocation owned by Python (through the library's helpers)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        call("fs_device_malloc", C.byref(p), C.c_size_t(nbytes))
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        call("fs_memcpy_htod", C.c_void_p(b.ptr), _host_ptr(a), C.c_size_t(a.nbytes))
        return b

    def to_host(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        call("fs_memcpy_dtoh", _host_ptr(out), C.c_void_p(self.ptr), C.c_size_t(out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().fs_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Solver:
    def __init__(self, method: int, nx: int, ny: int, nz: int, tau: float,
                 force=(0.0, 0.0, 0.0)):
        h = C.c_void_p()
        call("fs_solver_create", int(method), int(nx), int(ny), int(nz), float(tau),
             float(force[0]), float(force[1]), float(force[2]), C.byref(h))
        self.h = h
        self.method = int(method)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.n = self.nx * self.ny * self.nz

    # -- lifetime
    def close(self):
        if self.h:
            lib().fs_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- configuration
    def _u8(self, a) -> np.ndarray:
        a = np.ascontiguousarray(np.asarray(a, dtype=np.uint8).reshape(-1))
        if a.size != self.n:
            raise ValueError(f"expected {self.n} bytes, got {a.size}")
        return a

    def set_solid_mask(self, mask):
        a = self._u8(mask)
        call("fs_solver_set_solid_mask", self.h, _host_ptr(a))

    def set_legacy_mapping(self, is_legacy=None):
        if is_legacy is None:
            call("fs_solver_set_legacy_mapping", self.h, None, None)
        else:
            a = self._u8(is_legacy)
            call("fs_solver_set_legacy_mapping", self.h, _host_ptr(a), None)

    def set_flags(self, flags):
        a = self._u8(flags)
        call("fs_solver_set_flags", self.h, _host_ptr(a))

    def flags(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.uint8)
        call("fs_solver_copy_flags", self.h, _host_ptr(out))
        return out

    def set_inlet_profile(self, ux=None):
        if ux is None:
            call("fs_solver_set_inlet_profile", self.h, None)
        else:
            a = np.ascontiguousarray(np.asarray(ux, dtype=np.float32).reshape(-1))
            if a.size != self.ny * self.nz:
                raise ValueError("inlet profile needs ny*nz values")
            call("fs_solver_set_inlet_profile", self.h, _host_ptr(a))

    def set_force(self, fx, fy, fz):
        call("fs_solver_set_force", self.h, float(fx), float(fy), float(fz))

    def set_omegas(self, omega_legacy, omega_hydro, omega_shear):
        call("fs_solver_set_omegas", self.h, float(omega_legacy), float(omega_hydro),
             float(omega_shear))

    def set_bounce_rule(self, rule: int):
        call("fs_solver_set_bounce_rule", self.h, int(rule))

    # -- state
    def reset(self):
        call("fs_solver_reset", self.h)

    def reinit_from_macro(self, rho=None, ux=None, uy=None, uz=None):
        """Host numpy arrays (copied to the device) or device pointers (int)."""
        keep = []

        def dev(a):
            if a is None:
                return None
            if isinstance(a, int):
                return C.c_void_p(a)
            b = DeviceBuffer.from_host(np.asarray(a, dtype=np.float32).reshape(-1))
            keep.append(b)
            return C.c_void_p(b.ptr)

        call("fs_solver_reinit_from_macro", self.h, dev(rho), dev(ux), dev(uy), dev(uz))
        for b in keep:
            b.free()

    def step(self, nsteps: int = 1):
        call("fs_solver_step", self.h, int(nsteps))

    def step_async(self, nsteps: int):
        call("fs_solver_step_async", self.h, int(nsteps))

    def timed_steps(self, nsteps: int) -> float:
        """Run nsteps between HIP events on the solver's stream; device ms."""
        ms = C.c_float()
        call("fs_solver_timed_steps", self.h, int(nsteps), C.byref(ms))
        return float(ms.value)

    def set_variant(self, variant: int):
        call("fs_solver_set_variant", self.h, int(variant))

    @property
    def variant(self) -> int:
        v = C.c_int()
        call("fs_solver_get_variant", self.h, C.byref(v))
        return v.value

    def autotune(self, trial_steps: int = 10) -> int:
        """Time every kernel variant from the current state; keep the fastest."""
        v = C.c_int()
        call("fs_solver_autotune", self.h, int(trial_steps), C.byref(v))
        return v.value

    def synchronize(self):
        call("fs_solver_synchronize", self.h)

    @property
    def stream(self) -> int:
        p = C.c_void_p()
        call("fs_solver_get_stream", self.h, C.byref(p))
        return p.value or 0

    def set_stream(self, stream_handle):
        call("fs_solver_set_stream", self.h, C.c_void_p(stream_handle) if stream_handle else None)

    def macro_ptr(self, which: int) -> int:
        p = C.c_void_p()
        call("fs_solver_macro_ptr", self.h, int(which), C.byref(p))
        return p.value

    def macro(self, which: int) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float32)
        call("fs_solver_copy_macro", self.h, int(which), _host_ptr(out))
        return out

    def macros(self):
        return [self.macro(k) for k in range(4)]

    def state_size(self) -> int:
        n = C.c_size_t()
        call("fs_solver_state_size", self.h, C.byref(n))
        return n.value

    def state(self) -> np.ndarray:
        out = np.empty(self.state_size(), dtype=np.float32)
        call("fs_solver_copy_state", self.h, _host_ptr(out))
        return out

    def load_state(self, a):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))
        if a.size != self.state_size():
            raise ValueError("state size mismatch")
        call("fs_solver_load_state", self.h, _host_ptr(a))


def load_stl(path: str):
    """load_stl (stl_loader.cpp:102-122): (tri_xyz (T,9) float32, bbmin, bbmax)."""
    n = C.c_int()
    bmin = np.zeros(3, np.float32)
    bmax = np.zeros(3, np.float32)
    call("fs_load_stl", path.encode(), None, 0, C.byref(n), _host_ptr(bmin), _host_ptr(bmax))
    tris = np.zeros((max(n.value, 1), 9), np.float32)
    call("fs_load_stl", path.encode(), _host_ptr(tris), n.value, C.byref(n), _host_ptr(bmin),
         _host_ptr(bmax))
    return tris[: n.value], bmin, bmax


def voxelize(tris, bbmin, bbmax, nx, ny, nz, uniform_scale=0.5, translate=(0.5, 0.5, 0.5),
             refine=1, threshold=0.5) -> np.ndarray:
    """voxelize_mesh_to_mask on the GPU -> uint8 mask of nx*ny*nz (x fastest)."""
    t = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 9))
    bmin = np.ascontiguousarray(np.asarray(bbmin, np.float32))
    bmax = np.ascontiguousarray(np.asarray(bbmax, np.float32))
    out = np.empty(nx * ny * nz, np.uint8)
    call("fs_voxelize", _host_ptr(t), int(t.shape[0]), _host_ptr(bmin), _host_ptr(bmax),
         int(nx), int(ny), int(nz), float(uniform_scale), float(translate[0]),
         float(translate[1]), float(translate[2]), int(refine), float(threshold),
         _host_ptr(out))
    return out


def band(solid, nx, ny, nz, d0) -> np.ndarray:
    s = np.ascontiguousarray(np.asarray(solid, np.uint8).reshape(-1))
    out = np.empty(nx * ny * nz, np.uint8)
    call("fs_band", _host_ptr(s), int(nx), int(ny), int(nz), int(d0), _host_ptr(out))
    return out


__all__ = ["Solver", "Batch", "DeviceBuffer", "load_stl", "voxelize", "band", "BGK", "HOME",
           "HYBRID", "capi"]


class Batch:
    """Many independent domains stepped together (fs_batch_*).

    domains: list of (method, nx, ny, nz, tau[, (fx, fy, fz)]) tuples."""

    def __init__(self, domains):
        descs = (capi.DomainDesc * len(domains))()
        self.dims = []
        self.methods = []
        for i, d in enumerate(domains):
            m, nx, ny, nz, tau = d[:5]
            f = d[5] if len(d) > 5 else (0.0, 0.0, 0.0)
            descs[i] = capi.DomainDesc(int(m), int(nx), int(ny), int(nz), float(tau), *map(float, f))
            self.dims.append((int(nx), int(ny), int(nz)))
            self.methods.append(int(m))
        h = C.c_void_p()
        call("fs_batch_create", C.cast(descs, C.c_void_p), len(domains), C.byref(h))
        self.h = h

    def close(self):
        if self.h:
            lib().fs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.dims)

    def cells(self) -> int:
        n = C.c_size_t()
        call("fs_batch_info", self.h, None, C.byref(n))
        return n.value

    def _n(self, d):
        nx, ny, nz = self.dims[d]
        return nx * ny * nz

    def set_flags(self, d, flags):
        a = np.ascontiguousarray(np.asarray(flags, np.uint8).reshape(-1))
        assert a.size == self._n(d)
        call("fs_batch_set_flags", self.h, int(d), _host_ptr(a))

    def set_inlet_profile(self, d, ux=None):
        if ux is None:
            call("fs_batch_set_inlet_profile", self.h, int(d), None)
        else:
            a = np.ascontiguousarray(np.asarray(ux, np.float32).reshape(-1))
            call("fs_batch_set_inlet_profile", self.h, int(d), _host_ptr(a))

    def set_bounce_rule(self, d, rule):
        call("fs_batch_set_bounce_rule", self.h, int(d), int(rule))

    def reinit_from_macro(self, d, rho=None, ux=None, uy=None, uz=None):
        keep = []

        def dev(a):
            if a is None:
                return None
            b = DeviceBuffer.from_host(np.asarray(a, np.float32).reshape(-1))
            keep.append(b)
            return C.c_void_p(b.ptr)

        call("fs_batch_reinit_from_macro", self.h, int(d), dev(rho), dev(ux), dev(uy), dev(uz))
        for b in keep:
            b.free()

    def step(self, nsteps=1):
        call("fs_batch_step", self.h, int(nsteps))

    def timed_steps(self, nsteps) -> float:
        ms = C.c_float()
        call("fs_batch_timed_steps", self.h, int(nsteps), C.byref(ms))
        return float(ms.value)

    def set_launch(self, pop_blocks=0, home_blocks=0, use_graph=True):
        call("fs_batch_set_launch", self.h, int(pop_blocks), int(home_blocks), int(bool(use_graph)))

    def launch(self):
        p, h, g = C.c_int(), C.c_int(), C.c_int()
        call("fs_batch_get_launch", self.h, C.byref(p), C.byref(h), C.byref(g))
        return p.value, h.value, bool(g.value)

    def autotune(self, trial_steps=10):
        p, h = C.c_int(), C.c_int()
        call("fs_batch_autotune", self.h, int(trial_steps), C.byref(p), C.byref(h))
        return p.value, h.value

    def macro(self, d, which) -> np.ndarray:
        out = np.empty(self._n(d), np.float32)
        call("fs_batch_copy_macro", self.h, int(d), int(which), _host_ptr(out))
        return out

    def macros(self, d):
        return [self.macro(d, k) for k in range(4)]

    def state(self, d) -> np.ndarray:
        out = np.empty(self._n(d) * (10 if self.methods[d] == HOME else 19), np.float32)
        call("fs_batch_copy_state", self.h, int(d), _host_ptr(out))
        return out


class Adjoint:
    """Owner of one fs_adjoint handle (checkpoint memory of a T-step run of
    `solver`); forward_ptrs / backward_ptrs take the device pointers of
    fs_adjoint_forward / fs_adjoint_backward.  The torch front end is
    myfluidsimulation_amd.adjoint.DifferentiableRun."""

    def __init__(self, solver, steps: int, segment: int = 0, body=None):
        h = C.c_void_p()
        b = None
        if body is not None:
            b = np.ascontiguousarray(np.asarray(body, np.uint8).reshape(-1))
            if b.size != solver.n:
                raise ValueError(f"body: expected {solver.n} bytes")
        self._body = b
        call("fs_adjoint_create", solver.h, int(steps), int(segment),
             b.ctypes.data_as(C.c_void_p) if b is not None else None, C.byref(h))
        self.h = h
        self.solver = solver
        self.steps = int(steps)

    def info(self):
        T, S, K, nb = C.c_int(), C.c_int(), C.c_int(), C.c_size_t()
        call("fs_adjoint_info", self.h, C.byref(T), C.byref(S), C.byref(K), C.byref(nb))
        return dict(steps=T.value, segment=S.value, checkpoints=K.value, device_bytes=nb.value)

    def forward_ptrs(self, *ptrs):
        call("fs_adjoint_forward", self.h, *ptrs)

    def backward_ptrs(self, *ptrs):
        call("fs_adjoint_backward", self.h, *ptrs)

    # -- host (numpy) front end: staging buffers on the device, no torch
    def _stage(self, arrays):
        bufs = [None if a is None else
                DeviceBuffer.from_host(np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1)))
                for a in arrays]
        return bufs, [None if b is None else C.c_void_p(b.ptr) for b in bufs]

    def forward(self, rho0=None, ux0=None, uy0=None, uz0=None, inlet=None):
        """T steps from (rho0, u0) (N host floats each, None: rho 1 / u 0)
        with the inlet profile (ny*nz floats, None keeps the solver's).
        Returns (rho, ux, uy, uz, drag) as numpy arrays (N, N, N, N, T)."""
        n, s = self.solver.n, self.solver
        ins, iptr = self._stage([rho0, ux0, uy0, uz0, inlet])
        outs = [DeviceBuffer(4 * n) for _ in range(4)] + [DeviceBuffer(4 * self.steps)]
        try:
            call("fs_adjoint_forward", self.h, *iptr, *[C.c_void_p(b.ptr) for b in outs])
            s.synchronize()
            res = [b.to_host(np.float32, n) for b in outs[:4]]
            res.append(outs[4].to_host(np.float32, self.steps))
        finally:
            for b in ins + outs:
                if b is not None:
                    b.free()
        return tuple(res)

    def backward(self, g_rho=None, g_ux=None, g_uy=None, g_uz=None, g_drag=None,
                 want=(True, True, True, True, True)):
        """Gradients of the last forward pass, given dL/d(end macros) (N host
        floats each) and dL/d(drag) (T floats), any None for zero.  Returns
        (dL/drho0, dL/dux0, dL/duy0, dL/duz0, dL/dinlet) as numpy arrays, None
        where `want` is False."""
        n, s = self.solver.n, self.solver
        ins, iptr = self._stage([g_rho, g_ux, g_uy, g_uz, g_drag])
        sizes = [n, n, n, n, s.ny * s.nz]
        outs = [DeviceBuffer(4 * m) if w else None for m, w in zip(sizes, want)]
        try:
            call("fs_adjoint_backward", self.h, *iptr,
                 *[None if b is None else C.c_void_p(b.ptr) for b in outs])
            s.synchronize()
            res = tuple(None if b is None else b.to_host(np.float32, m)
                        for b, m in zip(outs, sizes))
        finally:
            for b in ins + outs:
                if b is not None:
                    b.free()
        return res

    def close(self):
        if getattr(self, "h", None):
            lib().fs_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
