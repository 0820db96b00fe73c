"""Differentiable runs: a T-step solver run as a torch function of its
initial macro fields and inlet profile, with gradients from the library's
hand-written adjoint kernels (fs_adjoint_* in include/fluidsim.h).

    run = DifferentiableRun(solver, steps=2000)            # sqrt(T) checkpoints
    rho, ux, uy, uz, drag = run(ux0, uy0, uz0, rho0=None, inlet=None)
    loss = 0.5 * ((ux - tx) ** 2 + (uy - ty) ** 2 + (uz - tz) ** 2).sum()
    loss.backward()                                         # ux0.grad, ...

Tensors are float32 on the GPU the solver runs on, N = nx*ny*nz cells (any
shape with N elements, x fastest) and ny*nz inlet values (index z*ny + y).
Outputs: the solver's macro fields after T steps (as Solver.macro reports
them) and drag[n], the x-momentum handed to body cells in step n.  The work
is enqueued on torch's current stream.  torch is imported here only: the
rest of the package does not need it.
"""
from __future__ import annotations

import ctypes as C

import torch

from .solver import Adjoint


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _flat(t, n, name):
    if t is None:
        return None
    if t.numel() != n:
        raise ValueError(f"{name}: expected {n} elements, got {t.numel()}")
    if t.dtype != torch.float32:
        raise TypeError(f"{name}: expected float32, got {t.dtype}")
    if torch.cuda.is_available() and not t.is_cuda:
        raise ValueError(f"{name}: expected a GPU tensor")
    return t.detach().reshape(-1).contiguous()


class _RunFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, run, rho0, ux0, uy0, uz0, inlet):
        outs = run._forward(rho0, ux0, uy0, uz0, inlet)
        ctx.set_materialize_grads(False)  # unused outputs give None: no adjoint source
        ctx.run = run
        ctx.generation = run.generation
        ctx.shapes = [None if t is None else t.shape for t in (rho0, ux0, uy0, uz0, inlet)]
        return outs

    @staticmethod
    def backward(ctx, g_rho, g_ux, g_uy, g_uz, g_drag):
        run = ctx.run
        if ctx.generation != run.generation:
            raise RuntimeError("DifferentiableRun was called again after this forward pass; "
                               "backward must come before the next call")
        need = ctx.needs_input_grad[1:]
        grads = run._backward((g_rho, g_ux, g_uy, g_uz, g_drag), need)
        out = [None]
        for g, shape in zip(grads, ctx.shapes):
            out.append(None if g is None else g.reshape(shape))
        return tuple(out)


class DifferentiableRun:
    """T steps of `solver` (a configured Solver: flags, inlet, forcing) as a
    differentiable function.  segment: steps between checkpoints (0 =
    ceil(sqrt(T))); body: N host bytes marking the solid cells whose links
    count toward the drag (None: every solid cell).  The solver's flags and
    parameters must stay as they are while the run is used."""

    def __init__(self, solver, steps: int, segment: int = 0, body=None):
        self.solver = solver
        self.adj = Adjoint(solver, steps, segment, body)
        self.steps = int(steps)
        self.n = solver.n
        self.n_inlet = solver.ny * solver.nz
        self.generation = 0

    def info(self):
        return self.adj.info()

    def __call__(self, ux0, uy0, uz0, rho0=None, inlet=None):
        return _RunFn.apply(self, rho0, ux0, uy0, uz0, inlet)

    # -- internals
    def _device(self, *ts):
        for t in ts:
            if t is not None:
                return t.device
        return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")

    def _bind_stream(self, dev):
        if dev.type == "cuda":
            self.solver.set_stream(torch.cuda.current_stream(dev).cuda_stream)

    def _forward(self, rho0, ux0, uy0, uz0, inlet):
        n = self.n
        ins = [_flat(t, n, name) for t, name in
               ((rho0, "rho0"), (ux0, "ux0"), (uy0, "uy0"), (uz0, "uz0"))]
        inl = _flat(inlet, self.n_inlet, "inlet")
        dev = self._device(*ins, inl)
        self._bind_stream(dev)
        outs = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(4)]
        drag = torch.empty(self.steps, dtype=torch.float32, device=dev)
        self.generation += 1
        self.adj.forward_ptrs(*[_ptr(t) for t in ins], _ptr(inl), *[_ptr(t) for t in outs],
                              _ptr(drag))
        return (*outs, drag)

    def _backward(self, gouts, need):
        n = self.n
        gs = [_flat(g, n, "grad") for g in gouts[:4]]
        gd = _flat(gouts[4], self.steps, "grad drag")
        dev = self._device(*gs, gd)
        self._bind_stream(dev)
        grads = [torch.empty(n, dtype=torch.float32, device=dev) if need[k] else None
                 for k in range(4)]
        gin = torch.empty(self.n_inlet, dtype=torch.float32, device=dev) if need[4] else None
        self.adj.backward_ptrs(*[_ptr(g) for g in gs], _ptr(gd), *[_ptr(g) for g in grads],
                               _ptr(gin))
        return (*grads, gin)

    def close(self):
        self.adj.close()
