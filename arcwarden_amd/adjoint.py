"""Gradients of the field energy of an electrostatic run: the library's
discrete adjoint (aw_adj_*, include/aw.h) behind a torch.autograd.Function.

The run is a species load whose x positions are displaced by
sum_m A_m sin(m 2 pi x / Lx), stepped nsteps times by the library's ES step;
the loss is EE(nsteps).  Parameters: the mode amplitudes A (n_modes) and the
species drift velocities ufl (nspecies x 3).  The forward pass keeps a
checkpoint every ``checkpoint_every`` steps; the backward pass recomputes each
segment from its checkpoint and sweeps it in reverse on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .sim import d2h


class AdjointRun:
    def __init__(self, grid: L.AwGrid, rp: L.AwRunParams, species: list, n_modes: int,
                 checkpoint_every: int = 10, tile: int = 0):
        self.grid, self.rp, self.species, self.n_modes = grid, rp, list(species), n_modes
        arr = (L.AwSpecies * len(self.species))(*self.species)
        h = C.c_void_p()
        L.call("aw_adj_create", C.byref(grid), C.byref(rp), arr, len(self.species), n_modes,
               checkpoint_every, tile, C.byref(h))
        self._h = h
        nc = grid.nx * grid.ny
        self.n = sum(s.ppc for s in self.species) * nc

    def close(self):
        if getattr(self, "_h", None):
            L.lib().aw_adj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, mode_amp, ufl=None, nsteps: int = 100) -> float:
        a = np.ascontiguousarray(mode_amp, np.float64).reshape(-1)
        assert a.size == self.n_modes
        u = None
        if ufl is not None:
            u = np.ascontiguousarray(ufl, np.float64).reshape(-1)
            assert u.size == 3 * len(self.species)
        loss = C.c_double()
        L.call("aw_adj_forward", self._h,
               a.ctypes.data_as(C.POINTER(C.c_double)) if a.size else None,
               u.ctypes.data_as(C.POINTER(C.c_double)) if u is not None else None,
               int(nsteps), C.byref(loss))
        return loss.value

    def backward(self):
        gm = np.zeros(max(1, self.n_modes))
        gu = np.zeros(3 * len(self.species))
        L.call("aw_adj_backward", self._h, gm.ctypes.data_as(C.POINTER(C.c_double)),
               gu.ctypes.data_as(C.POINTER(C.c_double)))
        return gm[:self.n_modes], gu.reshape(len(self.species), 3)

    def _dev4(self):
        ptrs = []
        for _ in range(4):
            p = C.c_void_p()
            L.call("aw_device_alloc", self.n * 4, C.byref(p))
            ptrs.append(p)
        return ptrs

    def _host4(self, ptrs, names):
        out = {k: d2h(p.value, self.n) for k, p in zip(names, ptrs)}
        for p in ptrs:
            L.call("aw_device_free", p)
        return out

    def state_grad(self) -> dict:
        ptrs = self._dev4()
        L.call("aw_adj_state_grad", self._h, *ptrs)
        return self._host4(ptrs, ("x", "y", "ux", "uy"))

    def initial_state(self) -> dict:
        ptrs = self._dev4()
        p = L.AwParticles()
        p.x, p.y, p.ux, p.uy = (q.value for q in ptrs)
        p.n = self.n
        L.call("aw_adj_initial_state", self._h, C.byref(p))
        return self._host4(ptrs, ("x", "y", "ux", "uy"))


def _autograd():
    import torch

    class FieldEnergy(torch.autograd.Function):
        """loss = EE(nsteps) of the run with mode amplitudes A and drifts ufl."""

        @staticmethod
        def forward(ctx, A, ufl, run: AdjointRun, nsteps: int):
            loss = run.forward(A.detach().cpu().double().numpy(),
                               ufl.detach().cpu().double().numpy(), nsteps)
            ctx.run = run
            ctx.shapes = (A, ufl)
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            gm, gu = ctx.run.backward()
            A, ufl = ctx.shapes
            g = float(grad_out)
            return (torch.as_tensor(gm * g, dtype=A.dtype, device=A.device),
                    torch.as_tensor(gu * g, dtype=ufl.dtype, device=ufl.device), None, None)

    return FieldEnergy


_FE = None


def field_energy(A, ufl, run: AdjointRun, nsteps: int):
    """Differentiable EE(nsteps) (torch tensors A [n_modes], ufl [nspecies, 3])."""
    global _FE
    if _FE is None:
        _FE = _autograd()
    return _FE.apply(A, ufl, run, nsteps)


def two_stream_adjoint(nx: int = 128, ppc: int = 64, v0: float = 0.2, vth: float = 0.02,
                       n_modes: int = 16, dx: float = 0.5, dt: float = 0.1,
                       seed: int = 20260630, checkpoint_every: int = 10):
    """The differentiable two-stream workload: nx x nx grid, two beams +-v0
    (density 1/2, ppc/2 each, thermal vth), modes A_m ~ 1e-3 U[-1, 1] drawn
    from numpy's generator seeded with ``seed``.  Returns (run, A, ufl)."""
    g = L.grid(nx, nx, nx * dx, nx * dx)
    rp = L.run_params(dt=dt, qm=-1.0, eps0=1.0, noisy_load=1, rng_seed=seed)
    sp = [L.species("beam_plus", 0.5, ppc // 2, (vth,) * 3, (v0, 0.0, 0.0)),
          L.species("beam_minus", 0.5, ppc // 2, (vth,) * 3, (-v0, 0.0, 0.0))]
    A = 1e-3 * np.random.default_rng(seed).uniform(-1.0, 1.0, n_modes)
    ufl = np.array([[v0, 0.0, 0.0], [-v0, 0.0, 0.0]])
    return AdjointRun(g, rp, sp, n_modes, checkpoint_every), A, ufl
