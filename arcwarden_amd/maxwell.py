"""The full-Maxwell (Yee) branch from Python: MaxwellSimulation over the
library's aw_maxwell_* engine (include/aw.h).

Reference: MaxwellSimulation (include/pic/simulation_maxwell.hpp:24-227).
The simulation owns its particle arrays (torch device tensors); load them
with ``load_species`` (Particles::initialize), ``load_mirror``
(Particles::initialize_mirror) or ``set_particles``, and turn on the delta-f
weights with ``enable_deltaf``.  ``step`` advances fields and particles by
one step; the particle order never changes.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .sim import d2h, h2d


def _torch():
    import torch
    return torch


class MaxwellSimulation:
    def __init__(self, grid: L.AwGrid, rp: L.AwRunParams):
        self.grid, self.rp = grid, rp
        h = C.c_void_p()
        L.call("aw_maxwell_create", C.byref(grid), C.byref(rp), C.byref(h))
        self._h = h
        self._mesh = L.AwYeeMesh()
        L.call("aw_maxwell_mesh", h, C.byref(self._mesh))
        self.n = 0
        self._p = {}
        self._w = None
        self.w_uniform = 0.0
        self._df = {}
        self.nstep = 0

    def close(self):
        if getattr(self, "_h", None):
            L.lib().aw_maxwell_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- particles --------------------------------------------------------------
    def _alloc(self, n: int, with_w: bool):
        torch = _torch()
        self.n = n
        self._p = {k: torch.zeros(max(n, 1), dtype=torch.float32, device="cuda")
                   for k in ("x", "y", "ux", "uy", "uz")}
        self._w = torch.zeros(max(n, 1), dtype=torch.float32, device="cuda") if with_w else None

    def particle_struct(self) -> L.AwParticles:
        p = L.AwParticles()
        if self.n:
            p.x, p.y, p.ux, p.uy, p.uz = (self._p[k].data_ptr()
                                          for k in ("x", "y", "ux", "uy", "uz"))
            p.w = self._w.data_ptr() if self._w is not None else None
        p.n = self.n
        p.w_uniform = self.w_uniform
        return p

    def load_species(self, species: list):
        """Particles::initialize(species, grid, rp): per-particle weights."""
        nc = self.grid.nx * self.grid.ny
        self._alloc(sum(s.ppc for s in species) * nc, True)
        sp = (L.AwSpecies * len(species))(*species)
        w = (C.c_double * len(species))()
        h = C.c_void_p()
        L.call("aw_ctx_create", C.byref(self.grid), 0, 0, C.byref(h))
        try:
            L.call("aw_load_species", h, sp, len(species), C.byref(self.rp), None, 0,
                   C.byref(self.particle_struct()), w, None)
            _torch().cuda.synchronize()
        finally:
            L.lib().aw_ctx_destroy(h)

    def load_mirror(self, species: L.AwSpecies):
        """Particles::initialize_mirror(q, grid, rp) (ny = 1, rp.b0_prof)."""
        n = C.c_int64()
        L.call("aw_mirror_count", C.byref(self.grid), C.byref(self.rp), C.byref(species),
               C.byref(n))
        self._alloc(n.value, True)
        L.call("aw_load_mirror", C.byref(self.grid), C.byref(self.rp), C.byref(species),
               C.byref(self.particle_struct()), None)

    def set_particles(self, arrays: dict, w=None, w_uniform: float = 0.0):
        n = int(np.asarray(arrays["x"]).size)
        self._alloc(n, w is not None)
        self.w_uniform = float(w_uniform)
        torch = _torch()
        for k in ("x", "y", "ux", "uy", "uz"):
            self._p[k][:n] = torch.from_numpy(np.ascontiguousarray(arrays[k], np.float32)).cuda()
        if w is not None:
            self._w[:n] = torch.from_numpy(np.ascontiguousarray(w, np.float32)).cuda()
        torch.cuda.synchronize()

    def enable_deltaf(self, wprec: int = 0, wd0=None):
        """Particles::enable_deltaf: wd = 0 (or wd0), plus the accumulators of
        df_wprec 1 (Kahan) / 2 (FP64)."""
        torch = _torch()
        n = max(self.n, 1)
        self._df = {"wd": torch.zeros(n, dtype=torch.float32, device="cuda")}
        if wd0 is not None:
            self._df["wd"][:self.n] = torch.from_numpy(np.ascontiguousarray(wd0, np.float32)).cuda()
        if wprec == 1:
            self._df["wc"] = torch.zeros(n, dtype=torch.float32, device="cuda")
        if wprec == 2:
            self._df["wdd"] = torch.zeros(n, dtype=torch.float64, device="cuda")
            if wd0 is not None:
                self._df["wdd"][:self.n] = self._df["wd"][:self.n].double()
        torch.cuda.synchronize()

    def _df_struct(self):
        if not self._df:
            return None
        d = L.AwDeltafState()
        d.wd = self._df["wd"].data_ptr()
        d.wc = self._df["wc"].data_ptr() if "wc" in self._df else None
        d.wdd = self._df["wdd"].data_ptr() if "wdd" in self._df else None
        return d

    # ---- stepping and state -------------------------------------------------------
    def step(self, tnow: float | None = None):
        p = self.particle_struct()
        df = self._df_struct()
        t = self.nstep * self.rp.dt if tnow is None else tnow
        L.call("aw_maxwell_step_at", self._h, C.byref(p), C.byref(df) if df else None,
               C.c_double(t))
        self.nstep += 1

    def synchronize(self):
        L.call("aw_maxwell_synchronize", self._h)

    def particles(self) -> dict:
        self.synchronize()
        out = {k: v[:self.n].cpu().numpy() for k, v in self._p.items()}
        if self._w is not None:
            out["w"] = self._w[:self.n].cpu().numpy()
        for k, v in self._df.items():
            out[k] = v[:self.n].cpu().numpy()
        return out

    def mesh(self) -> dict:
        """ex, ey, ez, bx, by, bz, jx, jy, jz as [ny, nx] host arrays."""
        self.synchronize()
        nc = self.grid.nx * self.grid.ny
        return {k: d2h(getattr(self._mesh, k), nc).reshape(self.grid.ny, self.grid.nx)
                for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz")}

    def mesh_flat(self) -> np.ndarray:
        m = self.mesh()
        return np.concatenate([m[k].ravel() for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx",
                                                      "jy", "jz")])

    def set_mesh(self, **grids):
        self.synchronize()
        for k, v in grids.items():
            h2d(getattr(self._mesh, k), np.ascontiguousarray(v, np.float32).ravel())

    def field_energy(self):
        we, wb = C.c_double(), C.c_double()
        L.call("aw_maxwell_field_energy", self._h, C.byref(we), C.byref(wb))
        return we.value, wb.value

    def residuals(self):
        b, g = C.c_double(), C.c_double()
        wd = self._df["wd"].data_ptr() if self._df else None
        L.call("aw_maxwell_residuals", self._h, C.byref(self.particle_struct()), wd, C.byref(b),
               C.byref(g))
        return b.value, g.value

    def wd_stats(self):
        if not self._df:
            return 0.0, 0.0, 0.0
        s, r, m = C.c_double(), C.c_double(), C.c_double()
        self.synchronize()
        L.call("aw_wd_stats", C.c_void_p(self._df["wd"].data_ptr()), self.n, C.byref(s),
               C.byref(r), C.byref(m), None)
        return s.value, r.value, m.value

    def diag_fields(self, t: float) -> dict:
        f = L.AwFields()
        L.call("aw_maxwell_diag_fields", self._h, C.c_double(t), C.byref(f))
        self.synchronize()
        nc = self.grid.nx * self.grid.ny
        return {k: d2h(getattr(f, k), nc).reshape(self.grid.ny, self.grid.nx)
                for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "ELx")}
