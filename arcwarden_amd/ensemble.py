"""Batched ensembles over aw_ensemble_* (include/aw.h): electrostatic runs of
any mix of grids, or a uniform batch of Darwin runs (SimConfig with
field_model = AW_MODEL_DARWIN).  Also the randomised ES ensemble workload:
runs of a two-beam plasma whose grid side, particles per cell, thermal speed
and beam speed are drawn from a seeded generator."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .sim import SimConfig, d2h


@dataclass
class RunSpec:
    grid: L.AwGrid
    rp: L.AwRunParams
    species: list
    has_b0: bool = False
    mode_amp: tuple = ()


def two_beam_run(side: int, ppc: int, vth: float, v0: float, *, dx: float = 0.5,
                 dt: float = 0.1, seed: int = 0, dump_every: int = 0) -> RunSpec:
    """Two counter-streaming beams (+-v0 along x, thermal spread vth in every
    component, density 1/2 each, ppc/2 particles per cell each) on a periodic
    side x side grid of cell size dx, noisy load."""
    g = L.grid(side, side, side * dx, side * dx)
    rp = L.run_params(dt=dt, qm=-1.0, eps0=1.0, noisy_load=1, rng_seed=seed,
                      dump_every=dump_every)
    sp = [L.species("beam_plus", 0.5, ppc // 2, (vth,) * 3, (v0, 0.0, 0.0)),
          L.species("beam_minus", 0.5, ppc // 2, (vth,) * 3, (-v0, 0.0, 0.0))]
    return RunSpec(g, rp, sp)


def random_ensemble(seed: int = 20260629, nruns: int = 24, sides=(64, 96, 128, 192, 256),
                    ppcs=(16, 64, 256), vth=(0.005, 0.05), v0=(0.1, 0.3), dx: float = 0.5,
                    dt: float = 0.1, dump_every: int = 0) -> list[RunSpec]:
    """The randomised ensemble: side and ppc uniform over the lists, vth
    log-uniform and v0 uniform over the ranges, one numpy generator seeded
    with ``seed`` drawing (in this order) all sides, ppcs, vths, v0s."""
    rng = np.random.default_rng(seed)
    s = rng.choice(sides, nruns)
    p = rng.choice(ppcs, nruns)
    t = np.exp(rng.uniform(np.log(vth[0]), np.log(vth[1]), nruns))
    v = rng.uniform(v0[0], v0[1], nruns)
    return [two_beam_run(int(s[i]), int(p[i]), float(t[i]), float(v[i]), dx=dx, dt=dt,
                         seed=seed + i, dump_every=dump_every) for i in range(nruns)]


class Ensemble:
    """aw_ensemble over a list of RunSpec; ``graph=True`` replays the step
    from HIP graphs."""

    def __init__(self, runs: list[RunSpec], config: SimConfig | None = None,
                 graph: bool = False):
        self.runs = runs
        descs = (L.AwRunDesc * len(runs))()
        self._keep = []
        for d, r in zip(descs, runs):
            d.grid = r.grid
            d.rp = r.rp
            arr = (L.AwSpecies * len(r.species))(*r.species)
            self._keep.append(arr)
            d.species = C.cast(arr, C.c_void_p)
            d.nspecies = len(r.species)
            d.has_b0 = int(r.has_b0)
            d.n_modes = len(r.mode_amp)
            for i, a in enumerate(r.mode_amp):
                d.mode_amp[i] = a
        cfg = (config or SimConfig()).to_c()
        h = C.c_void_p()
        L.call("aw_ensemble_create", descs, len(runs), C.byref(cfg), C.byref(h))
        self._h = h
        if graph:
            L.call("aw_ensemble_set_graph", h, 1)

    def close(self):
        if getattr(self, "_h", None):
            L.lib().aw_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self):
        L.call("aw_ensemble_init", self._h)

    def step(self, n: int):
        L.call("aw_ensemble_step", self._h, n)

    def run(self, first: int, nsteps: int):
        L.call("aw_ensemble_run", self._h, first, nsteps)

    def synchronize(self):
        L.call("aw_ensemble_synchronize", self._h)

    def info(self) -> dict:
        s = L.AwEnsembleStats()
        L.call("aw_ensemble_info", self._h, C.byref(s))
        return {k: getattr(s, k) for k, _ in L.AwEnsembleStats._fields_}

    def particles(self, run: int) -> dict:
        p = L.AwParticles()
        L.call("aw_ensemble_particles", self._h, run, C.byref(p))
        self.synchronize()
        out = {k: d2h(getattr(p, k), p.n) for k in ("x", "y", "ux", "uy", "uz")}
        out["w_uniform"] = p.w_uniform
        return out

    def rho(self, run: int) -> np.ndarray:
        ptr = C.c_void_p()
        L.call("aw_ensemble_rho", self._h, run, C.byref(ptr))
        self.synchronize()
        g = self.runs[run].grid
        return d2h(ptr.value, g.nx * g.ny).reshape(g.ny, g.nx)

    def fields(self, run: int) -> dict:
        """Host copies of the run's field grids (ES: Ex, Ey; Darwin: all)."""
        f = L.AwFields()
        L.call("aw_ensemble_fields", self._h, run, C.byref(f))
        self.synchronize()
        g = self.runs[run].grid
        return {k: d2h(getattr(f, k), g.nx * g.ny).reshape(g.ny, g.nx)
                for k, _ in L.AwFields._fields_ if getattr(f, k)}

    def history(self, run: int) -> list[dict]:
        cnt = C.c_int64()
        L.call("aw_ensemble_history", self._h, run, None, 0, C.byref(cnt))
        arr = (L.AwDiagSample * max(1, cnt.value))()
        L.call("aw_ensemble_history", self._h, run, arr, cnt.value, C.byref(cnt))
        return [{k: getattr(arr[i], k) for k, _ in L.AwDiagSample._fields_}
                for i in range(cnt.value)]
