"""Python host harness over the C ABI: device-array helpers and a thin
``Simulation`` wrapper with the reference's step contract
(include/pic/simulation.hpp:38-140: init / step / run / particles / fields /
diagnostics().history()).  All computation happens in libarcwarden_hip.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

_HIP = None
HIP_H2D, HIP_D2H, HIP_D2D = 1, 2, 3


def hip() -> C.CDLL:
    """The process's HIP runtime (already loaded by the library)."""
    global _HIP
    if _HIP is None:
        L.lib()
        _HIP = C.CDLL("libamdhip64.so.7")
        _HIP.hipMemcpy.restype = C.c_int
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipDeviceSynchronize.restype = C.c_int
    return _HIP


def memcpy(dst: int, src: int, nbytes: int, kind: int) -> None:
    if nbytes == 0:
        return
    rc = hip().hipMemcpy(C.c_void_p(dst), C.c_void_p(src), nbytes, kind)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed with {rc}")


def d2h(ptr: int, n: int, dtype=np.float32) -> np.ndarray:
    out = np.empty(n, dtype=dtype)
    memcpy(out.ctypes.data, ptr, out.nbytes, HIP_D2H)
    return out


def h2d(ptr: int, arr: np.ndarray) -> None:
    a = np.ascontiguousarray(arr)
    memcpy(ptr, a.ctypes.data, a.nbytes, HIP_H2D)


def particles_struct(x, y, ux, uy, uz, w=None, w_uniform=0.0) -> L.AwParticles:
    """AwParticles over torch (or any object with data_ptr()) device arrays."""
    p = L.AwParticles()
    p.x, p.y, p.ux, p.uy, p.uz = (t.data_ptr() for t in (x, y, ux, uy, uz))
    p.w = w.data_ptr() if w is not None else None
    p.n = int(x.numel())
    p.w_uniform = float(w_uniform)
    return p


@dataclass
class SimConfig:
    """SimConfig (include/pic/config.hpp:81-98) plus execution knobs."""
    has_b0: bool = True
    deposit_policy: int = L.AW_DEPOSIT_AUTO
    field_model: int = L.AW_MODEL_ES
    tile_x: int = 0
    tile_y: int = 0
    sort_every: int = 20
    stray_frac: float = 0.0
    mode_amp: list = field(default_factory=list)
    legacy_load: bool = False
    fixed_scale: int = 0
    blocks_per_tile: int = 0
    autotune: bool = False
    current_deposit: int = L.AW_CURRENT_CIC

    def to_c(self) -> L.AwSimConfig:
        c = L.AwSimConfig()
        c.has_b0 = int(self.has_b0)
        c.deposit_policy = self.deposit_policy
        c.field_model = self.field_model
        c.tile_x, c.tile_y = self.tile_x, self.tile_y
        c.sort_every = self.sort_every
        c.stray_frac = self.stray_frac
        c.n_modes = len(self.mode_amp)
        for i, a in enumerate(self.mode_amp):
            c.mode_amp[i] = a
        c.legacy_load = int(self.legacy_load)
        c.fixed_scale = self.fixed_scale
        c.blocks_per_tile = self.blocks_per_tile
        c.autotune = int(self.autotune)
        c.current_deposit = int(self.current_deposit)
        return c


class Simulation:
    """Simulation<Cfg>(grid, rp[, species]) over aw_sim_*."""

    def __init__(self, grid: L.AwGrid, rp: L.AwRunParams, species=None,
                 config: SimConfig | None = None):
        self.grid = grid
        self.rp = rp
        self.config = config or SimConfig(legacy_load=species is None)
        sp = list(species or [])
        arr = (L.AwSpecies * max(1, len(sp)))(*sp) if sp else None
        h = C.c_void_p()
        L.call("aw_sim_create", C.byref(grid), C.byref(rp), arr, len(sp),
               C.byref(self.config.to_c()), C.byref(h))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            L.lib().aw_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self):
        L.call("aw_sim_init", self._h)

    def step(self, n: int):
        L.call("aw_sim_step", self._h, n)

    def run(self, nsteps: int | None = None):
        L.call("aw_sim_run", self._h, int(self.rp.nsteps if nsteps is None else nsteps))

    def synchronize(self):
        L.call("aw_sim_synchronize", self._h)

    def request_sort(self):
        L.call("aw_sim_request_sort", self._h)

    @property
    def stream(self) -> int:
        s = C.c_void_p()
        L.call("aw_sim_stream", self._h, C.byref(s))
        return s.value or 0

    def particle_views(self) -> L.AwParticles:
        p = L.AwParticles()
        L.call("aw_sim_particles", self._h, C.byref(p))
        return p

    def particles(self) -> dict:
        """Host copies of x, y, ux, uy, uz (and w) -- synchronises."""
        self.synchronize()
        p = self.particle_views()
        out = {k: d2h(getattr(p, k), p.n) for k in ("x", "y", "ux", "uy", "uz")}
        if p.w:
            out["w"] = d2h(p.w, p.n)
        out["n"] = p.n
        out["w_uniform"] = p.w_uniform
        return out

    def save_checkpoint(self, path: str, step: int, deck: str = "", git_hash: str = "") -> None:
        """Restart file (ARCW v1) of the state after the steps run so far."""
        L.call("aw_sim_checkpoint_save", self._h, str(path).encode(), int(step),
               C.c_double(step * self.rp.dt), deck.encode(), git_hash.encode())

    def restore_checkpoint(self, path: str) -> None:
        """Instead of init(): load a file written by save_checkpoint for the same
        grid, species and configuration; continue with step(saved step)."""
        L.call("aw_sim_checkpoint_restore", self._h, str(path).encode())

    def set_particles(self, arrays: dict) -> None:
        """Writes x, y, ux, uy, uz (host arrays in the current particle order)
        and rebuilds the next step's sources from them."""
        self.synchronize()
        p = self.particle_views()
        for k in ("x", "y", "ux", "uy", "uz"):
            a = np.asarray(arrays[k], dtype=np.float32)
            assert a.size == p.n
            h2d(getattr(p, k), a)
        L.call("aw_sim_particles_changed", self._h)

    def fields(self) -> dict:
        self.synchronize()
        f = L.AwFields()
        L.call("aw_sim_fields", self._h, C.byref(f))
        nc = self.grid.nx * self.grid.ny
        return {k: d2h(getattr(f, k), nc).reshape(self.grid.ny, self.grid.nx)
                for k, _ in L.AwFields._fields_ if getattr(f, k)}

    def rho(self) -> np.ndarray:
        self.synchronize()
        s = L.AwSources()
        L.call("aw_sim_sources", self._h, C.byref(s))
        return d2h(s.rho, self.grid.nx * self.grid.ny).reshape(self.grid.ny, self.grid.nx)

    def sources_all(self) -> dict:
        """Host copies of every allocated source grid (rho, J, dcu, amu)."""
        self.synchronize()
        s = L.AwSources()
        L.call("aw_sim_sources", self._h, C.byref(s))
        nc = self.grid.nx * self.grid.ny
        return {k: d2h(getattr(s, k), nc).reshape(self.grid.ny, self.grid.nx)
                for k, _ in L.AwSources._fields_ if getattr(s, k)}

    def history(self) -> list[dict]:
        cnt = C.c_int64()
        L.call("aw_sim_history", self._h, None, 0, C.byref(cnt))
        buf = (L.AwDiagSample * max(1, cnt.value))()
        L.call("aw_sim_history", self._h, buf, cnt.value, C.byref(cnt))
        return [{k: getattr(buf[i], k) for k, _ in L.AwDiagSample._fields_}
                for i in range(cnt.value)]

    def diag_now(self, step: int = 0) -> dict:
        d = L.AwDiagSample()
        L.call("aw_sim_diag_now", self._h, step, C.byref(d))
        return {k: getattr(d, k) for k, _ in L.AwDiagSample._fields_}

    def set_profiling(self, on: bool):
        L.call("aw_sim_set_profiling", self._h, int(on))

    def kernel_time(self):
        ms, n, s = C.c_double(), C.c_int64(), C.c_int64()
        L.call("aw_sim_kernel_time", self._h, C.byref(ms), C.byref(n), C.byref(s))
        return ms.value, n.value, s.value

    def exec_config(self) -> dict:
        """Tile size and workgroups per tile in use (after autotuning)."""
        tx, ty, b = C.c_int32(), C.c_int32(), C.c_int32()
        L.call("aw_sim_exec_config", self._h, C.byref(tx), C.byref(ty), C.byref(b))
        return {"tile_x": tx.value, "tile_y": ty.value, "blocks_per_tile": b.value}
