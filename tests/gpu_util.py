"""Helpers for the GPU parity tests: torch-owned device arrays handed to the C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from arcwarden_amd import _lib as L
from arcwarden_amd.sim import particles_struct


def dev(a, dtype=None):
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(a).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Ctx:
    def __init__(self, g: L.AwGrid, tx: int = 0, ty: int = 0):
        self.g = g
        h = C.c_void_p()
        L.call("aw_ctx_create", C.byref(g), tx, ty, C.byref(h))
        self.h = h
        nt, ntx, a, b = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        L.call("aw_ctx_tiles", h, C.byref(nt), C.byref(ntx), C.byref(a), C.byref(b))
        self.ntiles, self.ntx, self.tx, self.ty = nt.value, ntx.value, a.value, b.value

    def close(self):
        if self.h:
            torch.cuda.synchronize()
            L.lib().aw_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DevParticles:
    """Particle SoA in torch device memory."""

    def __init__(self, state: dict, w=None, w_uniform=0.0):
        self.t = {k: dev(state[k], np.float32) for k in ("x", "y", "ux", "uy", "uz")}
        self.w = dev(w, np.float32) if w is not None else None
        self.w_uniform = w_uniform

    @classmethod
    def empty(cls, n, with_w=False, w_uniform=0.0):
        z = np.zeros(n, np.float32)
        return cls({k: z for k in ("x", "y", "ux", "uy", "uz")}, w=z if with_w else None,
                   w_uniform=w_uniform)

    def struct(self) -> L.AwParticles:
        return particles_struct(*(self.t[k] for k in ("x", "y", "ux", "uy", "uz")), w=self.w,
                                w_uniform=self.w_uniform)

    def numpy(self) -> dict:
        torch.cuda.synchronize()
        out = {k: v.cpu().numpy() for k, v in self.t.items()}
        if self.w is not None:
            out["w"] = self.w.cpu().numpy()
        return out


def random_particles(rng, nx, ny, n, umax=0.3, edges=True):
    x = rng.uniform(0, nx, n).astype(np.float32)
    y = rng.uniform(0, ny, n).astype(np.float32)
    if edges and n >= 8:
        top_x = np.nextafter(np.float32(nx), np.float32(0))
        top_y = np.nextafter(np.float32(ny), np.float32(0))
        x[:4] = [0.0, top_x, 0.5, top_x]
        y[:4] = [0.0, top_y, top_y, 0.0]
    st = {"x": x, "y": y}
    for k in ("ux", "uy", "uz"):
        st[k] = rng.uniform(-umax, umax, n).astype(np.float32)
    return st


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)
