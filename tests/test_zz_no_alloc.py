"""aw_deposit_rho_noalloc (the allocation-free deposit on a static device
scratch array) against aw_deposit_rho on the same particles: bit-identical
rho (both accumulate the same fixed-point integers), on two calls in a row
whose second input is smaller, so a stale scratch from the first call would
show; the whole-grid LDS policy too; and the size limit is an error, not a
fault."""
import ctypes as C

import numpy as np
import pytest
import torch

from arcwarden_amd import _lib as L

from gpu_util import Ctx, DevParticles, random_particles, stream

pytestmark = pytest.mark.gpu


def _lib_fn():
    return L.lib().aw_deposit_rho_noalloc


def _case(rng, nx, ny, n, policy, per_w, base):
    g = L.grid(nx, ny, nx * 0.5, ny * 0.37)
    st = random_particles(rng, nx, ny, n)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32) if per_w else None
    p = DevParticles(st, w=w, w_uniform=0.0 if per_w else 0.0123)
    init = rng.normal(0, 1, (ny, nx)).astype(np.float32) if base else np.zeros((ny, nx),
                                                                                np.float32)
    rho_ctx = torch.from_numpy(init.copy()).cuda()
    rho_na = torch.from_numpy(init.copy()).cuda()
    S = 38
    with Ctx(g) as ctx:
        L.call("aw_ctx_set_fixed_scale", ctx.h, S)
        L.call("aw_deposit_rho", ctx.h, C.byref(p.struct()), C.c_double(-1.0), policy,
               C.c_void_p(rho_ctx.data_ptr()), stream())
        torch.cuda.synchronize()
    L.check(_lib_fn()(C.byref(g), C.byref(p.struct()), C.c_double(-1.0), policy, S,
                      C.c_void_p(rho_na.data_ptr()), stream()))
    torch.cuda.synchronize()
    a, b = rho_ctx.cpu().numpy(), rho_na.cpu().numpy()
    assert np.array_equal(a, b), (nx, ny, n, policy)
    assert np.abs(a - init).max() > 0  # something was deposited
    return a


def test_noalloc_deposit_matches_context_deposit(gpu):
    rng = np.random.default_rng(5)
    # two calls in a row, the second smaller (grid and particle count)
    _case(rng, 200, 150, 300_000, L.AW_DEPOSIT_GLOBAL, False, True)
    _case(rng, 40, 24, 5_000, L.AW_DEPOSIT_GLOBAL, True, False)
    # the whole-grid LDS policy (the grid fits LDS), then AUTO on a larger grid
    _case(rng, 48, 40, 60_000, L.AW_DEPOSIT_LDS_GRID, False, False)
    _case(rng, 1024, 1024, 400_000, L.AW_DEPOSIT_AUTO, True, True)
    _case(rng, 33, 17, 9_000, L.AW_DEPOSIT_AUTO, False, True)


def test_noalloc_deposit_rejects_what_it_cannot_do(gpu):
    rng = np.random.default_rng(6)
    f = _lib_fn()
    st = random_particles(rng, 16, 16, 100)
    p = DevParticles(st, w_uniform=0.01)
    rho = torch.zeros(2048 * 1024, dtype=torch.float32, device="cuda")
    big = L.grid(2048, 1024, 1024.0, 512.0)   # 2^21 cells > the 2^20-cell scratch
    assert f(C.byref(big), C.byref(p.struct()), C.c_double(-1.0), L.AW_DEPOSIT_GLOBAL, 40,
             C.c_void_p(rho.data_ptr()), stream()) == L.AW_ERR_INVALID
    small = L.grid(16, 16, 8.0, 8.0)
    assert f(C.byref(small), C.byref(p.struct()), C.c_double(-1.0), L.AW_DEPOSIT_BINNED, 40,
             C.c_void_p(rho.data_ptr()), stream()) != 0
    torch.cuda.synchronize()
    assert float(rho.abs().max()) == 0.0
