"""GPU parity of the diagnostics reductions the modules use: the phase-space
histogram bit for bit against the oracle's binning (both the FP32 movie and
the FP64 snapshot rules, with wrapped positions, an oblique B0 and values
at the window edges), the FP64 sum of squares and the sample gather."""
import ctypes as C

import numpy as np
import pytest
import torch

from arcwarden_amd import _lib as L

from gpu_util import DevParticles, stream

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fp64", [0, 1])
def test_phase_hist_matches_oracle(gpu, oracle, fp64):
    rng = np.random.default_rng(4)
    nx, ny, n = 96, 10, 300_000
    g = L.grid(nx, ny, nx * 0.5, ny * 0.5)
    st = {"x": rng.uniform(-0.5, nx + 0.5, n).astype(np.float32),   # some outside one period
          "y": rng.uniform(0, ny, n).astype(np.float32),
          "ux": rng.normal(0, 0.06, n).astype(np.float32),
          "uy": np.zeros(n, np.float32),
          "uz": rng.normal(0, 0.06, n).astype(np.float32)}
    bx, bz, vth, nxb, nvb = np.cos(0.5), np.sin(0.5), 0.05, 256, 200
    p = DevParticles(st)
    counts = torch.zeros(nxb * nvb, dtype=torch.int32, device="cuda")
    L.call("aw_phase_hist", C.byref(p.struct()), C.byref(g), nxb, nvb, C.c_double(bx),
           C.c_double(bz), C.c_double(vth), C.c_double(-6.0), C.c_double(6.0), fp64,
           C.c_void_p(counts.data_ptr()), stream())
    torch.cuda.synchronize()
    got = counts.cpu().numpy().view(np.uint32).reshape(nvb, nxb)
    ref = oracle.phase_hist(oracle.grid(nx, ny, 0.5, 0.5), st, nxb, nvb, bx, bz, vth, -6.0,
                            6.0, fp64)
    assert np.array_equal(got, ref)
    assert got.sum() > 0.99 * n


def test_sum_sq_and_gather(gpu):
    rng = np.random.default_rng(8)
    a = rng.normal(0.3, 1.0, 1_000_003).astype(np.float32)
    t = torch.from_numpy(a).cuda()
    out = C.c_double()
    L.call("aw_sum_sq", C.c_void_p(t.data_ptr()), a.size, C.c_double(0.25), C.byref(out), None)
    ref = np.sum((a.astype(np.float64) - 0.25) ** 2)
    assert out.value == pytest.approx(ref, rel=1e-12)
    L.call("aw_sum_sq", C.c_void_p(t.data_ptr()), a.size, C.c_double(0.25), C.byref(out), None)
    assert out.value == pytest.approx(ref, rel=1e-12)
    idx = np.sort(rng.choice(a.size, 100_000, replace=False)).astype(np.int32)
    ti = torch.from_numpy(idx).cuda()
    dst = torch.empty(idx.size, dtype=torch.float32, device="cuda")
    L.call("aw_gather_f32", C.c_void_p(t.data_ptr()), C.c_void_p(ti.data_ptr()), idx.size,
           C.c_void_p(dst.data_ptr()), None)
    torch.cuda.synchronize()
    assert np.array_equal(dst.cpu().numpy(), a[idx])
