"""GPU loader parity: aw_load_species against the oracle's restatement of
species_init_kernel (include/pic/particles.hpp:333-379), slot for slot.

Both sides store particle (cell c, slot s) of species k at
base_k + s*nx*ny + c, so the arrays compare element by element.  Positions
without modes are pure FP32/integer arithmetic of the same draws and must be
bit-exact.  The mode displacement and the Maxwellian velocities go through a
double-precision sin / erfinv on each side (device ocml against the oracle's
Winitzki start + Halley steps), so they may differ by the last unit of the
FP32 result; the test bounds that at one FP32 ulp and checks that almost all
values agree bit for bit."""
import ctypes as C

import numpy as np
import pytest

from arcwarden_amd import _lib as L

from gpu_util import Ctx, DevParticles, stream

pytestmark = pytest.mark.gpu


def _load_gpu(g, rp, species, modes):
    n = sum(s.ppc for s in species) * g.nx * g.ny
    p = DevParticles.empty(n)
    sp = (L.AwSpecies * len(species))(*species)
    w = (C.c_double * len(species))()
    amp = (C.c_double * max(1, len(modes)))(*modes) if modes else None
    with Ctx(g) as ctx:
        L.call("aw_load_species", ctx.h, sp, len(species), C.byref(rp), amp, len(modes),
               C.byref(p.struct()), w, stream())
        out = p.numpy()
    return out, list(w)


def _ulps(a, b):
    """Distance in FP32 units of the last place between same-sign values."""
    ai = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    bi = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    same_sign = np.sign(np.asarray(a)) == np.sign(np.asarray(b))
    d = np.abs(ai - bi)
    return np.where(same_sign | (np.asarray(a) == np.asarray(b)), d, np.int64(1) << 40)


SPECIES = [
    # name, density, ppc, uth, ufl
    ("beam_plus", 0.5, 16, (0.05, 0.02, 0.01), (0.2, 0.0, 0.0)),
    ("beam_minus", 0.5, 8, (0.03, 0.03, 0.03), (-0.2, 0.05, 0.0)),
    ("hot", 0.25, 4, (0.4, 0.3, 0.2), (0.0, 0.0, 0.1)),
]


@pytest.mark.parametrize("noisy", [0, 1])
@pytest.mark.parametrize("modes", [[], [2e-2, -1e-2, 5e-3]])
def test_load_species_matches_oracle(gpu, oracle, noisy, modes):
    nx, ny, dx = 48, 20, 0.5
    g = L.grid(nx, ny, nx * dx, ny * dx)
    rp = L.run_params(dt=0.1, qm=-1.0, noisy_load=noisy, rng_seed=20260627)
    species = [L.species(nm, d, ppc, uth, ufl) for nm, d, ppc, uth, ufl in SPECIES]
    got, w = _load_gpu(g, rp, species, modes)
    og = oracle.grid(nx, ny, dx, dx)
    ref, wref = oracle.load_species(og, [(d, ppc, uth, ufl) for _, d, ppc, uth, ufl in SPECIES],
                                    bool(noisy), 20260627, mode_amp=modes or None)
    assert w == pytest.approx(wref, rel=0, abs=0)
    assert np.array_equal(got["y"], ref["y"])
    if not modes:
        assert np.array_equal(got["x"], ref["x"])
    else:
        d = _ulps(got["x"], ref["x"])
        assert d.max() <= 1, d.max()
        assert np.mean(d == 0) > 0.999
        # the displacement is really applied: the modes move x by up to the amplitudes
        base, _ = oracle.load_species(og, [(d_, ppc, uth, ufl) for _, d_, ppc, uth, ufl in SPECIES],
                                      bool(noisy), 20260627)
        shift = (got["x"] - base["x"] + 0.5 * nx) % nx - 0.5 * nx
        # max |sum_m A_m sin(m k1 x)| >= its rms >= |A_1|/sqrt(2)
        assert np.abs(shift).max() > 0.5 * abs(modes[0]) / dx
    for k in ("ux", "uy", "uz"):
        d = _ulps(got[k], ref[k])
        assert d.max() <= 1, (k, d.max())
        assert np.mean(d == 0) > 0.99, (k, np.mean(d == 0))
    # positions in range and each species' quiet load places exactly ppc per cell
    assert (got["x"] >= 0).all() and (got["x"] < nx).all()
    assert (got["y"] >= 0).all() and (got["y"] < ny).all()
    if not noisy and not modes:
        base = 0
        for _, _, ppc, _, _ in SPECIES:
            sl = slice(base, base + ppc * nx * ny)
            cells = np.floor(got["y"][sl]).astype(np.int64) * nx + np.floor(got["x"][sl]).astype(
                np.int64)
            assert (np.bincount(cells, minlength=nx * ny) == ppc).all()
            base += ppc * nx * ny
