"""GPU Darwin ensembles: a uniform batch of runs stepped by one aw_sim (every
particle kernel and batched rocFFT call spans the runs) gives each run the
same sources, fields, particles and diagnostics as that run stepped alone."""
import math

import numpy as np
import pytest

from arcwarden_amd import _lib as L
from arcwarden_amd.ensemble import Ensemble, RunSpec
from arcwarden_amd.sim import SimConfig, Simulation

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi


def whistler_like(vth, seed, nx=64, ny=32, ppc=32, modes=()):
    """A small pumped, magnetised Darwin run (shared step parameters; vth,
    seed and modes differ between runs)."""
    g = L.grid(nx, ny, nx * 0.5, ny * 0.5)
    rp = L.run_params(dt=0.05, qm=-1.0, eps0=1.0, c=5.0, n0=1.0, noisy_load=1, rng_seed=seed,
                      B0=(0.4, 0.0, 0.23), dump_every=3)
    rp.pump = 1
    rp.pump_ex, rp.pump_ey, rp.pump_ez = 1.39e-3, 1.81e-3, -1.81e-3
    rp.pump_k0 = TWO_PI * 3 / (nx * 0.5)
    rp.pump_w0 = 0.15
    rp.pump_trmp, rp.pump_toff = 2.0, 100.0
    sp = [L.species("electrons", 1.0, ppc, (vth,) * 3, (0.0, 0.0, 0.05))]
    return RunSpec(g, rp, sp, has_b0=True, mode_amp=tuple(modes))


def _multiset(st):
    cols = [st[k] for k in ("x", "y", "ux", "uy", "uz")]
    order = np.lexsort(cols[::-1])
    return np.stack([c[order] for c in cols])


@pytest.mark.parametrize("current", [L.AW_CURRENT_CIC, L.AW_CURRENT_ESIRKEPOV])
def test_darwin_batch_equals_single_runs(gpu, current):
    runs = [whistler_like(0.05, 11), whistler_like(0.12, 12, modes=(0.02, -0.01)),
            whistler_like(0.3, 13)]
    cfg = SimConfig(has_b0=True, field_model=L.AW_MODEL_DARWIN, sort_every=5,
                    stray_frac=1.0, current_deposit=current)
    nsteps = 10
    ens = Ensemble(runs, cfg)
    ens.init()
    ens.run(0, nsteps)
    info = ens.info()
    assert info["nruns"] == 3 and info["particles"] == 3 * 64 * 32 * 32
    for k, r in enumerate(runs):
        sim = Simulation(r.grid, r.rp, r.species,
                         SimConfig(has_b0=True, field_model=L.AW_MODEL_DARWIN, sort_every=5,
                                   stray_frac=1.0, current_deposit=current,
                                   mode_amp=list(r.mode_amp)))
        sim.init()
        for n in range(nsteps):
            sim.step(n)
        fs, fe = sim.fields(), ens.fields(k)
        for name in ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "ELx", "ELy", "ETx", "ETy", "ETz"):
            assert np.array_equal(fs[name], fe[name]), (k, name)
        assert np.array_equal(sim.sources_all()["rho"], ens.rho(k)), k
        ps, pe = sim.particles(), ens.particles(k)
        for name in ("x", "y", "ux", "uy", "uz"):   # same sort cadence: same order
            assert np.array_equal(ps[name], pe[name]), (k, name)
        hs, he = sim.history(), ens.history(k)
        assert [h["step"] for h in hs] == [h["step"] for h in he]
        for a, b in zip(hs, he):
            assert a["ee"] == b["ee"] and a["charge"] == b["charge"], k
            assert a["ke"] == pytest.approx(b["ke"], rel=1e-12)
        sim.close()
    ens.close()


def test_darwin_batch_adaptive_sort_keeps_results(gpu):
    """Re-sorting on strays (timing differs from a single run's) changes only
    the particle order."""
    runs = [whistler_like(1.0, 21), whistler_like(1.2, 22)]   # ~0.1 cell per step
    base = dict(has_b0=True, field_model=L.AW_MODEL_DARWIN)
    ens = Ensemble(runs, SimConfig(sort_every=0, stray_frac=1e-4, **base))
    ens.init()
    ens.run(0, 30)
    for k, r in enumerate(runs):
        sim = Simulation(r.grid, r.rp, r.species, SimConfig(sort_every=0, stray_frac=1.0, **base))
        sim.init()
        for n in range(30):
            sim.step(n)
        assert np.array_equal(sim.sources_all()["rho"], ens.rho(k))
        fs, fe = sim.fields(), ens.fields(k)
        for name in ("Ex", "Bz"):
            assert np.array_equal(fs[name], fe[name])
        assert np.array_equal(_multiset(sim.particles()), _multiset(ens.particles(k)))
        sim.close()
    assert ens.info()["sorts"] > 1
    ens.close()


def test_darwin_batch_rejects_mixed_runs(gpu):
    cfg = SimConfig(has_b0=True, field_model=L.AW_MODEL_DARWIN)
    a = whistler_like(0.05, 1)
    b = whistler_like(0.05, 2, nx=32)
    with pytest.raises(L.AwError):
        Ensemble([a, b], cfg)
    c = whistler_like(0.05, 3)
    c.rp.dt = 0.04
    with pytest.raises(L.AwError):
        Ensemble([a, c], cfg)
    d = whistler_like(0.05, 4, ppc=16)
    with pytest.raises(L.AwError):
        Ensemble([a, d], cfg)
