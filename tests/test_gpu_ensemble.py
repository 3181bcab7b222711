"""GPU tests of the batched ensemble: every run of an ensemble (eager and
HIP-graph replay) gives the same results as the same run stepped alone by
aw_sim -- rho bit for bit, the particle multiset bit for bit, the field
diagnostics exactly -- including runs that share a grid size (one batched
FFT) and a re-sort cadence shorter than the run."""
import numpy as np
import pytest

from arcwarden_amd import _lib as L
from arcwarden_amd.ensemble import Ensemble, random_ensemble, two_beam_run
from arcwarden_amd.sim import SimConfig, Simulation

pytestmark = pytest.mark.gpu


def _runs():
    spec = [(32, 16, 0.02, 0.2), (48, 64, 0.01, 0.15), (32, 32, 0.03, 0.25),
            (64, 16, 0.005, 0.3), (32, 64, 0.04, 0.1)]
    return [two_beam_run(s, p, t, v, seed=100 + i, dump_every=3)
            for i, (s, p, t, v) in enumerate(spec)]


def _multiset(st):
    cols = [st[k] for k in ("x", "y", "ux", "uy", "uz")]
    order = np.lexsort(cols[::-1])
    return np.stack([c[order] for c in cols])


@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_matches_single_runs(gpu, graph):
    runs = _runs()
    nsteps = 14
    ens = Ensemble(runs, SimConfig(sort_every=5), graph=graph)
    ens.init()
    info = ens.info()
    assert info["nruns"] == 5 and info["ngroups"] == 3  # 32^2 x3, 48^2, 64^2
    sims = []
    for r in runs:
        s = Simulation(r.grid, r.rp, r.species, SimConfig(has_b0=False, sort_every=5))
        s.init()
        sims.append(s)
    for n in range(nsteps):
        ens.step(n)
        for s in sims:
            s.step(n)
        if n < 6 or n == nsteps - 1:  # every replay checked against the eager result
            for i, s in enumerate(sims):
                assert np.array_equal(ens.rho(i), s.rho()), (n, i)
    for i, s in enumerate(sims):
        assert np.array_equal(_multiset(ens.particles(i)), _multiset(s.particles())), i
        f, fs = ens.fields(i), s.fields()
        assert np.array_equal(f["Ex"], fs["Ex"]) and np.array_equal(f["Ey"], fs["Ey"]), i
        he, hs = ens.history(i), s.history()
        assert [h["step"] for h in he] == [h["step"] for h in hs] == list(range(0, nsteps, 3))
        # same fields and particles; the sums differ only by reduction order
        for a, b in zip(he, hs):
            assert a["max_e"] == b["max_e"]
            for k in ("ke", "ee", "charge"):
                assert a[k] == pytest.approx(b[k], rel=1e-12), k
        s.close()
    info = ens.info()
    assert info["sorts"] == 2
    if graph:
        assert info["graph"] == 1 and info["graph_launches"] == nsteps - 1
    ens.close()


@pytest.mark.parametrize("graph", [False, True])
def test_five_grid_sizes_share_the_solve_streams(gpu, graph):
    """Five grid sizes: the per-size solves run on the main stream and two side
    streams, so two streams carry two sizes each; every run still equals the
    same run stepped alone (rho, E and the particle multiset bit for bit)."""
    spec = [(16, 16, 0.02, 0.2), (24, 16, 0.01, 0.15), (32, 16, 0.03, 0.25), (40, 16, 0.005, 0.3),
            (48, 16, 0.04, 0.1), (16, 32, 0.02, 0.12)]
    runs = [two_beam_run(sd, p, t, v, seed=400 + i) for i, (sd, p, t, v) in enumerate(spec)]
    nsteps = 9
    ens = Ensemble(runs, SimConfig(sort_every=4), graph=graph)
    ens.init()
    assert ens.info()["ngroups"] == 5
    ens.run(0, nsteps)
    for i, r in enumerate(runs):
        s = Simulation(r.grid, r.rp, r.species, SimConfig(has_b0=False, sort_every=4))
        s.init()
        s.run(nsteps)
        assert np.array_equal(ens.rho(i), s.rho()), i
        f, fs = ens.fields(i), s.fields()
        assert np.array_equal(f["Ex"], fs["Ex"]) and np.array_equal(f["Ey"], fs["Ey"]), i
        assert np.array_equal(_multiset(ens.particles(i)), _multiset(s.particles())), i
        s.close()
    ens.close()


@pytest.mark.parametrize("graph", [False, True])
def test_stray_sort_is_per_run(gpu, graph):
    """Only the run whose particles strayed beyond the threshold is re-sorted
    (fast beams: 0.2 cells per step); the slow runs keep their order.  Every
    run still equals the same run stepped alone, and the single runs' own
    stray-driven sorts reorder nothing that rho, E or the particle multiset
    can see."""
    runs = [two_beam_run(32, 16, 0.01, 0.01, seed=301), two_beam_run(32, 16, 0.01, 1.0, seed=302),
            two_beam_run(48, 16, 0.01, 0.01, seed=303), two_beam_run(64, 16, 0.01, 0.01, seed=304)]
    nsteps = 30
    cfg = dict(sort_every=0, stray_frac=0.01)
    ens = Ensemble(runs, SimConfig(**cfg), graph=graph)
    ens.init()
    ens.run(0, nsteps)
    info = ens.info()
    assert info["ngroups"] == 3
    assert info["sorts"] >= 1 and info["run_sorts"] == info["sorts"], info
    for i, r in enumerate(runs):
        s = Simulation(r.grid, r.rp, r.species, SimConfig(has_b0=False, **cfg))
        s.init()
        s.run(nsteps)
        assert np.array_equal(ens.rho(i), s.rho()), i
        f, fs = ens.fields(i), s.fields()
        assert np.array_equal(f["Ex"], fs["Ex"]) and np.array_equal(f["Ey"], fs["Ey"]), i
        pe, ps = ens.particles(i), s.particles()
        assert np.array_equal(_multiset(pe), _multiset(ps)), i
        if i != 1:  # never re-sorted, here or alone: the same order
            for k in ("x", "y", "ux", "uy", "uz"):
                assert np.array_equal(pe[k], ps[k]), (i, k)
        s.close()
    ens.close()


@pytest.mark.parametrize("graph", [False, True])
def test_single_run_ensemble(gpu, graph):
    """One run: one grid size, so no side streams and no fork or join; it
    equals the run stepped alone."""
    r = two_beam_run(32, 16, 0.02, 0.2, seed=500, dump_every=4)
    ens = Ensemble([r], SimConfig(sort_every=3), graph=graph)
    ens.init()
    ens.run(0, 7)
    s = Simulation(r.grid, r.rp, r.species, SimConfig(has_b0=False, sort_every=3))
    s.init()
    s.run(7)
    assert ens.info()["nruns"] == 1 and ens.info()["ngroups"] == 1
    assert np.array_equal(ens.rho(0), s.rho())
    assert np.array_equal(_multiset(ens.particles(0)), _multiset(s.particles()))
    assert [h["step"] for h in ens.history(0)] == [h["step"] for h in s.history()]
    s.close()
    ens.close()


def test_random_ensemble_generator_and_invariants(gpu):
    """The 24-run workload generator (seeded) and a short run of a scaled-down
    draw: charge conserved in every run."""
    runs = random_ensemble(nruns=24)
    sides = [r.grid.nx for r in runs]
    assert len(runs) == 24 and set(sides) <= {64, 96, 128, 192, 256}
    small = random_ensemble(nruns=6, sides=(32, 48), ppcs=(8, 16), dump_every=2)
    ens = Ensemble(small, SimConfig(sort_every=4), graph=True)
    ens.init()
    ens.run(0, 9)
    for i in range(len(small)):
        h = ens.history(i)
        q = [x["charge"] for x in h]
        assert max(q) - min(q) <= 1e-6 * abs(q[0])
        assert q[0] == pytest.approx(-small[i].grid.nx * small[i].grid.ny * 0.25, rel=1e-6)
    ens.close()


def test_ensemble_autotune_changes_no_result(gpu):
    """cfg.autotune picks each run's tile and workgroups per tile (the
    autotuner of a run stepped alone); rho, fields and the particle multisets
    stay bit-identical to the heuristic configuration."""
    runs = _runs()
    a = Ensemble(runs, SimConfig(sort_every=5), graph=True)
    b = Ensemble(runs, SimConfig(sort_every=5, autotune=True), graph=True)
    for e in (a, b):
        e.init()
        e.run(0, 12)
    for i in range(len(runs)):
        assert np.array_equal(a.rho(i), b.rho(i)), i
        fa, fb = a.fields(i), b.fields(i)
        assert np.array_equal(fa["Ex"], fb["Ex"]) and np.array_equal(fa["Ey"], fb["Ey"]), i
        assert np.array_equal(_multiset(a.particles(i)), _multiset(b.particles(i))), i
    a.close()
    b.close()


def test_ensemble_init_twice_is_refused(gpu):
    e = Ensemble(_runs()[:2], SimConfig(sort_every=5))
    e.init()
    with pytest.raises(L.AwError, match="twice"):
        e.init()
    e.close()
