"""GPU parity at the workloads themselves (not scaled-down draws):

- the Darwin whistler-pump run (512x128, 256 ppc = 16,777,216 electrons,
  Esirkepov current, pump, B0 at 30 degrees): one step against the FP64
  oracle -- every source, every field, the particle state -- and the 3-run
  batch against the runs stepped alone, bit for bit;
- the Darwin whistler eigenmode: the cold parallel whistler rotates at
  w = w_ce k^2 c^2 / (1 + k^2 c^2) (tests/test_yee_vs_darwin_lowfreq.cu, the
  Darwin half), within 2%;
- the 24-run batched ES ensemble (seed 20260629, 89,210,880 particles):
  every run against the same run stepped alone, and one run of each grid
  size against the FP64 oracle's step."""
import math

import numpy as np
import pytest

from arcwarden_amd import _lib as L
from arcwarden_amd.ensemble import Ensemble, RunSpec, random_ensemble
from arcwarden_amd.sim import SimConfig, Simulation
from arcwarden_amd.workloads import WHISTLER_RATIOS, whistler_run

from gpu_util import rel_l2

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi


def _pump_grid(rp, g, t):
    x = np.arange(g.nx) * g.dx
    th = rp.pump_k0 * x - rp.pump_w0 * t
    r = t / rp.pump_trmp if t < rp.pump_trmp else 1.0
    return [np.tile(v, (g.ny, 1)) for v in (rp.pump_ex * np.cos(th) * r,
                                           -rp.pump_ey * np.sin(th) * r,
                                           rp.pump_ez * np.cos(th) * r)]


def _periodic_rel_l2(a, b, n):
    d = (np.asarray(a, np.float64) - np.asarray(b, np.float64) + 0.5 * n) % n - 0.5 * n
    return np.linalg.norm(d) / np.linalg.norm(np.asarray(b, np.float64))


def _darwin_cfg(**kw):
    return SimConfig(has_b0=True, field_model=L.AW_MODEL_DARWIN,
                     current_deposit=L.AW_CURRENT_ESIRKEPOV, **kw)


def test_config5_whistler_step_matches_fp64_oracle(gpu, oracle):
    g, rp, sp = whistler_run(WHISTLER_RATIOS[0], dump_every=0)
    # no re-sort inside the measured steps: s0, s1, s2 keep one particle order
    sim = Simulation(g, rp, sp, _darwin_cfg(sort_every=0, stray_frac=1.0))
    sim.init()
    n_step = 2
    for n in range(n_step):
        sim.step(n)
    s0 = sim.particles()
    w = s0["w_uniform"]
    assert s0["n"] == 512 * 128 * 256
    sim.step(n_step)
    s1 = sim.particles()
    sim.step(n_step + 1)   # solves its fields from the sources of s1 and pushes s1
    s2 = sim.particles()
    srcs = sim.sources_all()
    f = sim.fields()
    sim.close()
    og = oracle.grid(g.nx, g.ny, g.dx, g.dy)
    # sources of step n_step + 1: rho, amu CIC at x^{n+1}; J the Esirkepov
    # current of the move x^n -> x^{n+1}
    jx64, jy64, jz64 = oracle.esirkepov_f64(og, s0["x"], s0["y"], s1["x"], s1["y"], s1["uz"],
                                            -1.0, rp.dt, w_uniform=w)
    for name, ref in (("Jx", jx64), ("Jy", jy64), ("Jz", jz64)):
        assert rel_l2(srcs[name], ref) < 1e-6, name
    rho64 = oracle.deposit_rho_f64(og, s1["x"], s1["y"], -1.0, w_uniform=w)
    assert rel_l2(srcs["rho"], rho64) < 1e-6
    amu64 = oracle.deposit_amu_f64(og, s1, -1.0, w_uniform=w)
    for q in range(4):
        assert rel_l2(srcs[f"amu{q}"], amu64[q]) < 2e-6, q
    # fields: E_L, B (staggered current shifted onto the nodes) + B0, pump, dcu, E_T
    el64 = oracle.solve_es_f64(og, rho64)
    bself = oracle.solve_darwin_b_f64(og, jx64, jy64, jz64, c=rp.c, staggered=True)
    b64 = [b + rp.B0[i] for i, b in enumerate(bself)]
    t = (n_step + 1) * rp.dt
    p = _pump_grid(rp, g, t)
    ed64 = (el64[0] + p[0], el64[1] + p[1], p[2])
    dcu64 = oracle.deposit_dcu_f64(og, s1, -1.0, [a.astype(np.float32) for a in ed64],
                                   [b.astype(np.float32) for b in b64], w_uniform=w)
    for name, ref in zip(("dcux", "dcuy", "dcuz"), dcu64):
        assert rel_l2(srcs[name], ref) < 1e-5, name
    et64 = oracle.solve_darwin_et_f64(og, dcu64, amu64, c=rp.c, n0=rp.n0)
    e64 = (el64[0] + et64[0] + p[0], el64[1] + et64[1] + p[1], et64[2] + p[2])
    assert rel_l2(f["ELx"], el64[0]) < 1e-5 and rel_l2(f["ELy"], el64[1]) < 1e-5
    for q, name in enumerate(("Bx", "By", "Bz")):
        # the self field alone (B0 would hide its error).  The engine holds
        # the total B0 + B_self in FP32, which resolves the self field to
        # about one ulp of |B0| (3e-8 at 0.43, 1.4e-3 of this run's 2e-5
        # self field): the bound is 1e-5 of the self field plus two ulps
        err = f[name].astype(np.float64) - (rp.B0[q] + bself[q])
        ulp = float(np.spacing(np.float32(abs(rp.B0[q]) + np.abs(bself[q]).max())))
        rms = lambda a: float(np.sqrt(np.mean(np.square(a))))  # noqa: E731
        print(name, "rms error", rms(err), "self field rms", rms(bself[q]), "ulp", ulp)
        assert rms(err) < 1e-5 * rms(bself[q]) + 2 * ulp, name
    for name, ref in zip(("ETx", "ETy", "ETz"), et64):
        assert rel_l2(f[name], ref) < 1e-4, name
    for name, ref in zip(("Ex", "Ey", "Ez"), e64):
        assert rel_l2(f[name], ref) < 1e-5, name
    # the EM push of s1 with the FP64 fields
    ref = oracle.push_em_f64(og, -1.0, rp.dt, s1, e64, b64)
    for k in ("ux", "uy", "uz"):
        assert rel_l2(s2[k], ref[k]) < 1e-6, k
    assert _periodic_rel_l2(s2["x"], ref["x"], g.nx) < 1e-6
    assert _periodic_rel_l2(s2["y"], ref["y"], g.ny) < 1e-6


def test_config5_batch_equals_single_runs(gpu):
    """The three whistler-pump runs (v_r/v_th 3.2, 2.1, 1.0) at full size in
    one batched aw_sim against each run stepped alone: every field grid, rho
    and the particles bit for bit, and the energies."""
    nsteps = 6
    specs = []
    for r in WHISTLER_RATIOS:
        g, rp, sp = whistler_run(r, dump_every=3)
        specs.append((g, rp, sp))
    ens = Ensemble([RunSpec(g, rp, sp, has_b0=True) for g, rp, sp in specs],
                   _darwin_cfg(sort_every=4, stray_frac=1.0))
    ens.init()
    ens.run(0, nsteps)
    assert ens.info()["particles"] == 3 * 512 * 128 * 256
    for k, (g, rp, sp) in enumerate(specs):
        sim = Simulation(g, rp, sp, _darwin_cfg(sort_every=4, stray_frac=1.0))
        sim.init()
        for n in range(nsteps):
            sim.step(n)
        fs, fe = sim.fields(), ens.fields(k)
        for name in ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "ELx", "ELy", "ETx", "ETy", "ETz"):
            assert np.array_equal(fs[name], fe[name]), (k, name)
        assert np.array_equal(sim.sources_all()["rho"], ens.rho(k)), k
        ps, pe = sim.particles(), ens.particles(k)
        for name in ("x", "y", "ux", "uy", "uz"):
            assert np.array_equal(ps[name], pe[name]), (k, name)
        del ps, pe
        hs, he = sim.history(), ens.history(k)
        assert [h["step"] for h in hs] == [h["step"] for h in he] == [0, 3]
        for a, b in zip(hs, he):
            assert a["ee"] == b["ee"] and a["charge"] == b["charge"], k
        sim.close()
    ens.close()


def _project(st, k, dx):
    th = k * st["x"].astype(np.float64) * dx
    return np.sum((st["uy"].astype(np.float64) + 1j * st["uz"].astype(np.float64))
                  * np.exp(1j * th))


def _fit_omega(C, dt_s):
    ph = np.unwrap(np.angle(np.asarray(C)))
    t = np.arange(len(C)) * dt_s
    return abs(np.polyfit(t, ph, 1)[0])


@pytest.mark.parametrize("current", [L.AW_CURRENT_CIC, L.AW_CURRENT_ESIRKEPOV])
def test_darwin_whistler_eigenmode_frequency(gpu, current):
    """tests/test_yee_vs_darwin_lowfreq.cu, Darwin half: cold electrons in
    B0 = w_ce x, a right-hand circular velocity perturbation u_y + i u_z =
    v1 e^{-ikx} at k c = 3 (mode 2); the projection sum (u_y + i u_z) e^{ikx}
    rotates at w = w_ce k^2 c^2/(1 + k^2 c^2) = 0.45, fitted over about five
    periods, within 2%.  The grid is 8 cells high instead of 1 (the physics
    does not depend on y)."""
    nx, ny, ppc, mode = 128, 8, 64, 2
    c, wce, uth, v1, k = 10.0, 0.5, 0.01, 0.05, 0.3
    Lx = TWO_PI * mode / k
    dx = Lx / nx
    g = L.grid(nx, ny, Lx, ny * dx)
    rp = L.run_params(dt=0.05, qm=-1.0, eps0=1.0, c=c, n0=1.0, noisy_load=0,
                      B0=(wce, 0.0, 0.0), wce=wce, dump_every=100000)
    sp = [L.species("e", 1.0, ppc, (uth,) * 3, (0.0, 0.0, 0.0))]
    sim = Simulation(g, rp, sp, SimConfig(has_b0=True, field_model=L.AW_MODEL_DARWIN,
                                          current_deposit=current))
    sim.init()
    st = sim.particles()
    th = k * st["x"].astype(np.float64) * dx
    st["uy"] = (st["uy"] + v1 * np.cos(th)).astype(np.float32)
    st["uz"] = (st["uz"] - v1 * np.sin(th)).astype(np.float32)
    sim.set_particles(st)
    nsteps, stride = int(round(70.0 / rp.dt)), 2
    C = []
    for n in range(nsteps):
        sim.step(n)
        if n % stride == 0:
            C.append(_project(sim.particles(), k, dx))
    sim.close()
    w = _fit_omega(C, stride * rp.dt)
    kc2 = (k * c) ** 2
    w_dar = wce * kc2 / (1.0 + kc2)
    print("whistler", current, w, w_dar, abs(w - w_dar) / w_dar)
    assert abs(w - w_dar) / w_dar < 0.02
    # the mode keeps its amplitude (no numerical damping or growth beyond 10%)
    assert abs(abs(C[-1]) / abs(C[0]) - 1.0) < 0.1


def _hash_multiset(st):
    """Sorted 64-bit hashes of the particles' (x, y, ux, uy, uz) bit patterns."""
    h = np.zeros(st["x"].size, np.uint64)
    for k in ("x", "y", "ux", "uy", "uz"):
        v = st[k].view(np.uint32).astype(np.uint64)
        h = (h * np.uint64(0x100000001B3)) ^ (v * np.uint64(0x9E3779B97F4A7C15))
    return np.sort(h)


def test_config3_ensemble_matches_single_runs_and_oracle(gpu, oracle):
    """The 24-run ensemble (89,210,880 particles), HIP-graph replay, 20 steps
    with stray-triggered per-run re-sorts.  Step 0 of one run of each grid
    size agrees with the FP64 oracle (rho 1e-6, E 1e-5, particle state
    1e-6); after 20 steps every run's rho, fields, particle multiset and
    diagnostics equal the same run stepped alone by aw_sim."""
    runs = random_ensemble(nruns=24, dump_every=5)
    nsteps = 20
    ens = Ensemble(runs, SimConfig(), graph=True)
    ens.init()
    info = ens.info()
    assert info["nruns"] == 24 and info["particles"] == 89_210_880
    first = {}
    for i, r in enumerate(runs):
        first.setdefault(r.grid.nx, i)
    assert len(first) == 5
    before = {i: ens.particles(i) for i in first.values()}
    ens.step(0)
    ens.synchronize()
    for i, s0 in before.items():
        r, g = runs[i], runs[i].grid
        og = oracle.grid(g.nx, g.ny, g.dx, g.dy)
        rho64 = oracle.deposit_rho_f64(og, s0["x"], s0["y"], -1.0, w_uniform=s0["w_uniform"])
        assert rel_l2(ens.rho(i), rho64) < 1e-6, i
        ex64, ey64 = oracle.solve_es_f64(og, rho64)
        fe = ens.fields(i)
        assert rel_l2(fe["Ex"], ex64) < 1e-5 and rel_l2(fe["Ey"], ey64) < 1e-5, i
        pp = oracle.push_params(-1.0, r.rp.dt, (0, 0, 0), r.has_b0)
        ref = oracle.push_es_f64(og, pp, s0, ex64, ey64, True)
        pe = ens.particles(i)   # no re-sort before step 0: same order as s0
        assert _periodic_rel_l2(pe["x"], ref["x"], g.nx) < 1e-6, i
        assert _periodic_rel_l2(pe["y"], ref["y"], g.ny) < 1e-6, i
        for k in ("ux", "uy"):
            assert rel_l2(pe[k], ref[k]) < 1e-6, (i, k)
        del pe, ref
    del before
    ens.run(1, nsteps - 1)
    assert ens.info()["graph_launches"] >= nsteps - 1
    for i, r in enumerate(runs):
        sim = Simulation(r.grid, r.rp, r.species, SimConfig(has_b0=r.has_b0))
        sim.init()
        for n in range(nsteps):
            sim.step(n)
        pe, ps = ens.particles(i), sim.particles()
        same_order = all(np.array_equal(pe[k], ps[k]) for k in ("x", "y", "ux", "uy", "uz"))
        if not same_order:  # the re-sort points may differ; the deposits are order-free
            assert np.array_equal(_hash_multiset(pe), _hash_multiset(ps)), i
        del pe, ps
        assert np.array_equal(ens.rho(i), sim.rho()), i
        fe, fs = ens.fields(i), sim.fields()
        assert np.array_equal(fe["Ex"], fs["Ex"]) and np.array_equal(fe["Ey"], fs["Ey"]), i
        he, hs = ens.history(i), sim.history()
        assert [h["step"] for h in he] == [h["step"] for h in hs] == [0, 5, 10, 15]
        for a, b in zip(he, hs):
            assert a["max_e"] == b["max_e"], i
            for key in ("ke", "ee", "charge"):
                assert a[key] == pytest.approx(b[key], rel=1e-12), (i, key)
        sim.close()
    ens.close()
