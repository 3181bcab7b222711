#!/usr/bin/env python3
"""Headline benchmark: electrostatic 2D3V PIC steps on one MI355X.

Workload (the measured configuration): ES thermal plasma, 512x512 periodic
grid, dx = 0.5, CIC, 256 electrons per cell (67,108,864), Maxwellian
v_th = 0.05 in 3 components, uniform B0 = 0.5 z (magnetised Boris), FP32 SoA
particles, dt = 0.1.  A step is one Simulation::step: rho^n -> FFT Poisson ->
fused gather+push+move+wrap+deposit, the stable tile re-sort whenever more than
2% of the particles left their tile apron (--sort-every N: every N steps), and
the scalar diagnostics every 10 steps, all inside the timed region.  The same
run with a re-sort every 20 steps is reported beside it.

Prints ONE JSON line.  ``value`` is particle-steps/s over the whole timed
region; ``roofline`` prices the dominant kernel (push_deposit) at its
algorithmic particle traffic against 8 TB/s: x, y, ux, uy read and written
(32 B per particle; with B0 along z, uz is a constant of the motion), plus
4 B to read uz on the diagnostics steps, which sum the kinetic energy;
``cpu_baseline`` times the oracle's FP32 same-algorithm ES step with OpenMP
on this host's cores on the same particle state.  ``ensemble`` compares a
24-run batched ensemble with the same runs stepped one by one; ``adjoint``
times the field-energy gradient of a 128^2 two-stream run against its plain
forward run; ``darwin`` times the three-run Darwin whistler-pump scan with
the Esirkepov current deposit (``darwin_cic_current``: with the reference's
CIC current); ``yee_eaw_case7`` runs the reference's own Yee benchmark deck (and
``yee_eaw_case7_flat`` the same run without its tile sort)
and ``yee`` its delta-f whistler line; ``checkpoint`` times a restart file
write and restore of the headline run, outside the timed region.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE = 32.0  # read + write x, y, ux, uy (FP32); uz invariant for B0 || z
BYTES_UZ_READ = 4.0        # uz read on diagnostics steps (kinetic energy)


def traffic_from_profiles():
    """HBM bytes per push_deposit launch measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, KiB/dispatch, scaled by the copy-kernel
    calibration) -- read from the committed summary if present."""
    f = ROOT / "profiles" / "push_deposit_es_512x512_ppc256_pmc.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def cpu_threads() -> int:
    """Threads for the host baselines: OMP_NUM_THREADS when set (the GPU box
    sets it to its CPU share per GPU, 16), else every hardware thread."""
    try:
        return max(1, int(os.environ["OMP_NUM_THREADS"]))
    except (KeyError, ValueError):
        return os.cpu_count() or 1


def cpu_info() -> dict:
    """The host the CPU baselines ran on: model name, hardware threads and
    OMP_NUM_THREADS."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "hardware_concurrency": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(state, g_nx, g_ny, dx, w_uniform, rho, budget_s):
    import numpy as np
    sys.path.insert(0, str(ROOT / "oracle"))
    import pic_oracle as orc
    orc.build()
    threads = cpu_threads()
    g = orc.grid(g_nx, g_ny, dx, dx)
    pp = orc.push_params(-1.0, 0.1, (0.0, 0.0, 0.5), True)
    st = {k: np.ascontiguousarray(state[k], np.float32) for k in ("x", "y", "ux", "uy", "uz")}
    r = np.ascontiguousarray(rho, np.float32)
    n = st["x"].size
    steps, t0 = 0, time.perf_counter()
    while True:
        orc.es_step_f32(g, pp, st, w_uniform, 40, r, nthreads=threads)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 10:
            break
    out = {"value": n * steps / el, "unit": "particle-steps/s", "cores": threads,
           "kind": "port",
           "sample": f"oracle FP32 ES step (solve+push+wrap+deposit), OpenMP x{threads}, "
                     f"{steps} step(s) of the full {g_nx}x{g_ny}x256ppc particle set "
                     f"({n} particles) taken from the GPU state after init"}
    out.update(cpu_info())
    # the same step on one thread (the scalar host path), one step of the state
    st1 = {k: v.copy() for k, v in st.items()}
    t0 = time.perf_counter()
    orc.es_step_f32(g, pp, st1, w_uniform, 40, r.copy(), nthreads=1)
    el1 = time.perf_counter() - t0
    out["scalar"] = {"value": n / el1, "unit": "particle-steps/s", "cores": 1,
                     "sample": f"the same oracle step on one thread, 1 step of {n} particles"}
    return out


def two_stream_bench(steps: int = 200, cpu: bool = True):
    """The plain ES two-stream run (64x64, dx 0.5, beams +-0.2 with vth 0.01,
    64 ppc = 262,144 particles, x-displacement 1e-3 sin(k1 x), dt 0.1, seed
    20260627): 200 GPU steps, then the oracle's scalar FP32 step (one thread,
    the same algorithm) for the same 200 steps from the same initial state."""
    import numpy as np
    from arcwarden_amd import _lib as L
    from arcwarden_amd.sim import SimConfig, Simulation
    g = L.grid(64, 64, 32.0, 32.0)
    rp = L.run_params(dt=0.1, qm=-1.0, eps0=1.0, noisy_load=1, rng_seed=20260627,
                      dump_every=steps)
    sp = [L.species("beam_plus", 0.5, 32, (0.01, 0.01, 0.01), (0.2, 0, 0)),
          L.species("beam_minus", 0.5, 32, (0.01, 0.01, 0.01), (-0.2, 0, 0))]
    sim = Simulation(g, rp, sp, SimConfig(has_b0=False, mode_amp=[1e-3], sort_every=0))
    sim.init()
    state, rho = sim.particles(), sim.rho()
    n = state["x"].size
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run(steps)
    el = time.perf_counter() - t0
    ee_gpu = sim.history()[-1]["ee"]
    sim.close()
    out = {"workload": "ES two-stream 64x64, 64 ppc (262,144), 200 steps",
           "value": n * steps / el, "unit": "particle-steps/s", "ms_per_step": el / steps * 1e3}
    if cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pic_oracle as orc
        orc.build()
        og = orc.grid(64, 64, 0.5, 0.5)
        pp = orc.push_params(-1.0, 0.1, (0.0, 0.0, 0.0), False)
        st = {k: np.ascontiguousarray(state[k], np.float32) for k in ("x", "y", "ux", "uy", "uz")}
        r = np.ascontiguousarray(rho, np.float32)
        c0 = time.perf_counter()
        for _ in range(steps):
            orc.es_step_f32(og, pp, st, state["w_uniform"], 40, r, nthreads=1)
        cel = time.perf_counter() - c0
        out["cpu_baseline"] = {"value": n * steps / cel, "unit": "particle-steps/s", "cores": 1,
                               "kind": "port",
                               "sample": "oracle scalar FP32 ES step, one thread, the full "
                                         "200-step run from the GPU's initial state",
                               **cpu_info()}
        out["gpu_vs_cpu_scalar"] = (n * steps / el) / (n * steps / cel)
    out["ee_final"] = ee_gpu
    return out


def _ensemble_speedup(runs, steps: int, graph: bool):
    """Sequential (one aw_sim per run, run after run) and batched (one
    aw_ensemble) wall time of the same runs over `steps` steps."""
    from arcwarden_amd.ensemble import Ensemble
    from arcwarden_amd.sim import SimConfig, Simulation
    seq = 0.0
    for r in runs:
        sim = Simulation(r.grid, r.rp, r.species, SimConfig(has_b0=False, sort_every=0))
        sim.init()
        sim.synchronize()
        t0 = time.perf_counter()
        sim.run(steps)
        seq += time.perf_counter() - t0
        sim.close()
    ens = Ensemble(runs, SimConfig(sort_every=0), graph=graph)
    ens.init()
    ens.synchronize()
    t0 = time.perf_counter()
    ens.run(0, steps)
    bat = time.perf_counter() - t0
    info = ens.info()
    ens.close()
    return seq, bat, info


def ensemble_bench(steps: int, graph: bool = True, cpu: bool = False, cpu_steps: int = 5):
    """The 24-run batched ensemble (two-beam plasmas, seed 20260629: side in
    {64..256}, ppc in {16, 64, 256}, vth log-uniform [0.005, 0.05], v0 in
    [0.1, 0.3], dx 0.5, dt 0.1) stepped as one aw_ensemble versus the same
    runs stepped one after another by aw_sim.  Loading is outside both timed
    regions; each timed region ends with a device synchronisation.  cpu: the
    oracle's FP32 ES step (OpenMP) over cpu_steps steps of each of four runs
    (the first four distinct (side, ppc) shapes in draw order), from their GPU
    initial states."""
    from arcwarden_amd.ensemble import random_ensemble
    from arcwarden_amd.sim import SimConfig, Simulation
    runs = random_ensemble(dump_every=100)
    ntot = sum(r.grid.nx * r.grid.ny * sum(s.ppc for s in r.species) for r in runs)
    seq, bat, info = _ensemble_speedup(runs, steps, graph)
    print(f"# ensemble: sequential {seq:.3f}s, batched {bat:.3f}s for {steps} steps "
          f"(graph={graph})", file=sys.stderr, flush=True)
    # the same generator restricted to small runs (sides 64-128, ppc 16/64),
    # where each run alone is launch-bound: what batching gains there
    small = random_ensemble(dump_every=100, sides=(64, 96, 128), ppcs=(16, 64))
    nsmall = sum(r.grid.nx * r.grid.ny * sum(s.ppc for s in r.species) for r in small)
    sseq, sbat, _ = _ensemble_speedup(small, steps, graph)
    print(f"# small-run ensemble: sequential {sseq:.3f}s, batched {sbat:.3f}s",
          file=sys.stderr, flush=True)
    cpu_out = None
    if cpu:
        import numpy as np
        sys.path.insert(0, str(ROOT / "oracle"))
        import pic_oracle as orc
        orc.build()
        threads = cpu_threads()
        cps, cel, desc, shapes, subset = 0, 0.0, [], set(), []
        for r in runs:
            shape = (r.grid.nx, sum(sp.ppc for sp in r.species))
            if shape not in shapes and len(subset) < 4:
                shapes.add(shape)
                subset.append(r)
        for r in subset:
            sim = Simulation(r.grid, r.rp, r.species, SimConfig(has_b0=False, sort_every=0))
            sim.init()
            state, rho = sim.particles(), sim.rho()
            sim.close()
            g = r.grid
            og = orc.grid(g.nx, g.ny, g.dx, g.dy)
            pp = orc.push_params(r.rp.qm, r.rp.dt, (0.0, 0.0, 0.0), False)
            st = {k: np.ascontiguousarray(state[k], np.float32)
                  for k in ("x", "y", "ux", "uy", "uz")}
            rr = np.ascontiguousarray(rho, np.float32)
            c0 = time.perf_counter()
            for _ in range(cpu_steps):
                orc.es_step_f32(og, pp, st, state["w_uniform"], 40, rr, nthreads=threads)
            cel += time.perf_counter() - c0
            cps += st["x"].size * cpu_steps
            desc.append(f"{g.nx}^2 x {sum(sp.ppc for sp in r.species)} ppc")
        cpu_out = {"value": cps / cel, "unit": "particle-steps/s", "cores": threads,
                   "kind": "port",
                   "sample": f"oracle FP32 ES step (solve+push+wrap+deposit), OpenMP x{threads}, "
                             f"{cpu_steps} steps of each of four runs of the ensemble "
                             f"({', '.join(desc)}) from their GPU initial states",
                   **cpu_info()}
    return {"runs": len(runs), "particles": ntot, "steps": steps,
            "value": ntot * steps / bat, "unit": "particle-steps/s",
            "sequential_value": ntot * steps / seq, "speedup_vs_sequential": seq / bat,
            "graph": bool(graph), "size_groups": info["ngroups"],
            "sort_events": info["sorts"], "run_sorts": info["run_sorts"],
            "cpu_baseline": cpu_out,
            "small_runs": {"workload": "the same generator and seed with sides 64-128 and "
                                       "ppc 16/64 (24 runs)",
                           "particles": nsmall, "value": nsmall * steps / sbat,
                           "sequential_value": nsmall * steps / sseq,
                           "speedup_vs_sequential": sseq / sbat},
            "workload": "24 two-beam ES runs, seed 20260629, sides 64-256, ppc 16/64/256, "
                        "vth log-U[0.005,0.05], v0 U[0.1,0.3], dx 0.5, dt 0.1"}


def adjoint_bench(nsteps: int = 100, cpu_steps: int = 0):
    """Field-energy gradient of the differentiable two-stream run (128^2, dx
    0.5, 64 ppc as beams +-0.2 with vth 0.02, 16 x-displacement modes
    A_m ~ 1e-3 U[-1,1] seed 20260630, dt 0.1, 100 steps, checkpoints every 10):
    forward with checkpoints, backward sweep, and the plain forward run of
    the same state by aw_sim, each timed with a device synchronisation.
    cpu_steps > 0 also times the oracle's FP64 host adjoint (forward + reverse
    sweep, OpenMP) of the same initial state over that many steps."""
    from arcwarden_amd import _lib as L
    from arcwarden_amd.adjoint import two_stream_adjoint
    from arcwarden_amd.sim import SimConfig, Simulation
    run, A, ufl = two_stream_adjoint()
    run.forward(A, ufl, nsteps)  # warm-up (plans, allocations)
    run.backward()
    t0 = time.perf_counter()
    loss = run.forward(A, ufl, nsteps)
    t1 = time.perf_counter()
    gA, gU = run.backward()
    t2 = time.perf_counter()
    rp = L.run_params(dt=0.1, qm=-1.0, eps0=1.0, noisy_load=1, rng_seed=20260630)
    sim = Simulation(run.grid, rp, run.species, SimConfig(has_b0=False, mode_amp=list(A)))
    sim.init()
    sim.synchronize()
    t3 = time.perf_counter()
    sim.run(nsteps)
    t4 = time.perf_counter()
    sim.close()
    cpu = None
    if cpu_steps > 0:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pic_oracle as orc
        orc.build()
        st = run.initial_state()
        g = run.grid
        og = orc.grid(g.nx, g.ny, g.dx, g.dy)
        w = run.species[0].density * g.dx * g.dy / run.species[0].ppc
        threads = cpu_threads()
        c0 = time.perf_counter()
        orc.es_adjoint_f64(og, st, w, cpu_steps, nthreads=threads)
        cel = time.perf_counter() - c0
        cpu = {"value": run.n * cpu_steps / cel, "unit": "particle-steps/s (forward+reverse)",
               "cores": threads, "kind": "port",
               "sample": f"oracle FP64 host adjoint (forward + reverse sweep), OpenMP "
                         f"x{threads}, {cpu_steps} steps of the same 1,048,576-particle "
                         f"initial state", **cpu_info()}
    run.close()
    plain = t4 - t3
    n = run.n
    return {"workload": "two-stream 128^2, 64 ppc (1,048,576), 16 modes, 100 steps, "
                        "checkpoint every 10, loss EE(100)",
            "loss": loss, "forward_ms": 1e3 * (t1 - t0), "backward_ms": 1e3 * (t2 - t1),
            "plain_forward_ms": 1e3 * plain,
            "gradient_cost_vs_plain": (t2 - t0) / plain,
            "backward_vs_plain": (t2 - t1) / plain,
            "value": n * nsteps / (t2 - t0), "unit": "particle-steps/s (forward+reverse)",
            "cpu_baseline": cpu,
            "dL_dv0": float(gU[0, 0] - gU[1, 0]), "dL_dA1": float(gA[0])}


def yee_bench(steps: int = 300, warmup: int = 5):
    """The full-Maxwell (Yee) branch on the reference's chen2026_case1_df deck
    (tests/golden/decks): a 1D field-aligned line of 5000 cells, dipole
    background fitted from the deck, loss-cone delta-f hot electrons at 800
    ppc from the mirror loader, FP32 weights, cold-fluid background,
    reflecting walls with damping layers, 3-pass binomial current smoothing.
    Set up as tools/chirp2d.cu:44-99 does: the delta-f f0 from the species
    (temperatures, loss-cone shape, drift time), wd(0) uniform with rms
    wdnoise, and band-limited random-phase By, Bz of rms bnoise*wce over
    k in [0.2, 1.5] (numpy's generator here, so the noise samples differ from
    the reference's rand()).  Whole-step throughput over `steps` steps."""
    import numpy as np
    from arcwarden_amd.deck import load_deck
    from arcwarden_amd.maxwell import MaxwellSimulation
    d = load_deck(ROOT / "tests" / "golden" / "decks" / "chen2026_case1_df.ini")
    rp, sp = d.rp, d.species[0]
    rp.deltaf = 1
    rp.df_tpar = sp.uth[0] ** 2
    rp.df_tperp = sp.uth[1] ** 2
    rp.df_taud, rp.df_dist = sp.taud, sp.dist
    rp.df_rho, rp.df_kappa = sp.lc_rho, sp.lc_kappa
    g = d.grid()
    sim = MaxwellSimulation(g, rp)
    sim.load_mirror(sp)
    n = sim.n
    rng = np.random.default_rng(int(rp.rng_seed) + 7)
    a = sp.wdnoise * np.sqrt(3.0)
    sim.enable_deltaf(rp.df_wprec, wd0=rng.uniform(-a, a, n).astype(np.float32))
    lx = g.nx * g.dx
    m1 = max(1, int(np.ceil(0.2 * lx / (2 * np.pi))))
    m2 = min(g.nx // 2 - 1, int(1.5 * lx / (2 * np.pi)))
    amp = d.bnoise * rp.wce / np.sqrt(m2 - m1 + 1)
    i = np.arange(g.nx)
    nb, nz = np.zeros(g.nx), np.zeros(g.nx)
    for m in range(m1, m2 + 1):
        k = 2 * np.pi * m / lx
        p1, p2 = rng.uniform(0, 2 * np.pi, 2)
        nb += amp * np.cos(k * (i + 0.5) * g.dx + p1)
        nz += amp * np.cos(k * i * g.dx + p2)
    sim.set_mesh(by=nb, bz=nz)
    for _ in range(warmup):
        sim.step()
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    sim.synchronize()
    el = time.perf_counter() - t0
    we, wb = sim.field_energy()
    wsum, wrms, wmax = sim.wd_stats()
    sim.close()
    # per particle-step: the push reads x, y, ux, uy, uz, w, wd (28 B) and
    # writes ux, uy, uz, wd, the old and new x, y, the deposit velocity
    # uz/gamma (the walls damp u after it is taken) and weight (40 B); the
    # current reads the move, that velocity and weight (24 B) and writes the
    # wrapped x, y (8 B)
    bpp = 100
    out = {"workload": "Yee delta-f whistler line, chen2026_case1_df deck: 5000x1 cells, "
                       "dipole background, loss-cone hot electrons 800 ppc, walls, "
                       "jfilter 3, wd and By/Bz noise seeds", "particles": n, "steps": steps,
           "value": n * steps / el, "unit": "particle-steps/s", "ms_per_step": el / steps * 1e3,
           "step_bytes_per_particle": bpp,
           "step_GBs": n * bpp * steps / el / 1e9,
           "field_energy": [we, wb], "wd_rms": wrms}
    print(f"# yee: {n} particles, {out['ms_per_step']:.4f} ms/step", file=sys.stderr, flush=True)
    return out


def yee_eaw_bench(steps: int = 40, warmup: int = 5, tile_sort: bool = True):
    """The reference's published Yee benchmark, the eaw_case7_tiled deck
    (tests/golden/decks): 625x625 periodic cells, 400 ppc (156.25 M
    electrons, T_perp/T_par = 5), B0 = 0.25 along x, c = 1, dt 0.0145,
    3-pass binomial current smoothing, tile re-sort every 20 steps, no
    diagnostics inside the timed steps.  The reference quotes 1.51e10
    particle-steps/s for this run without diagnostics on an RTX 5090, and
    4.5e9 for the same run without the tile sort (eaw_case7.ini, the flat
    path) (docs/PROFILE_BASELINE.md:39-50).  tile_sort=False runs the flat
    configuration."""
    from arcwarden_amd.deck import load_deck
    from arcwarden_amd.maxwell import MaxwellSimulation
    d = load_deck(ROOT / "tests" / "golden" / "decks" / "eaw_case7_tiled.ini")
    if not tile_sort:
        d.rp.tile_sort = 0
    sim = MaxwellSimulation(d.grid(), d.rp)
    sim.load_species(d.species)
    n = sim.n
    for _ in range(warmup):
        sim.step()
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    sim.synchronize()
    el = time.perf_counter() - t0
    we, wb = sim.field_energy()
    sim.close()
    ref = 1.51e10 if tile_sort else 4.5e9
    # particle bytes per step: one kernel reads x, y, u, w (24 B) and writes
    # x, y, u (20 B); without the tile sort it also reads the bin permutation
    # (4 B), and the binning every 8 steps (x, y read twice, the permutation
    # written: 20 B) adds 2.5 B a step
    bpp = 44 if tile_sort else 50.5
    out = {"workload": "Yee eaw_case7 deck: 625x625 cells, 400 ppc, B0 0.25 x, jfilter 3, " +
                       ("tile_sort 20" if tile_sort else "no tile sort"),
           "particles": n, "steps": steps,
           "value": n * steps / el, "unit": "particle-steps/s", "ms_per_step": el / steps * 1e3,
           "reference_rtx5090": ref, "vs_reference": n * steps / el / ref,
           "step_bytes_per_particle": bpp, "step_GBs": n * bpp * steps / el / 1e9,
           "field_energy": [we, wb]}
    print(f"# yee eaw_case7 (tile_sort {int(tile_sort)}): {n} particles, "
          f"{out['ms_per_step']:.4f} ms/step", file=sys.stderr, flush=True)
    return out


def checkpoint_bench(sim_factory, sim, step: int):
    """Write the running simulation's restart file (ARCW v1: particles, the
    sources of the next step, the last fields) to a temporary directory, then
    restore it into a fresh simulation of the same configuration; bytes and
    wall time of each."""
    import tempfile
    d = tempfile.mkdtemp(prefix="aw_ck_")
    path = os.path.join(d, "run.arcw")
    try:
        sim.synchronize()
        t0 = time.perf_counter()
        sim.save_checkpoint(path, step)
        t1 = time.perf_counter()
        size = os.path.getsize(path)
        other = sim_factory()
        t2 = time.perf_counter()
        other.restore_checkpoint(path)
        other.synchronize()
        t3 = time.perf_counter()
        other.close()
    finally:
        if os.path.exists(path):
            os.remove(path)
        os.rmdir(d)
    return {"bytes": size, "write_s": t1 - t0, "read_s": t3 - t2,
            "write_GBs": size / (t1 - t0) / 1e9, "read_GBs": size / (t3 - t2) / 1e9,
            "where": "the box's local temporary directory"}


def darwin_bench(steps: int, current: int, autotune: bool = True, batch: bool = True):
    """The Darwin whistler-pump scan (v_r/v_th in {3.2, 2.1, 1.0}, 512x128,
    256 ppc = 16,777,216 electrons per run) stepped run after run by aw_sim
    (spectral Darwin model with the given current deposit, autotuned tile
    and workgroups per tile, re-sort on 2% strays, diagnostics every 100
    steps).  Loading is outside the timed
    region.  The EM push kernel is timed with HIP events on the sim stream
    and priced at 40 B per particle (x, y, ux, uy, uz read and written)."""
    from arcwarden_amd import _lib as L
    from arcwarden_amd.sim import SimConfig, Simulation
    from arcwarden_amd.workloads import whistler_run
    tot_ps, tot_s, kms, kl, n_last, sorts = 0.0, 0.0, 0.0, 0, 0, 0
    ratios = (3.2, 2.1, 1.0)
    ee = []
    for r in ratios:
        g, rp, sp = whistler_run(r)
        sim = Simulation(g, rp, sp, SimConfig(has_b0=True, field_model=L.AW_MODEL_DARWIN,
                                              sort_every=0, current_deposit=current,
                                              autotune=autotune))
        sim.init()
        n = sim.particle_views().n
        exec_cfg = sim.exec_config()
        sim.synchronize()
        sim.set_profiling(True)
        t0 = time.perf_counter()
        sim.run(steps)
        el = time.perf_counter() - t0
        ms, launches, ns = sim.kernel_time()
        kms += ms
        kl += launches
        sorts += ns
        hist = sim.history()
        ee.append(hist[-1]["ee"] if hist else None)
        sim.close()
        tot_ps += n * steps
        tot_s += el
        n_last = n
        print(f"# darwin current={current} v_r/v_th={r}: {el / steps * 1e3:.3f} ms/step",
              file=sys.stderr, flush=True)
    avg_k = kms / 1e3 / max(kl, 1)
    achieved = 40.0 * n_last / avg_k / 1e9
    # the same three runs as one batched Darwin ensemble (one launch per
    # particle kernel and per transform for all runs)
    from arcwarden_amd.ensemble import Ensemble, RunSpec
    out = {"workload": "Darwin whistler pump, 3 runs v_r/v_th 3.2/2.1/1.0, 512x128, dx 0.5, "
                       "256 ppc (16,777,216 each), c 10, w_ce 0.5 at 30 deg, pump w0 0.15 "
                       "M=18 dB=1e-3 B0, dt 0.05, run after run",
           "current_deposit": "esirkepov" if current == L.AW_CURRENT_ESIRKEPOV else "cic",
           "steps": steps, "value": tot_ps / tot_s, "unit": "particle-steps/s",
           "ms_per_step": tot_s / (len(ratios) * steps) * 1e3, "sorts": sorts,
           "em_push": {"avg_kernel_ms": avg_k * 1e3, "bytes_per_particle": 40.0,
                       "achieved_GBs": achieved, "frac": achieved / HBM_PEAK_GBS},
           "final_ee": ee, "exec": exec_cfg}
    if not batch:
        return out
    specs = []
    for r in ratios:
        g, rp, sp = whistler_run(r)
        specs.append(RunSpec(g, rp, sp, has_b0=True))
    ens = Ensemble(specs, SimConfig(has_b0=True, field_model=L.AW_MODEL_DARWIN, sort_every=0,
                                    current_deposit=current, autotune=autotune))
    ens.init()
    ens.synchronize()
    t0 = time.perf_counter()
    ens.run(0, steps)
    bat = time.perf_counter() - t0
    ens_sorts = ens.info()["sorts"]
    ens.close()
    print(f"# darwin current={current} batched: {bat / steps * 1e3:.3f} ms/step (3 runs)",
          file=sys.stderr, flush=True)
    out["batched"] = {"value": tot_ps / bat, "ms_per_step_all_runs": bat / steps * 1e3,
                      "speedup_vs_run_after_run": tot_s / bat, "sorts": ens_sorts}
    return out


def policy_bench(nx: int, ppc: int, reps: int = 5):
    """Deposit / push microbench on the headline state (the reference's
    deposit_bench): nx^2 grid, ppc electrons per cell loaded as in the main
    run, tile-sorted (8x8 tiles).  Times, with HIP events on one stream, the
    CIC charge deposit under each policy (global atomics, whole grid in LDS
    when the int64 grid fits the 160 KiB of LDS, tile-binned LDS), the push
    alone (Pusher::boris, no deposit), the fused push+deposit, and the
    Esirkepov current deposit of one step's moves under each policy.  Bytes
    are algorithmic: deposit 8 B/particle (x, y), push 40 B (x, y, ux, uy, uz
    read and written), fused 32 B, Esirkepov 20 B (x0, y0, x1, y1, uz read;
    w uniform)."""
    import ctypes as C
    import torch
    from arcwarden_amd import _lib as L
    from arcwarden_amd.sim import particles_struct
    dx = 0.5
    g = L.grid(nx, nx, nx * dx, nx * dx)
    rp = L.run_params(dt=0.1, qm=-1.0, eps0=1.0, noisy_load=1, rng_seed=20260628,
                      B0=(0.0, 0.0, 0.5))
    sp = L.species("electrons", 1.0, ppc, (0.05, 0.05, 0.05), (0.0, 0.0, 0.0))
    n = nx * nx * ppc
    ncell = nx * nx
    s = torch.cuda.current_stream()
    st = C.c_void_p(s.cuda_stream)
    h = C.c_void_p()
    L.call("aw_ctx_create", C.byref(g), 8, 8, C.byref(h))
    out = {}

    def soa():
        return [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(5)]

    def timed(fn, nbytes):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        best = float("inf")
        for _ in range(reps):
            e0.record(s)
            fn()
            e1.record(s)
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        return {"ms": best, "particles_per_s": n / (best * 1e-3),
                "GBs": nbytes * n / (best * 1e-3) / 1e9,
                "frac": nbytes * n / (best * 1e-3) / 1e9 / HBM_PEAK_GBS}

    try:
        nt, ntx, a, b = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        L.call("aw_ctx_tiles", h, C.byref(nt), C.byref(ntx), C.byref(a), C.byref(b))
        off = torch.empty(nt.value + 1, dtype=torch.int32, device="cuda")
        A, B = soa(), soa()
        wout = (C.c_double * 1)()
        L.call("aw_load_species", h, C.byref(sp), 1, C.byref(rp), None, 0,
               C.byref(particles_struct(*A)), wout, st)
        pa = particles_struct(*A, w_uniform=wout[0])
        pb = particles_struct(*B, w_uniform=wout[0])
        L.call("aw_sort_by_tile", h, C.byref(pa), C.byref(pb), C.c_void_p(off.data_ptr()), st)
        rho = torch.zeros(ncell, dtype=torch.float32, device="cuda")
        E = torch.randn(2 * ncell, device="cuda") * 1e-3
        Ex, Ey = C.c_void_p(E.data_ptr()), C.c_void_p(E.data_ptr() + 4 * ncell)
        J = [torch.zeros(ncell, dtype=torch.float32, device="cuda") for _ in range(3)]
        pols = {"global": L.AW_DEPOSIT_GLOBAL, "binned": L.AW_DEPOSIT_BINNED}
        if ncell * 8 <= 160 * 1024:
            pols["lds_grid"] = L.AW_DEPOSIT_LDS_GRID
        else:
            out["lds_grid"] = f"not applicable: the {ncell * 8 >> 10} KiB int64 grid exceeds LDS"
        # sorted state in B
        for name, pol in pols.items():
            def dep(pol=pol):
                rho.zero_()
                L.call("aw_deposit_rho", h, C.byref(pb), C.c_double(-1.0), pol,
                       C.c_void_p(rho.data_ptr()), st)
            out[f"deposit_{name}"] = timed(dep, 8.0)

        def fused():
            rho.zero_()
            L.call("aw_push_deposit_es", h, C.byref(pb), C.c_void_p(off.data_ptr()), Ex, Ey,
                   C.byref(rp), 1, C.c_void_p(rho.data_ptr()), None, st)
        out["push_deposit_fused"] = timed(fused, 32.0)
        out["push"] = timed(lambda: L.call("aw_push_es", h, C.byref(pb), Ex, Ey, C.byref(rp),
                                           1, st), 40.0)
        L.call("aw_migrate", h, C.byref(pb), None, st)
        # one more fused step from saved positions gives the Esirkepov moves
        x0, y0 = B[0].clone(), B[1].clone()
        L.call("aw_push_deposit_es", h, C.byref(pb), C.c_void_p(off.data_ptr()), Ex, Ey,
               C.byref(rp), 1, C.c_void_p(rho.data_ptr()), None, st)
        for name, pol in pols.items():
            def esk(pol=pol):
                for j in J:
                    j.zero_()
                L.call("aw_deposit_esirkepov", h, C.c_void_p(x0.data_ptr()),
                       C.c_void_p(y0.data_ptr()), C.byref(pb), C.c_double(-1.0),
                       C.c_double(0.1), pol, *(C.c_void_p(j.data_ptr()) for j in J), st)
            out[f"esirkepov_{name}"] = timed(esk, 20.0)
        torch.cuda.synchronize()
    finally:
        L.lib().aw_ctx_destroy(h)
    out["workload"] = (f"{nx}x{nx}, {ppc} ppc ({n} particles), tile-sorted 8x8 tiles, "
                       f"best of {reps} after one warm-up")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--ppc", type=int, default=256)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--sort-every", type=int, default=0,
                    help="re-sort cadence in steps; 0: re-sort only when more than 2%% of "
                         "the particles left their tile apron")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=500,
                    help="steps of the batched-ensemble comparison (0: skip)")
    ap.add_argument("--no-fixed-cadence", action="store_true",
                    help="skip the comparison run with a re-sort every 20 steps")
    ap.add_argument("--adjoint", type=int, default=1,
                    help="1: time the 128^2 two-stream adjoint (0: skip)")
    ap.add_argument("--darwin-steps", type=int, default=2000,
                    help="steps per run of the Darwin whistler-pump scan (0: skip)")
    ap.add_argument("--two-stream", type=int, default=1,
                    help="1: time the 64x64 two-stream run against the scalar host step")
    ap.add_argument("--darwin-autotune", type=int, default=1,
                    help="1: autotune the Darwin runs' tile / workgroups per tile")
    ap.add_argument("--no-darwin-batch", action="store_true",
                    help="skip the batched Darwin ensemble of the three runs")
    ap.add_argument("--no-policies", action="store_true",
                    help="skip the deposit/push policy microbench")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the default tile / workgroups-per-tile heuristic")
    ap.add_argument("--tile", type=int, default=0,
                    help="fixed square tile size (implies --no-autotune; 0: default)")
    ap.add_argument("--bpt", type=int, default=0,
                    help="fixed workgroups per tile (implies --no-autotune; 0: default)")
    ap.add_argument("--yee-steps", type=int, default=300,
                    help="steps of the Yee delta-f deck (0: skip)")
    ap.add_argument("--eaw-steps", type=int, default=40,
                    help="steps of the eaw_case7_tiled Yee deck (0: skip)")
    ap.add_argument("--checkpoint", type=int, default=1,
                    help="1: time a restart file write and restore of the headline run")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 1 GiB device copy (torch clone) so that a "
                         "rocprofv3 --pmc run can scale FETCH_SIZE/WRITE_SIZE by a "
                         "kernel of known traffic")
    args = ap.parse_args()
    if args.tile or args.bpt:
        args.no_autotune = True
    if args.calibrate_copy:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        for _ in range(3):
            dst = src.clone()
        torch.cuda.synchronize()
        del src, dst
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(ensembles would shard one run per GPU; a single run by particles, with an "
              "exact all-reduce of the int64 charge grid each step)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from arcwarden_amd import _lib as L
    from arcwarden_amd.sim import SimConfig, Simulation

    nx = ny = args.nx
    dx = 0.5
    g = L.grid(nx, ny, nx * dx, ny * dx)
    rp = L.run_params(dt=0.1, qm=-1.0, eps0=1.0, noisy_load=1, rng_seed=20260628,
                      B0=(0.0, 0.0, 0.5), dump_every=10)
    sp = [L.species("electrons", 1.0, args.ppc, (0.05, 0.05, 0.05), (0.0, 0.0, 0.0))]

    def make_sim(sort_every):
        return Simulation(g, rp, sp, SimConfig(has_b0=True, sort_every=sort_every,
                                               tile_x=args.tile, tile_y=args.tile,
                                               blocks_per_tile=args.bpt,
                                               autotune=not args.no_autotune))

    ck_out = {}

    def timed_run(sort_every, keep_state, checkpoint=False):
        sim = make_sim(sort_every)
        sim.init()
        n = sim.particle_views().n
        exec_cfg = sim.exec_config()
        print(f"# init done: {n} particles on {nx}x{ny}; exec {exec_cfg}", file=sys.stderr,
              flush=True)
        state = rho0 = None
        if keep_state:
            state = sim.particles()
            rho0 = sim.rho()
        for k in range(args.warmup):
            sim.step(k)
        sim.synchronize()
        sim.set_profiling(True)
        t0 = time.perf_counter()
        for k in range(args.warmup, args.warmup + args.steps):
            sim.step(k)
        sim.synchronize()
        t1 = time.perf_counter()
        kms, launches, sorts = sim.kernel_time()
        hist = sim.history()
        if checkpoint:  # outside the timed region
            ck_out.update(checkpoint_bench(lambda: make_sim(sort_every), sim,
                                           args.warmup + args.steps))
        sim.close()
        return dict(n=n, exec_cfg=exec_cfg, elapsed=t1 - t0, kms=kms, launches=launches,
                    sorts=sorts, hist=hist, state=state, rho0=rho0)

    main_run = timed_run(args.sort_every, not args.no_cpu, checkpoint=args.checkpoint)
    n, exec_cfg, elapsed = main_run["n"], main_run["exec_cfg"], main_run["elapsed"]
    kms, launches, sorts, hist = (main_run[k] for k in ("kms", "launches", "sorts", "hist"))
    state, rho0 = main_run["state"], main_run["rho0"]
    cpu = None
    value = n * args.steps / elapsed
    avg_kernel_s = kms / max(1, launches) / 1e3
    dumps = sum(1 for k in range(args.warmup, args.warmup + args.steps) if k % 10 == 0)
    bytes_per_launch = n * (BYTES_PER_PARTICLE + BYTES_UZ_READ * dumps / max(1, args.steps))
    achieved = bytes_per_launch / avg_kernel_s / 1e9
    print(f"# timed {args.steps} steps in {elapsed:.4f}s; push_deposit avg "
          f"{avg_kernel_s * 1e3:.4f} ms over {launches} launches; sorts={sorts}",
          file=sys.stderr, flush=True)
    fixed = None
    if args.sort_every != 20 and not args.no_fixed_cadence:
        r20 = timed_run(20, False)
        fixed = {"sort_every": 20, "value": n * args.steps / r20["elapsed"],
                 "ms_per_step": r20["elapsed"] / args.steps * 1e3, "sorts": r20["sorts"]}
        print(f"# sort every 20: {fixed['ms_per_step']:.4f} ms/step", file=sys.stderr,
              flush=True)
    ens = (ensemble_bench(args.ensemble_steps, cpu=not args.no_cpu)
           if args.ensemble_steps > 0 else None)
    adj = adjoint_bench(cpu_steps=0 if args.no_cpu else 10) if args.adjoint else None
    ts2 = two_stream_bench(cpu=not args.no_cpu) if args.two_stream else None
    dar = dar_cic = None
    if args.darwin_steps > 0:
        from arcwarden_amd import _lib as L
        kw = dict(autotune=bool(args.darwin_autotune), batch=not args.no_darwin_batch)
        dar = darwin_bench(args.darwin_steps, L.AW_CURRENT_ESIRKEPOV, **kw)
        dar_cic = darwin_bench(args.darwin_steps, L.AW_CURRENT_CIC, **kw)
    yee = yee_bench(args.yee_steps) if args.yee_steps > 0 else None
    yee_eaw = yee_eaw_bench(args.eaw_steps) if args.eaw_steps > 0 else None
    yee_eaw_flat = (yee_eaw_bench(max(1, args.eaw_steps // 2), tile_sort=False)
                    if args.eaw_steps > 0 else None)
    pol = None if args.no_policies else policy_bench(nx, args.ppc)
    if not args.no_cpu:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(state, nx, ny, dx, state["w_uniform"], rho0, args.cpu_seconds)
    charge_drift = abs(hist[-1]["charge"] - hist[0]["charge"]) / abs(hist[0]["charge"])
    out = {
        "metric": "particle-steps/s on 1 MI355X, %HBM roofline; batched & adjoint vs plain",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Maxwellian load, seed 20260628)",
        "config": {"workload": f"ES thermal plasma {nx}x{ny}, dx=0.5, CIC, {args.ppc} ppc "
                               f"({n} electrons), vth=0.05, B0=0.5z, dt=0.1; "
                               + (f"re-sort every {args.sort_every} steps"
                                  if args.sort_every > 0 else
                                  "re-sort when >2% of particles left their tile apron")
                               + ", diagnostics every 10",
                   "particles": n, "grid": [nx, ny], "ppc": args.ppc,
                   "tile": [exec_cfg["tile_x"], exec_cfg["tile_y"]],
                   "blocks_per_tile": exec_cfg["blocks_per_tile"],
                   "autotune": not args.no_autotune},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles(),
                     "kernel": "push_deposit_kernel",
                     "bytes_per_launch": bytes_per_launch,
                     "avg_kernel_ms": avg_kernel_s * 1e3},
        "cpu_baseline": cpu,
        "ensemble": ens,
        "adjoint": adj,
        "two_stream": ts2,
        "darwin": dar,
        "darwin_cic_current": dar_cic,
        "yee": yee,
        "yee_eaw_case7": yee_eaw,
        "yee_eaw_case7_flat": yee_eaw_flat,
        "checkpoint": ck_out or None,
        "policies": pol,
        "checks": {"charge_drift": charge_drift, "sorts": sorts,
                   "fixed_sort_cadence": fixed},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
