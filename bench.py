#!/usr/bin/env python3
"""Headline benchmark: electrostatic 2D3V PIC steps on one MI355X.

Workload (the measured configuration): ES thermal plasma, 512x512 periodic
grid, dx = 0.5, CIC, 256 electrons per cell (67,108,864), Maxwellian
v_th = 0.05 in 3 components, uniform B0 = 0.5 z (magnetised Boris), FP32 SoA
particles, dt = 0.1.  A step is one Simulation::step: rho^n -> FFT Poisson ->
fused gather+push+move+wrap+deposit, with the stable tile re-sort every 20
steps and the scalar diagnostics every 10 steps, all inside the timed region.

Prints ONE JSON line.  ``value`` is particle-steps/s over the whole timed
region; ``roofline`` prices the dominant kernel (push_deposit) at its
algorithmic particle traffic against 8 TB/s: x, y, ux, uy read and written
(32 B per particle; with B0 along z, uz is a constant of the motion), plus
4 B to read uz on the diagnostics steps, which sum the kinetic energy;
``cpu_baseline`` times the oracle's FP32 same-algorithm ES step with OpenMP
on this host's cores on the same particle state.
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


def cpu_baseline(state, g_nx, g_ny, dx, w_uniform, rho, budget_s):
    import numpy as np
    sys.path.insert(0, str(ROOT / "oracle"))
    import pic_oracle as orc
    orc.build()
    threads = min(16, os.cpu_count() or 1)
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
    return {"value": n * steps / el, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle FP32 ES step (solve+push+wrap+deposit), OpenMP x{threads}, "
                      f"{steps} step(s) of the full {g_nx}x{g_ny}x256ppc particle set "
                      f"({n} particles) taken from the GPU state after init"}


def ensemble_bench(steps: int, graph: bool = True):
    """The 24-run batched ensemble (two-beam plasmas, seed 20260629: side in
    {64..256}, ppc in {16, 64, 256}, vth log-uniform [0.005, 0.05], v0 in
    [0.1, 0.3], dx 0.5, dt 0.1) stepped as one aw_ensemble versus the same
    runs stepped one after another by aw_sim.  Loading is outside both timed
    regions; each timed region ends with a device synchronisation."""
    from arcwarden_amd.ensemble import Ensemble, random_ensemble
    from arcwarden_amd.sim import SimConfig, Simulation
    runs = random_ensemble(dump_every=100)
    ntot = sum(r.grid.nx * r.grid.ny * sum(s.ppc for s in r.species) for r in runs)
    seq = 0.0
    for r in runs:
        sim = Simulation(r.grid, r.rp, r.species, SimConfig(has_b0=False, sort_every=20))
        sim.init()
        sim.synchronize()
        t0 = time.perf_counter()
        sim.run(steps)
        seq += time.perf_counter() - t0
        sim.close()
    print(f"# ensemble: sequential {seq:.3f}s for {steps} steps", file=sys.stderr, flush=True)
    ens = Ensemble(runs, SimConfig(sort_every=20), graph=graph)
    ens.init()
    ens.synchronize()
    t0 = time.perf_counter()
    ens.run(0, steps)
    bat = time.perf_counter() - t0
    info = ens.info()
    ens.close()
    print(f"# ensemble: batched {bat:.3f}s (graph={graph})", file=sys.stderr, flush=True)
    return {"runs": len(runs), "particles": ntot, "steps": steps,
            "value": ntot * steps / bat, "unit": "particle-steps/s",
            "sequential_value": ntot * steps / seq, "speedup_vs_sequential": seq / bat,
            "graph": bool(graph), "size_groups": info["ngroups"],
            "workload": "24 two-beam ES runs, seed 20260629, sides 64-256, ppc 16/64/256, "
                        "vth log-U[0.005,0.05], v0 U[0.1,0.3], dx 0.5, dt 0.1"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--ppc", type=int, default=256)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--sort-every", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=500,
                    help="steps of the batched-ensemble comparison (0: skip)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the default tile / workgroups-per-tile heuristic")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 1 GiB device copy (torch clone) so that a "
                         "rocprofv3 --pmc run can scale FETCH_SIZE/WRITE_SIZE by a "
                         "kernel of known traffic")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        for _ in range(3):
            dst = src.clone()
        torch.cuda.synchronize()
        del src, dst
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the ensemble and particle-decomposition sharding is described in DESIGN.md)",
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
    sim = Simulation(g, rp, sp, SimConfig(has_b0=True, sort_every=args.sort_every,
                                          autotune=not args.no_autotune))
    sim.init()
    n = sim.particle_views().n
    exec_cfg = sim.exec_config()
    print(f"# init done: {n} particles on {nx}x{ny}; exec {exec_cfg}", file=sys.stderr,
          flush=True)

    cpu = None
    if not args.no_cpu:
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
    sim.close()
    elapsed = t1 - t0
    value = n * args.steps / elapsed
    avg_kernel_s = kms / max(1, launches) / 1e3
    dumps = sum(1 for k in range(args.warmup, args.warmup + args.steps) if k % 10 == 0)
    bytes_per_launch = n * (BYTES_PER_PARTICLE + BYTES_UZ_READ * dumps / max(1, args.steps))
    achieved = bytes_per_launch / avg_kernel_s / 1e9
    print(f"# timed {args.steps} steps in {elapsed:.4f}s; push_deposit avg "
          f"{avg_kernel_s * 1e3:.4f} ms over {launches} launches; sorts={sorts}",
          file=sys.stderr, flush=True)
    ens = ensemble_bench(args.ensemble_steps) if args.ensemble_steps > 0 else None
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
                               f"({n} electrons), vth=0.05, B0=0.5z, dt=0.1; sort every "
                               f"{args.sort_every} steps, diagnostics every 10",
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
        "checks": {"charge_drift": charge_drift, "sorts": sorts},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
