"""GPU tests of the full-Maxwell (Yee) branch (aw_maxwell_*):

- one step against the oracle: the mesh updates and the push bit for bit
  (FP32, same operation sequence), the current of the move within 1e-6 of
  the FP64 Esirkepov restatement, across B0 orientations, the relativistic
  push, delta-f weights (three precisions), the mirror background on a
  one-row grid, reflecting walls and the pump, and the current smoothing in
  LDS and in global passes;
- the tile_sort path (particles kept in tile order, the current deposited
  from the tile slices of the last sort) bit for bit against the unsorted
  run, full-f and delta-f;
- the conservation diagnostics against the oracle, and the reference's
  physics gates: vacuum dispersion (tests/test_yee_vacuum.cu, 3e-4), cold
  Langmuir on the Maxwell path (tests/test_yee_langmuir.cu, 2%), the
  parallel whistler on Yee and Darwin (tests/test_yee_vs_darwin_lowfreq.cu,
  2% each and across models, div B and Gauss drift below 1e-5), the delta-f
  equilibrium hold (tests/test_deltaf_equilibrium.cu) and the mirror load's
  marker counts."""
import math

import numpy as np
import pytest

from arcwarden_amd import _lib as L
from arcwarden_amd.maxwell import MaxwellSimulation
from arcwarden_amd.sim import SimConfig, Simulation

from gpu_util import rel_l2

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi

# RunParams member -> the oracle's Yee parameter block
_YEE_KEYS = ("qm", "dt", "c", "rel", "pump", "deltaf", "df_wprec", "df_dist", "b0_prof",
             "bnd_x", "bnd_nd", "pump_ex", "pump_ey", "pump_ez", "pump_k0", "pump_w0",
             "pump_trmp", "pump_toff", "df_tpar", "df_tperp", "df_kappa", "df_rho", "df_taud",
             "b0_a", "b0_xc", "b0_sref", "bnd_numax")


def _oracle_params(oracle, rp):
    kw = {k: getattr(rp, k) for k in _YEE_KEYS}
    kw["B0"] = tuple(rp.B0)
    kw["b0_c"] = tuple(rp.b0_c)
    return oracle.yee_params(**kw)


CASES = {
    # name: (nx, ny, dx, dy, run-parameter overrides)
    "oblique_b0": (40, 24, 0.5, 0.4, dict(B0=(0.3, 0.1, 0.25))),
    "relativistic": (32, 16, 0.5, 0.5, dict(B0=(0.0, 0.0, 0.4), rel=1)),
    "deltaf": (32, 16, 0.5, 0.5, dict(B0=(0.5, 0.0, 0.0), deltaf=1, df_tpar=0.01,
                                      df_tperp=0.02)),
    "deltaf_kahan_losscone": (32, 16, 0.5, 0.5, dict(B0=(0.5, 0.0, 0.0), deltaf=1,
                                                     df_tpar=0.01, df_tperp=0.02, df_wprec=1,
                                                     df_dist=1, df_rho=0.6, df_kappa=0.3,
                                                     df_taud=40.0)),
    "deltaf_fp64": (32, 16, 0.5, 0.5, dict(B0=(0.5, 0.0, 0.0), deltaf=1, df_tpar=0.01,
                                           df_tperp=0.02, df_wprec=2)),
    "mirror_one_row": (64, 1, 0.5, 0.5, dict(B0=(0.5, 0.0, 0.0), b0_prof=1, b0_a=2e-3,
                                             b0_xc=16.0)),
    "walls_smoothing": (48, 8, 0.5, 0.5, dict(B0=(0.2, 0.0, 0.3), bnd_x=2, bnd_nd=6,
                                              bnd_numax=0.5, jfilter=2)),
    # grids above the in-LDS smoothing size: the multi-launch passes
    "smoothing_3_passes_large": (128, 80, 0.5, 0.5, dict(B0=(0.0, 0.3, 0.2), jfilter=3)),
    "smoothing_1_pass_large": (128, 80, 0.5, 0.5, dict(B0=(0.0, 0.3, 0.2), jfilter=1)),
    "pump": (48, 12, 0.5, 0.5, dict(B0=(0.4, 0.0, 0.2), pump=1, pump_ex=1e-3, pump_ey=2e-3,
                                    pump_ez=-2e-3, pump_k0=TWO_PI * 3 / 24.0, pump_w0=0.15,
                                    pump_trmp=0.5, pump_toff=100.0)),
}


def _setup(case, seed=3):
    nx, ny, dx, dy, over = CASES[case]
    g = L.grid(nx, ny, nx * dx, ny * dy)
    rp = L.run_params(dt=0.04, qm=-1.0, eps0=1.0, c=5.0)
    for k, v in over.items():
        if k == "B0":
            for i in range(3):
                rp.B0[i] = v[i]
        else:
            setattr(rp, k, v)
    rng = np.random.default_rng(seed)
    n = 6000
    st = {"x": rng.uniform(0, nx, n).astype(np.float32),
          "y": rng.uniform(0, ny, n).astype(np.float32)}
    for k in ("ux", "uy", "uz"):
        st[k] = rng.normal(0, 0.3, n).astype(np.float32)
    if over.get("bnd_x"):  # particles near both walls
        st["x"][:50] = rng.uniform(0, 0.2, 50).astype(np.float32)
        st["ux"][:50] = -np.abs(st["ux"][:50]) - 0.5
        st["x"][50:100] = rng.uniform(nx - 0.2, nx, 50).astype(np.float32)
        st["ux"][50:100] = np.abs(st["ux"][50:100]) + 0.5
    w = rng.uniform(0.5, 1.5, n).astype(np.float32) * (dx * dy / 4)
    nc = nx * ny
    F0 = rng.normal(0, 1e-2, 9 * nc).astype(np.float32)
    F0[6 * nc:] = 0.0
    return g, rp, st, w, F0


@pytest.mark.parametrize("case", sorted(CASES))
def test_yee_step_matches_oracle(gpu, oracle, case):
    g, rp, st, w, F0 = _setup(case)
    nx, ny = g.nx, g.ny
    nc = nx * ny
    sim = MaxwellSimulation(g, rp)
    sim.set_particles(st, w=w)
    wd0 = None
    if rp.deltaf:
        wd0 = np.random.default_rng(5).uniform(-0.05, 0.05, st["x"].size).astype(np.float32)
        sim.enable_deltaf(rp.df_wprec, wd0=wd0)
    names = ("ex", "ey", "ez", "bx", "by", "bz")
    sim.set_mesh(**{k: F0[q * nc:(q + 1) * nc] for q, k in enumerate(names)})
    t0 = 0.7
    sim.step(t0)
    got = sim.particles()
    M = sim.mesh_flat()
    sim.close()
    og = oracle.grid(nx, ny, g.dx, g.dy)
    yp = _oracle_params(oracle, rp)
    F1 = oracle.yee_faraday(og, F0, 0.5 * rp.dt)
    ref, x0, y0, vz, wdep = oracle.yee_push(
        og, yp, t0, F1, st, w=w, wd=wd0, wc=np.zeros_like(wd0) if wd0 is not None else None,
        wdd=wd0.astype(np.float64) if wd0 is not None else None)
    # the pump's double cos/sin, the damping layers' expf and the loss-cone
    # expf are the device's libm, which may differ from the host's in the
    # last place: those cases compare within FP32 rounding, the rest bit for bit
    exact = not (rp.pump or rp.bnd_x == 2 or (rp.deltaf and rp.df_dist == 1))
    for k in ("ux", "uy", "uz"):
        if exact:
            assert np.array_equal(got[k], ref[k]), (case, k)
        else:
            assert rel_l2(got[k], ref[k]) < 1e-7, (case, k)
    rx, ry = oracle.migrate(og, ref["x"], ref["y"])[:2]
    if exact:
        assert np.array_equal(got["x"], rx) and np.array_equal(got["y"], ry), case
    else:
        assert rel_l2(got["x"], rx) < 1e-7 and rel_l2(got["y"], ry) < 1e-7, case
    if rp.deltaf:
        if exact:
            assert np.array_equal(got["wd"], ref["wd"]), case
        else:
            assert rel_l2(got["wd"], ref["wd"]) < 1e-6, case
    # the current of the move, then smoothing, against the FP64 restatement
    wdep_or_w = wdep if rp.deltaf else w
    # Jz takes uz / gamma after the kick, before the walls damp u
    jz_vel = vz if (rp.rel or rp.bnd_x) else ref["uz"]
    j64 = oracle.esirkepov_f64(og, x0, y0, ref["x"], ref["y"], jz_vel, rp.qm, rp.dt,
                               w=wdep_or_w)
    for q, jref in enumerate(j64):
        jr = jref.astype(np.float32).ravel()
        for _ in range(rp.jfilter):
            jr = oracle.yee_binomial(og, jr)
        jg = M[(6 + q) * nc:(7 + q) * nc]
        assert rel_l2(jg, jr) < 1e-6, (case, q, rel_l2(jg, jr))
    # given the GPU's J, the second half of B and the E update are bit-exact
    F2 = F1.copy()
    F2[6 * nc:] = M[6 * nc:]
    F2 = oracle.yee_faraday(og, F2, 0.5 * rp.dt)
    F2 = oracle.yee_ampere(og, F2, rp.c * rp.c, rp.dt)
    got6 = M[:6 * nc].reshape(6, ny, nx)
    ref6 = F2[:6 * nc].reshape(6, ny, nx)
    if rp.bnd_x:  # the damping layers scale the columns within bnd_nd of the walls
        nd = rp.bnd_nd
        got6, ref6 = got6[:, :, nd:nx - nd], ref6[:, :, nd:nx - nd]
    assert np.array_equal(got6, ref6), case


def test_yee_vacuum_step_and_diagnostics(gpu, oracle):
    """No particles: one step is the oracle's Faraday / Faraday / Ampere bit for
    bit; field energy and div B match the FP64 oracle."""
    rng = np.random.default_rng(9)
    nx, ny, dx, dy = 36, 20, 0.5, 0.3
    g = L.grid(nx, ny, nx * dx, ny * dy)
    rp = L.run_params(dt=0.02, qm=-1.0, c=3.0)
    nc = nx * ny
    F0 = rng.normal(0, 1, 9 * nc).astype(np.float32)
    F0[6 * nc:] = 0.0
    sim = MaxwellSimulation(g, rp)
    sim.set_particles({k: np.zeros(0, np.float32) for k in ("x", "y", "ux", "uy", "uz")})
    names = ("ex", "ey", "ez", "bx", "by", "bz")
    sim.set_mesh(**{k: F0[q * nc:(q + 1) * nc] for q, k in enumerate(names)})
    for _ in range(3):
        sim.step()
    M = sim.mesh_flat()
    we, wb = sim.field_energy()
    divb, _ = sim.residuals()
    sim.close()
    og = oracle.grid(nx, ny, dx, dy)
    F = F0
    for _ in range(3):
        F = oracle.yee_faraday(og, F, 0.5 * rp.dt)
        F = oracle.yee_faraday(og, F, 0.5 * rp.dt)
        F = oracle.yee_ampere(og, F, rp.c ** 2, rp.dt)
    assert np.array_equal(M[:6 * nc], F[:6 * nc])
    e64 = oracle.yee_energy(og, F, rp.c ** 2)
    assert we == pytest.approx(e64[0], rel=1e-12) and wb == pytest.approx(e64[1], rel=1e-12)
    assert divb == pytest.approx(oracle.yee_divb_rms(og, F), rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("mx,my", [(2, 0), (1, 2)])
def test_yee_vacuum_dispersion(gpu, mx, my):
    """tests/test_yee_vacuum.cu: a standing Ez mode oscillates at the Yee
    dispersion sin^2(w dt/2)/(c dt)^2 = sin^2(kx dx/2)/dx^2 + sin^2(ky dy/2)/dy^2
    within 3e-4 over about 40 half periods."""
    nx, ny, dx, dy, c = 32, 16, 1.0, 1.0, 1.0
    dt = 0.5 / (c * math.sqrt(1 / dx ** 2 + 1 / dy ** 2))
    g = L.grid(nx, ny, nx * dx, ny * dy)
    rp = L.run_params(dt=dt, qm=-1.0, c=c)
    kx, ky = TWO_PI * mx / (nx * dx), TWO_PI * my / (ny * dy)
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dy)
    mode = np.cos(kx * X + ky * Y)
    s = math.sqrt(math.sin(kx * dx / 2) ** 2 / dx ** 2 + math.sin(ky * dy / 2) ** 2 / dy ** 2)
    w_yee = 2.0 / dt * math.asin(c * dt * s)
    sim = MaxwellSimulation(g, rp)
    sim.set_particles({k: np.zeros(0, np.float32) for k in ("x", "y", "ux", "uy", "uz")})
    sim.set_mesh(ez=mode.astype(np.float32))
    nsteps = int(20 * TWO_PI / w_yee / dt)
    prev, cross = None, []
    for n in range(1, nsteps + 1):
        sim.step()
        a = float(np.sum(sim.mesh()["ez"].astype(np.float64) * mode))
        if prev is not None and np.sign(a) != np.sign(prev):
            cross.append((n - 1) * dt + dt * prev / (prev - a))
        prev = a
    sim.close()
    w = math.pi * (len(cross) - 1) / (cross[-1] - cross[0])
    assert abs(w - w_yee) / w_yee < 3e-4, (w, w_yee)


def test_yee_langmuir(gpu):
    """tests/test_yee_langmuir.cu: cold electrons on a quiet 8 x 8 sub-lattice
    per cell, ux = v0 sin(k x), E = B = 0: the charge-conserving current
    builds Ex, and the mode oscillates at w_pe = 1 within 2%."""
    nx, ny, side, dx, dy, c = 64, 8, 8, 1.0, 1.0, 10.0
    g = L.grid(nx, ny, nx * dx, ny * dy)
    rp = L.run_params(qm=-1.0, c=c, dt=0.4 / (c * math.sqrt(2.0)))
    kx, v0 = TWO_PI / (nx * dx), 0.01
    sub = (np.arange(side) + 0.5) / side
    i, j, a, b = np.meshgrid(np.arange(nx), np.arange(ny), sub, sub, indexing="ij")
    x = (i + a).ravel().astype(np.float32)
    y = (j + b).ravel().astype(np.float32)
    n = x.size
    st = {"x": x, "y": y, "ux": (v0 * np.sin(kx * x * dx)).astype(np.float32),
          "uy": np.zeros(n, np.float32), "uz": np.zeros(n, np.float32)}
    sim = MaxwellSimulation(g, rp)
    sim.set_particles(st, w=np.full(n, dx * dy / side ** 2, np.float32))
    nsteps = int(10 * TWO_PI / rp.dt)
    prof = np.sin(kx * (np.arange(nx) + 0.5) * dx)
    prev, cross = None, []
    for k in range(1, nsteps + 1):
        sim.step()
        A = float(np.sum(sim.mesh()["ex"].astype(np.float64) * prof[None, :]))
        if prev is not None and ((prev < 0 <= A) or (prev > 0 >= A)):
            cross.append((k - 1 + prev / (prev - A)) * rp.dt)
        prev = A
    sim.close()
    assert len(cross) >= 5
    w = math.pi * (len(cross) - 1) / (cross[-1] - cross[0])
    assert abs(w - 1.0) < 0.02, w


def _whistler_grid():
    nx, ny, mode, k = 128, 1, 2, 0.3
    Lx = TWO_PI * mode / k
    return L.grid(nx, ny, Lx, Lx / nx), k


def _perturb(st, k, dx, v1):
    th = k * st["x"].astype(np.float64) * dx
    st["uy"] = (st["uy"] + v1 * np.cos(th)).astype(np.float32)
    st["uz"] = (st["uz"] - v1 * np.sin(th)).astype(np.float32)
    return st


def _project(st, k, dx):
    th = k * st["x"].astype(np.float64) * dx
    return np.sum((st["uy"].astype(np.float64) + 1j * st["uz"].astype(np.float64))
                  * np.exp(1j * th))


def _fit(C, dt_s):
    ph = np.unwrap(np.angle(np.asarray(C)))
    return abs(np.polyfit(np.arange(len(C)) * dt_s, ph, 1)[0])


def test_yee_vs_darwin_whistler(gpu):
    """tests/test_yee_vs_darwin_lowfreq.cu on this build: the cold parallel
    whistler at k c = 3, w_ce = 0.5 (one-row grid, 64 quiet electrons per
    cell, v1 = 0.05).  Darwin rotates at w_ce k^2c^2/(1 + k^2c^2) and Yee at
    the root of c^2k^2 = w^2 + w/(w_ce - w), each within 2%; the two agree
    within 2%; the Yee run keeps div B and the Gauss drift below 1e-5."""
    c, wce, uth, v1 = 10.0, 0.5, 0.01, 0.05
    g, k = _whistler_grid()
    dx = g.dx
    kc2 = (k * c) ** 2
    w_dar = wce * kc2 / (1 + kc2)
    w_max = w_dar
    for _ in range(50):
        f = kc2 - w_max ** 2 - w_max / (wce - w_max)
        df = -2 * w_max - wce / (wce - w_max) ** 2
        w_max -= f / df
    sp = [L.species("e", 1.0, 64, (uth,) * 3, (0.0, 0.0, 0.0))]
    tend = 70.0
    # Darwin branch
    rp = L.run_params(dt=0.05, qm=-1.0, eps0=1.0, c=c, n0=1.0, noisy_load=0,
                      B0=(wce, 0.0, 0.0), wce=wce, dump_every=10 ** 6)
    sim = Simulation(g, rp, sp, SimConfig(has_b0=True, field_model=L.AW_MODEL_DARWIN))
    sim.init()
    sim.set_particles(_perturb(sim.particles(), k, dx, v1))
    Cd = []
    for n in range(int(round(tend / rp.dt))):
        sim.step(n)
        if n % 2 == 0:
            Cd.append(_project(sim.particles(), k, dx))
    sim.close()
    w_d = _fit(Cd, 2 * rp.dt)
    # Yee branch
    rp = L.run_params(dt=0.02, qm=-1.0, eps0=1.0, c=c, noisy_load=0, B0=(wce, 0.0, 0.0),
                      wce=wce)
    ysim = MaxwellSimulation(g, rp)
    ysim.load_species(sp)
    ysim.set_particles(_perturb(ysim.particles(), k, dx, v1), w=ysim.particles()["w"])
    divb0, _ = ysim.residuals()
    Cy = []
    for n in range(int(round(tend / rp.dt))):
        ysim.step()
        if n % 5 == 0:
            Cy.append(_project(ysim.particles(), k, dx))
    divb1, gauss1 = ysim.residuals()
    ysim.close()
    w_y = _fit(Cy, 5 * rp.dt)
    print("whistler darwin", w_d, w_dar, "yee", w_y, w_max, "divb", divb0, divb1,
          "gauss", gauss1)
    assert abs(w_d - w_dar) / w_dar < 0.02
    assert abs(w_y - w_max) / w_max < 0.02
    assert abs(w_y - w_d) / w_d < 0.02
    assert divb1 < 1e-5 and gauss1 < 1e-5


def test_deltaf_equilibrium_hold(gpu):
    """tests/test_deltaf_equilibrium.cu: a quiet isotropic Maxwellian with
    delta-f weights, B0 = w_ce x, a 1e-6 transverse B seed: the weights stay
    bounded (max |wd| < 1e-2) with no secular growth (rms grows less than 3x
    from the middle to the end)."""
    nx, ny, ppc, c, wce, uth, dx = 64, 16, 100, 10.0, 0.5, 0.05, 0.1
    g = L.grid(nx, ny, nx * dx, ny * dx)
    rp = L.run_params(dt=0.4 * dx / c, c=c, qm=-1.0, eps0=1.0, B0=(wce, 0.0, 0.0), wce=wce,
                      noisy_load=0, deltaf=1, df_tpar=uth * uth, df_tperp=uth * uth)
    sim = MaxwellSimulation(g, rp)
    sim.load_species([L.species("e", 1.0, ppc, (uth,) * 3, (0, 0, 0))])
    sim.enable_deltaf(0)
    rng = np.random.default_rng(7)
    sim.set_mesh(by=rng.uniform(-1e-6, 1e-6, (ny, nx)), bz=rng.uniform(-1e-6, 1e-6, (ny, nx)))
    nsteps = 4000
    for n in range(1, nsteps + 1):
        sim.step()
        if n == nsteps // 2:
            _, rms_half, mx_half = sim.wd_stats()
    _, rms_end, mx_end = sim.wd_stats()
    sim.close()
    print("deltaf hold", rms_half, rms_end, mx_half, mx_end)
    assert mx_end < 1e-2
    assert rms_end / rms_half < 3.0


def test_mirror_load_counts_and_moments(gpu, oracle):
    """Particles::initialize_mirror: per-cell marker counts equal the oracle's
    offsets exactly, positions stay in their cells, and the velocity moments
    follow the mapped bi-Maxwellian (<u_par^2> = T_par, <u_perp^2> = 2 T_perp(x))."""
    nx, dx = 64, 0.5
    g = L.grid(nx, 1, nx * dx, dx)
    rp = L.run_params(dt=0.05, qm=-1.0, c=10.0, B0=(0.5, 0.0, 0.0), b0_prof=1, b0_a=4e-3,
                      b0_xc=nx * dx / 2, rng_seed=11)
    q = L.species("e", 1.0, 400, (0.1, 0.2, 0.2), (0, 0, 0))
    sim = MaxwellSimulation(g, rp)
    sim.load_mirror(q)
    st = sim.particles()
    sim.close()
    yp = _oracle_params(oracle, rp)
    off = oracle.mirror_offsets(oracle.grid(nx, 1, dx, dx), yp, 400, 0.1, 0.2)
    assert st["x"].size == off[-1]
    cells = np.floor(st["x"]).astype(np.int64)
    assert (np.bincount(cells, minlength=nx) == np.diff(off)).all()
    assert (st["y"] == 0.5).all() and np.allclose(st["w"], dx * dx / 400)
    assert np.mean(st["ux"].astype(np.float64) ** 2) == pytest.approx(0.01, rel=0.03)
    xc = (np.arange(nx) + 0.5) * dx
    b = 1 + 4e-3 * (xc - nx * dx / 2) ** 2
    tp = 1 / ((1 - 1 / b) / 0.01 + (1 / b) / 0.04)
    up2 = st["uy"].astype(np.float64) ** 2 + st["uz"].astype(np.float64) ** 2
    per_cell = np.bincount(cells, weights=up2, minlength=nx) / np.diff(off)
    assert np.allclose(per_cell, 2 * tp, rtol=0.15)


def _tile_sort_run(tile_sort, deltaf, steps=15, seed=11, nx=64, ny=48):
    dx = 0.5
    g = L.grid(nx, ny, nx * dx, ny * dx)
    rp = L.run_params(dt=0.06, qm=-1.0, eps0=1.0, c=5.0)
    rp.B0[0], rp.B0[2] = 0.4, 0.1
    rp.jfilter, rp.tile_sort = 3, tile_sort
    if deltaf:
        rp.B0[2] = 0.0
        rp.deltaf, rp.df_tpar, rp.df_tperp = 1, 1.5, 2.0
    rng = np.random.default_rng(seed)
    n = {48: 200_000, 1: 60_000}.get(ny, 10 * nx * ny)
    st = {"x": rng.uniform(0, nx, n).astype(np.float32),
          "y": rng.uniform(0, ny, n).astype(np.float32)}
    for k in ("ux", "uy", "uz"):  # up to ~0.8 cells a step, ~2 cells between sorts
        st[k] = rng.normal(0, 1.5, n).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32) * (dx * dx / 64)
    sim = MaxwellSimulation(g, rp)
    sim.set_particles(st, w=w)
    if deltaf:
        sim.enable_deltaf(0, wd0=rng.uniform(-0.05, 0.05, n).astype(np.float32))
    nc = nx * ny
    F0 = rng.normal(0, 1e-2, 6 * nc).astype(np.float32)
    names = ("ex", "ey", "ez", "bx", "by", "bz")
    sim.set_mesh(**{k: F0[q * nc:(q + 1) * nc] for q, k in enumerate(names)})
    for _ in range(steps):
        sim.step()
    M = sim.mesh_flat()
    P = sim.particles()
    sim.close()
    return M, P


def _as_set(P, keys):
    cols = np.stack([P[k] for k in keys])
    order = np.lexsort(cols[::-1])
    return cols[:, order]


@pytest.mark.parametrize("deltaf,nx,ny", [(False, 64, 48), (True, 64, 48), (False, 512, 1),
                                          (True, 512, 1), (False, 256, 160), (True, 256, 160)])
def test_tile_sort_matches_unsorted(gpu, deltaf, nx, ny):
    """MaxwellSimulation::step_at with rp.tile_sort (include/pic/simulation_maxwell.hpp:
    60-72): re-sorting the particles every 5 steps only reorders them -- the
    fields after 15 steps are the same bits, and so is the particle set.  The
    tile-sorted run pushes and deposits in one kernel per tile (the plain
    instance without delta-f, the general one with it; one-row grids sum
    their aliased stencil rows); the unsorted one in a push and a deposit
    (64 x 48, 512 x 1: the current's whole grid fits LDS), or with the same
    kernel over tile bins of the unmoved arrays (256 x 160, 10 particles per
    cell)."""
    M0, P0 = _tile_sort_run(0, deltaf, nx=nx, ny=ny)
    M1, P1 = _tile_sort_run(5, deltaf, nx=nx, ny=ny)
    nc = nx * ny
    print(f"tile_sort deltaf={deltaf} {nx}x{ny}: max |J| {np.abs(M0[-3 * nc:]).max():.3e}",
          flush=True)
    assert np.abs(M0[-3 * nc:]).max() > 0
    assert np.array_equal(M0, M1)
    keys = ("x", "y", "ux", "uy", "uz", "w") + (("wd",) if deltaf else ())
    assert np.array_equal(_as_set(P0, keys), _as_set(P1, keys))
    # the sorted run keeps the particles near tile order (the unsorted one
    # steps back to a lower tile at about every other particle)
    ntx = (nx + 15) // 16
    t = (P1["y"] // 16).astype(np.int64) * ntx + (P1["x"] // 16).astype(np.int64)
    assert np.mean(np.diff(t) < 0) < 0.25


def test_tile_sort_refuses_weight_precision_accumulators(gpu):
    g = L.grid(32, 32, 16.0, 16.0)
    rp = L.run_params(dt=0.04, qm=-1.0, eps0=1.0, c=5.0)
    rp.B0[0] = 0.5
    rp.deltaf, rp.df_tpar, rp.df_tperp, rp.df_wprec, rp.tile_sort = 1, 0.01, 0.02, 1, 4
    rng = np.random.default_rng(2)
    n = 1000
    st = {"x": rng.uniform(0, 32, n), "y": rng.uniform(0, 32, n),
          "ux": rng.normal(0, 0.1, n), "uy": rng.normal(0, 0.1, n), "uz": rng.normal(0, 0.1, n)}
    sim = MaxwellSimulation(g, rp)
    sim.set_particles(st, w=np.full(n, 0.25, np.float32))
    sim.enable_deltaf(1)
    with pytest.raises(Exception, match="tile_sort"):
        sim.step()
    sim.close()


@pytest.mark.parametrize("nx", [256, 20000])
def test_antenna_cold_fluid_walls_match_oracle(gpu, oracle, nx):
    """The field side of MaxwellSimulation::step_at without particles
    (include/pic/simulation_maxwell.hpp:52-109): B half step, the antenna
    column (yee2d.hpp:582-607), the cold fluid with the profile's B0(x)
    (:536-560), B half step, E step, the damping layers (:514-526), three
    steps against the oracle.  The antenna's exp/cos/sin, the fluid's sincosf
    and the masks' exp are libm calls, so the comparison is within FP32
    rounding."""
    dx = 0.25
    g = L.grid(nx, 1, nx * dx, dx)
    rp = L.run_params(dt=0.2, qm=-1.0, eps0=1.0, c=1.0)
    rp.B0[0] = 0.3
    rp.b0_prof, rp.b0_a, rp.b0_xc = 1, 1e-4, nx * dx / 2
    rp.cold_nc, rp.bnd_x, rp.bnd_nd, rp.bnd_numax = 0.8, 2, 20, 0.4
    rp.ant_amp, rp.ant_x0, rp.ant_sigma, rp.ant_w0 = 1e-3, nx / 3, 4.0, 0.25
    rp.ant_trmp, rp.ant_toff = 0.3, 100.0
    rng = np.random.default_rng(7)
    F = rng.normal(0, 1e-3, 9 * nx).astype(np.float32)
    F[6 * nx:] = 0.0
    sim = MaxwellSimulation(g, rp)
    names = ("ex", "ey", "ez", "bx", "by", "bz")
    sim.set_mesh(**{k: F[q * nx:(q + 1) * nx] for q, k in enumerate(names)})
    og = oracle.grid(nx, 1, dx, dx)
    yp = _oracle_params(oracle, rp)
    vcy = np.zeros(nx, np.float32)
    vcz = np.zeros(nx, np.float32)
    h = np.float32(0.5 * rp.dt)
    for k in range(3):
        t = k * rp.dt
        F = oracle.yee_faraday(og, F, h)
        F[6 * nx:] = 0.0
        F = oracle.yee_antenna(og, F, rp.ant_amp, rp.ant_x0, rp.ant_sigma, rp.ant_w0,
                               rp.ant_trmp, rp.ant_toff, t)
        F, vcy, vcz = oracle.yee_cold_fluid(og, yp, rp.cold_nc, F, vcy, vcz)
        F = oracle.yee_faraday(og, F, h)
        F = oracle.yee_ampere(og, F, np.float32(rp.c * rp.c), np.float32(rp.dt))
        F, vcy, vcz = oracle.yee_damp(og, rp.bnd_nd, rp.bnd_numax, rp.dt, F, vcy, vcz)
        sim.step()
        M = sim.mesh_flat()
        for q in range(9):
            a, b = M[q * nx:(q + 1) * nx], F[q * nx:(q + 1) * nx]
            print(f"nx={nx} step {k} grid {q}: rel {rel_l2(a, b):.2e}", flush=True)
            assert rel_l2(a, b) < 2e-6, (nx, k, q)
    sim.close()


@pytest.mark.parametrize("tile_sort", [0, 4])
def test_current_scale_guard_reports_a_runaway_particle(gpu, tile_sort):
    """The current's fixed-point scale assumes |uz/gamma| <= 2 max(1, c) and
    weights no larger than the run's largest: a particle past the first is
    reported by the next synchronize (once), and the run goes on."""
    nx = ny = 32
    g = L.grid(nx, ny, 16.0, 16.0)
    rp = L.run_params(dt=0.04, qm=-1.0, eps0=1.0, c=5.0)
    rp.B0[0], rp.tile_sort = 0.3, tile_sort
    rng = np.random.default_rng(4)
    n = 4000
    st = {"x": rng.uniform(0, nx, n), "y": rng.uniform(0, ny, n),
          "ux": rng.normal(0, 0.1, n), "uy": rng.normal(0, 0.1, n), "uz": rng.normal(0, 0.1, n)}
    sim = MaxwellSimulation(g, rp)
    sim.set_particles(st, w=np.full(n, 0.25, np.float32))
    sim.step()
    sim.synchronize()  # within the scale
    st = sim.particles()
    st["uz"][17] = 11.0  # past 2 c = 10
    sim.set_particles(st, w=np.full(n, 0.25, np.float32))
    sim.step()
    with pytest.raises(L.AwError, match="fixed-point scale"):
        sim.synchronize()
    sim.synchronize()  # reported once
    sim.close()
