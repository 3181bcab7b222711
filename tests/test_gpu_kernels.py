"""GPU parity of the individual operators against the CPU oracle.

Integer outputs (cells, tile bins, stable permutations) and the FP32
same-algorithm particle updates are compared bit for bit; the fixed-point
deposits are bit-exact against the oracle's fixed-point restatement and
within 1e-7 relative L2 of its FP64 deposit; spectral solves are compared
against closed forms and the FP64 oracle."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from arcwarden_amd import _lib as L

from gpu_util import Ctx, DevParticles, dev, host, random_particles, rel_l2, stream

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi


def _grid(nx, ny, dx, dy):
    return L.AwGrid(nx, ny, dx, dy, nx * dx, ny * dy)


# ---- integer layer ----------------------------------------------------------------------

@pytest.mark.parametrize("nx,ny,n", [(50, 30, 100_000), (64, 64, 262_144), (16, 4, 1), (8, 8, 0)])
def test_cells_bins_and_sorts_bit_exact(gpu, oracle, nx, ny, n):
    rng = np.random.default_rng(nx * 1000 + n)
    st = random_particles(rng, nx, ny, n)
    g = _grid(nx, ny, 0.5, 0.5)
    og = oracle.grid(nx, ny, 0.5, 0.5)
    with Ctx(g) as ctx:
        p = DevParticles(st)
        ps = p.struct()
        cell = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
        L.call("aw_cell_index", ctx.h, C.byref(ps), C.c_void_p(cell.data_ptr()), stream())
        off = torch.empty(ctx.ntiles + 1, dtype=torch.int32, device="cuda")
        idx = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
        L.call("aw_bin_tiles", ctx.h, C.byref(ps), C.c_void_p(off.data_ptr()),
               C.c_void_p(idx.data_ptr()), stream())
        nkeys = ctx.ntiles * ctx.tx * ctx.ty
        coff = torch.empty(nkeys + 1, dtype=torch.int32, device="cuda")
        cperm = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
        L.call("aw_sort_cells", ctx.h, C.byref(ps), C.c_void_p(cperm.data_ptr()),
               C.c_void_p(coff.data_ptr()), stream())
        q = DevParticles.empty(n)
        qs = q.struct()
        soff = torch.empty(ctx.ntiles + 1, dtype=torch.int32, device="cuda")
        L.call("aw_sort_by_tile", ctx.h, C.byref(ps), C.byref(qs), C.c_void_p(soff.data_ptr()),
               stream())
        torch.cuda.synchronize()
        ocell = oracle.cell_index(og, st["x"], st["y"])
        assert np.array_equal(host(cell)[:n], ocell)
        operm, ooff = oracle.counting_sort(oracle.tile_ids(og, ctx.tx, ctx.ty, st["x"], st["y"]),
                                           ctx.ntiles)
        assert np.array_equal(host(off), ooff)
        assert np.array_equal(host(idx)[:n], operm)
        assert np.array_equal(host(soff), ooff)
        cp, co = oracle.counting_sort(oracle.cell_keys(og, ctx.tx, ctx.ty, st["x"], st["y"]), nkeys)
        assert np.array_equal(host(cperm)[:n], cp)
        assert np.array_equal(host(coff), co)
        moved = q.numpy()
        for k in ("x", "y", "ux", "uy", "uz"):
            assert np.array_equal(moved[k], st[k][operm])


@pytest.mark.parametrize("nx,ny,tx,ty", [(1000, 24, 7, 5), (250, 250, 3, 3), (96, 40, 12, 40),
                                         (64, 64, 1, 64), (512, 512, 2, 2)])
def test_tile_sort_odd_tiles_bit_exact(gpu, oracle, nx, ny, tx, ty):
    """The one-pass tile sort and tile binning (tile ids from an FP32
    reciprocal corrected to the exact quotient) on tile sides that are not
    powers of two, with particles on both edges of every cell: the same
    stable permutation and offsets as the oracle's counting sort.  512x512
    with 2x2 tiles has 65,536 tiles, whose counters exceed LDS: the radix
    sort path."""
    i = np.arange(nx, dtype=np.float32)
    j = np.arange(ny, dtype=np.float32)
    lo_x, hi_x = i, np.nextafter(i + 1, np.float32(0)).astype(np.float32)
    lo_y, hi_y = j, np.nextafter(j + 1, np.float32(0)).astype(np.float32)
    X, Y = [], []
    for xs in (lo_x, hi_x):
        for ys in (lo_y, hi_y):
            gx, gy = np.meshgrid(xs, ys)
            X.append(gx.ravel())
            Y.append(gy.ravel())
    rng = np.random.default_rng(nx + tx)
    order = rng.permutation(4 * nx * ny)
    st = {"x": np.concatenate(X)[order].astype(np.float32),
          "y": np.concatenate(Y)[order].astype(np.float32)}
    for k in ("ux", "uy", "uz"):
        st[k] = rng.uniform(-1, 1, st["x"].size).astype(np.float32)
    og = oracle.grid(nx, ny, 0.5, 0.5)
    with Ctx(_grid(nx, ny, 0.5, 0.5), tx, ty) as ctx:
        assert (ctx.tx, ctx.ty) == (tx, ty)
        p = DevParticles(st)
        q = DevParticles.empty(st["x"].size)
        soff = torch.empty(ctx.ntiles + 1, dtype=torch.int32, device="cuda")
        L.call("aw_sort_by_tile", ctx.h, C.byref(p.struct()), C.byref(q.struct()),
               C.c_void_p(soff.data_ptr()), stream())
        boff = torch.empty(ctx.ntiles + 1, dtype=torch.int32, device="cuda")
        bidx = torch.empty(st["x"].size, dtype=torch.int32, device="cuda")
        L.call("aw_bin_tiles", ctx.h, C.byref(p.struct()), C.c_void_p(boff.data_ptr()),
               C.c_void_p(bidx.data_ptr()), stream())
        torch.cuda.synchronize()
        operm, ooff = oracle.counting_sort(oracle.tile_ids(og, tx, ty, st["x"], st["y"]),
                                           ctx.ntiles)
        assert np.array_equal(host(soff), ooff)
        assert np.array_equal(host(boff), ooff)
        assert np.array_equal(host(bidx), operm)
        moved = q.numpy()
        for k in ("x", "y", "ux", "uy", "uz"):
            assert np.array_equal(moved[k], st[k][operm]), k


def test_migrate_bit_exact(gpu, oracle):
    rng = np.random.default_rng(5)
    nx, ny, n = 40, 24, 50_000
    st = random_particles(rng, nx, ny, n)
    st["x"] = (st["x"] + rng.choice([-3 * nx, -nx, -0.3, 0, 0.4, nx, 2.5 * nx], n)).astype(np.float32)
    st["y"] = (st["y"] + rng.choice([-ny, -1e-7, 0, ny, 5 * ny], n)).astype(np.float32)
    with Ctx(_grid(nx, ny, 1.0, 1.0)) as ctx:
        p = DevParticles(st)
        cell = torch.empty(n, dtype=torch.int32, device="cuda")
        L.call("aw_migrate", ctx.h, C.byref(p.struct()), C.c_void_p(cell.data_ptr()), stream())
        xo, yo, co = oracle.migrate(oracle.grid(nx, ny, 1.0, 1.0), st["x"], st["y"])
        out = p.numpy()
        assert np.array_equal(out["x"], xo) and np.array_equal(out["y"], yo)
        assert np.array_equal(host(cell), co)


# ---- deposits --------------------------------------------------------------------------

@pytest.mark.parametrize("per_w", [False, True])
def test_deposit_policies_bit_exact(gpu, oracle, per_w):
    rng = np.random.default_rng(11)
    nx, ny, n = 48, 40, 200_000
    st = random_particles(rng, nx, ny, n)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32) if per_w else None
    wu = 0.0 if per_w else 3.0e-3
    g = _grid(nx, ny, 0.5, 0.25)
    og = oracle.grid(nx, ny, 0.5, 0.25)
    S = 40
    ref_fix, _ = oracle.deposit_rho_fix(og, st["x"], st["y"], -1.0, S, w=w, w_uniform=wu)
    ref64 = oracle.deposit_rho_f64(og, st["x"], st["y"], -1.0, w=w, w_uniform=wu)
    with Ctx(g) as ctx:
        L.call("aw_ctx_set_fixed_scale", ctx.h, S)
        p = DevParticles(st, w=w, w_uniform=wu)
        for pol in (L.AW_DEPOSIT_GLOBAL, L.AW_DEPOSIT_LDS_GRID, L.AW_DEPOSIT_BINNED,
                    L.AW_DEPOSIT_AUTO):
            rho = torch.zeros(nx * ny, dtype=torch.float32, device="cuda")
            L.call("aw_deposit_rho", ctx.h, C.byref(p.struct()), C.c_double(-1.0), pol,
                   C.c_void_p(rho.data_ptr()), stream())
            got = host(rho).reshape(ny, nx)
            assert np.array_equal(got, ref_fix), f"policy {pol}"
            assert rel_l2(got, ref64) < 1e-7


def test_deposit_closed_forms(gpu):
    g = _grid(8, 8, 0.125, 0.125)
    w = 1.0 * 0.125 * 0.125 / 32
    with Ctx(g) as ctx:
        for (x, y) in [(2.3, 3.7), (7.6, 7.6)]:
            st = {"x": np.array([x], np.float32), "y": np.array([y], np.float32),
                  "ux": np.zeros(1, np.float32), "uy": np.zeros(1, np.float32),
                  "uz": np.zeros(1, np.float32)}
            p = DevParticles(st, w_uniform=w)
            rho = torch.zeros(64, dtype=torch.float32, device="cuda")
            L.call("aw_deposit_rho", ctx.h, C.byref(p.struct()), C.c_double(-1.0),
                   L.AW_DEPOSIT_AUTO, C.c_void_p(rho.data_ptr()), stream())
            r = host(rho).reshape(8, 8)
            assert np.count_nonzero(r) == 4
            assert r.sum() * 0.125 * 0.125 == pytest.approx(-w, rel=1e-6)
        assert r[0, 0] != 0 and r[7, 0] != 0 and r[0, 7] != 0


# ---- push --------------------------------------------------------------------------

def _efield(rng, nx, ny):
    return (rng.normal(0, 0.05, (ny, nx)).astype(np.float32),
            rng.normal(0, 0.05, (ny, nx)).astype(np.float32))


@pytest.mark.parametrize("has_b0", [True, False])
def test_push_and_half_back_bit_exact(gpu, oracle, has_b0):
    rng = np.random.default_rng(21)
    nx, ny, n = 64, 48, 100_000
    st = random_particles(rng, nx, ny, n)
    ex, ey = _efield(rng, nx, ny)
    B0 = (0.25, -0.1, 0.43)
    rp = L.run_params(dt=0.1, qm=-1.0, B0=B0)
    og = oracle.grid(nx, ny, 0.5, 0.7)
    pp = oracle.push_params(-1.0, 0.1, B0, has_b0)
    with Ctx(_grid(nx, ny, 0.5, 0.7)) as ctx:
        p = DevParticles(st)
        E = dev(ex), dev(ey)
        L.call("aw_push_es", ctx.h, C.byref(p.struct()), C.c_void_p(E[0].data_ptr()),
               C.c_void_p(E[1].data_ptr()), C.byref(rp), int(has_b0), stream())
        ref = oracle.push_es_f32(og, pp, st, ex, ey, False)
        out = p.numpy()
        for k in ("x", "y", "ux", "uy", "uz"):
            assert np.array_equal(out[k], ref[k]), k
        q = DevParticles(st)
        L.call("aw_half_step_back", ctx.h, C.byref(q.struct()), C.c_void_p(E[0].data_ptr()),
               C.c_void_p(E[1].data_ptr()), C.byref(rp), int(has_b0), stream())
        ref = oracle.half_back_f32(og, pp, st, ex, ey)
        out = q.numpy()
        for k in ("x", "y", "ux", "uy", "uz"):
            assert np.array_equal(out[k], ref[k]), k


# B0 along z takes the kernel variant that neither rotates nor stores uz
@pytest.mark.parametrize("B0", [(0.0, 0.0, 0.5), (0.25, -0.1, 0.43)])
@pytest.mark.parametrize("nx,ny,tile,stale", [(64, 48, 16, False), (64, 48, 16, True),
                                              (50, 30, 8, True), (64, 4, 16, False)])
def test_fused_push_deposit_bit_exact(gpu, oracle, nx, ny, tile, stale, B0):
    rng = np.random.default_rng(31 + nx + tile)
    n = 150_000
    st = random_particles(rng, nx, ny, n, umax=2.0)
    ex, ey = _efield(rng, nx, ny)
    dx, dy, dt = 0.5, 0.5, 0.1
    rp = L.run_params(dt=dt, qm=-1.0, B0=B0)
    og = oracle.grid(nx, ny, dx, dy)
    pp = oracle.push_params(-1.0, dt, B0, True)
    wu = 2.0e-3
    S = 40
    with Ctx(_grid(nx, ny, dx, dy), tile, tile) as ctx:
        L.call("aw_ctx_set_fixed_scale", ctx.h, S)
        src = DevParticles(st, w_uniform=wu)
        dst = DevParticles.empty(n, w_uniform=wu)
        off = torch.empty(ctx.ntiles + 1, dtype=torch.int32, device="cuda")
        L.call("aw_sort_by_tile", ctx.h, C.byref(src.struct()), C.byref(dst.struct()),
               C.c_void_p(off.data_ptr()), stream())
        sorted_st = dst.numpy()
        if stale:  # particles far from their (stale) tile exercise the fallbacks
            sel = rng.random(n) < 0.05
            shift = rng.uniform(-12, 12, n).astype(np.float32)
            xs = np.mod(sorted_st["x"] + np.where(sel, shift, 0), nx).astype(np.float32)
            xs[xs >= nx] = 0.0
            sorted_st["x"] = xs
            dst = DevParticles(sorted_st, w_uniform=wu)
        E = dev(ex), dev(ey)
        rho = torch.zeros(nx * ny, dtype=torch.float32, device="cuda")
        strays = C.c_int64(-1)
        L.call("aw_push_deposit_es", ctx.h, C.byref(dst.struct()), C.c_void_p(off.data_ptr()),
               C.c_void_p(E[0].data_ptr()), C.c_void_p(E[1].data_ptr()), C.byref(rp), 1,
               C.c_void_p(rho.data_ptr()), C.byref(strays), stream())
        ref = oracle.push_es_f32(og, pp, sorted_st, ex, ey, True)
        out = dst.numpy()
        for k in ("x", "y", "ux", "uy", "uz"):
            assert np.array_equal(out[k], ref[k]), k
        ref_rho, _ = oracle.deposit_rho_fix(og, ref["x"], ref["y"], -1.0, S, w_uniform=wu)
        assert np.array_equal(host(rho).reshape(ny, nx), ref_rho)
        if stale:
            assert strays.value > 0
        else:
            assert strays.value >= 0


# ---- spectral -------------------------------------------------------------------------------

def test_fft_roundtrip_and_parseval(gpu):
    nx, ny = 64, 32
    g = _grid(nx, ny, TWO_PI / nx, TWO_PI / ny)
    x = np.arange(nx) * g.dx
    y = np.arange(ny) * g.dy
    X, Y = np.meshgrid(x, y)
    f = (0.7 * np.sin(X) + 0.5 * np.cos(2 * X + Y) - 0.3 * np.sin(3 * Y) + 1.0).astype(np.float32)
    with Ctx(g) as ctx:
        fd = dev(f)
        k = torch.empty((ny, nx // 2 + 1, 2), dtype=torch.float32, device="cuda")
        L.call("aw_fft_r2c", ctx.h, C.c_void_p(fd.data_ptr()), C.c_void_p(k.data_ptr()), stream())
        F = host(k)
        F = F[..., 0].astype(np.float64) + 1j * F[..., 1]
        w = np.full(F.shape[1], 2.0)
        w[0] = w[nx // 2] = 1.0
        ek = (np.abs(F) ** 2 * w[None, :]).sum() / f.size
        ex = (f.astype(np.float64) ** 2).sum()
        assert abs(ek - ex) / ex < 1e-4
        back = torch.empty_like(fd)
        L.call("aw_fft_c2r", ctx.h, C.c_void_p(k.data_ptr()), C.c_void_p(back.data_ptr()), stream())
        assert np.max(np.abs(host(back) - f)) < 1e-4


def test_poisson_sine(gpu):
    nx, ny = 64, 32
    g = _grid(nx, ny, TWO_PI / nx, TWO_PI / ny)
    x = np.arange(nx) * g.dx
    rho = np.tile(np.sin(x), (ny, 1)).astype(np.float32)
    with Ctx(g) as ctx:
        ex = torch.empty(nx * ny, dtype=torch.float32, device="cuda")
        ey = torch.empty_like(ex)
        L.call("aw_solve_es", ctx.h, C.c_void_p(dev(rho).data_ptr()), C.c_void_p(ex.data_ptr()),
               C.c_void_p(ey.data_ptr()), C.c_double(1.0), stream())
        e = host(ex).reshape(ny, nx)
        assert np.max(np.abs(e - (-np.cos(x))[None, :])) < 1e-4
        assert np.max(np.abs(host(ey))) < 1e-4
        assert np.max(np.abs(e)) == pytest.approx(1.0, abs=1e-3)


@pytest.mark.parametrize("nx,ny", [(64, 64), (96, 48), (512, 512), (16, 32), (128, 32), (2, 8)])
def test_solve_es_vs_fp64_oracle(gpu, oracle, nx, ny):
    rng = np.random.default_rng(nx + ny)
    rho = rng.normal(size=(ny, nx)).astype(np.float32)
    g = _grid(nx, ny, 0.5, 0.5)
    ex64, ey64 = oracle.solve_es_f64(oracle.grid(nx, ny, 0.5, 0.5), rho)
    with Ctx(g) as ctx:
        ex = torch.empty(nx * ny, dtype=torch.float32, device="cuda")
        ey = torch.empty_like(ex)
        L.call("aw_solve_es", ctx.h, C.c_void_p(dev(rho).data_ptr()), C.c_void_p(ex.data_ptr()),
               C.c_void_p(ey.data_ptr()), C.c_double(1.0), stream())
        assert rel_l2(host(ex), ex64) < 1e-6
        assert rel_l2(host(ey), ey64) < 1e-6


def test_reduce_energies_closed_form(gpu):
    """tests/test_diagnostics.cu: uniform Ex=.3, Ey=-.4, rho=2 on 16x8, dx=dy=.25."""
    g = _grid(16, 8, 0.25, 0.25)
    with Ctx(g) as ctx:
        nc = 128
        ex, ey, rho = (dev(np.full(nc, v, np.float32)) for v in (0.3, -0.4, 2.0))
        st = {k: np.full(10, 0.5, np.float32) for k in ("x", "y", "ux", "uy", "uz")}
        p = DevParticles(st, w_uniform=0.25)
        d = L.AwDiagSample()
        L.call("aw_reduce_energies", ctx.h, C.byref(p.struct()), C.c_void_p(ex.data_ptr()),
               C.c_void_p(ey.data_ptr()), C.c_void_p(rho.data_ptr()), C.c_double(1.0),
               C.byref(d), stream())
        assert d.ee == pytest.approx(1.0, abs=1e-6)
        assert d.max_e == pytest.approx(0.5, abs=1e-7)
        assert d.charge == pytest.approx(16.0, abs=1e-5)
        assert d.ke == pytest.approx(0.25 * 0.5 * 10 * 0.75, rel=1e-12)


def test_every_particle_operator_accepts_empty_input(gpu):
    """n = 0 through every particle entry point: each returns AW_OK, leaves
    the grids it accumulates into untouched, and reports zero energy."""
    nx, ny = 16, 8
    nc = nx * ny
    g = _grid(nx, ny, 0.5, 0.5)
    rp = L.run_params(dt=0.1, qm=-1.0, eps0=1.0, c=5.0, B0=(0.1, 0.0, 0.2))
    with Ctx(g) as ctx:
        p = DevParticles.empty(0, w_uniform=1e-3)
        ps = p.struct()
        q = DevParticles.empty(0, w_uniform=1e-3)
        qs = q.struct()
        grids = {k: torch.full((nc,), 7.0, dtype=torch.float32, device="cuda")
                 for k in ("rho", "Jx", "Jy", "Jz", "dcux", "dcuy", "dcuz", "amu0", "amu1",
                           "amu2", "amu3", "jx", "jy", "jz")}
        fld = {k: torch.full((nc,), 0.01, dtype=torch.float32, device="cuda")
               for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz")}
        keep = {k: v.clone() for k, v in grids.items()}
        x0 = torch.zeros(1, dtype=torch.float32, device="cuda")
        cell = torch.zeros(1, dtype=torch.int32, device="cuda")
        off = torch.full((ctx.ntiles + 1,), -1, dtype=torch.int32, device="cuda")
        s = stream()
        vp = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
        for pol in (L.AW_DEPOSIT_GLOBAL, L.AW_DEPOSIT_LDS_GRID, L.AW_DEPOSIT_BINNED,
                    L.AW_DEPOSIT_AUTO):
            L.call("aw_deposit_rho", ctx.h, C.byref(ps), -1.0, pol, vp(grids["rho"]), s)
            L.call("aw_deposit_esirkepov", ctx.h, vp(x0), vp(x0), C.byref(ps), -1.0, 0.1, pol,
                   vp(grids["jx"]), vp(grids["jy"]), vp(grids["jz"]), s)
        L.call("aw_cell_index", ctx.h, C.byref(ps), vp(cell), s)
        L.call("aw_migrate", ctx.h, C.byref(ps), vp(cell), s)
        L.call("aw_sort_by_tile", ctx.h, C.byref(ps), C.byref(qs), vp(off), s)
        L.call("aw_push_es", ctx.h, C.byref(ps), vp(fld["Ex"]), vp(fld["Ey"]), C.byref(rp), 1, s)
        L.call("aw_half_step_back", ctx.h, C.byref(ps), vp(fld["Ex"]), vp(fld["Ey"]),
               C.byref(rp), 1, s)
        strays = C.c_int64(-1)
        L.call("aw_push_deposit_es", ctx.h, C.byref(ps), vp(off), vp(fld["Ex"]), vp(fld["Ey"]),
               C.byref(rp), 1, vp(grids["rho"]), C.byref(strays), s)
        src = L.AwSources()
        for k in ("rho", "Jx", "Jy", "Jz", "dcux", "dcuy", "dcuz", "amu0", "amu1", "amu2",
                  "amu3"):
            setattr(src, k, grids[k].data_ptr())
        f = L.AwFields()
        for k, t in fld.items():
            setattr(f, k, t.data_ptr())
        L.call("aw_deposit_rho_j_amu", ctx.h, C.byref(ps), -1.0, C.byref(src), s)
        L.call("aw_deposit_dcu", ctx.h, C.byref(ps), -1.0, C.byref(f), C.byref(src), s)
        L.call("aw_push_boris_em", ctx.h, C.byref(ps), C.byref(f), C.byref(rp), s)
        d = L.AwDiagSample()
        L.call("aw_reduce_energies", ctx.h, C.byref(ps), vp(fld["Ex"]), vp(fld["Ey"]), None,
               1.0, C.byref(d), s)
        torch.cuda.synchronize()
        for k in grids:
            assert torch.equal(grids[k], keep[k]), k
        assert np.array_equal(host(off), np.zeros(ctx.ntiles + 1, np.int32))
        assert strays.value == 0
        assert d.ke == 0.0 and d.charge == 0.0
        assert d.ee == pytest.approx(0.5 * 2 * 0.01 ** 2 * nc * 0.25, rel=1e-6)
