"""GPU tests of the standalone adjoint primitives: aw_adj_push_reverse,
aw_adj_gather_t and aw_adj_deposit_t are the transposes of one push, of the
CIC gather and of the charge deposit.  Each is checked against central
differences of the FP64 oracle's forward map (exact up to rounding for the
piecewise-bilinear CIC weights when no particle crosses a cell face), and
gather_t against the FP64 deposit of the per-particle vectors, which is what
the transpose of an interpolation is."""
import ctypes as C

import numpy as np
import pytest
import torch

from arcwarden_amd import _lib as L

from gpu_util import Ctx, DevParticles, dev, host, random_particles, rel_l2, stream

pytestmark = pytest.mark.gpu

NX, NY, DX, DY = 16, 12, 0.5, 0.4
Q = 2.0 ** -10  # position quantum: x +- Q stays exact in float32


def _interior(rng, n, umax=0.3):
    """Positions whose fractional parts lie in [0.15, 0.85] on the 2^-10
    lattice, in every cell including the last (periodic neighbours)."""
    def coord(m):
        f = np.round(rng.uniform(0.15, 0.85, n) / Q) * Q
        return (rng.integers(0, m, n) + f).astype(np.float32)
    st = {"x": coord(NX), "y": coord(NY)}
    for k in ("ux", "uy", "uz"):
        st[k] = rng.uniform(-umax, umax, n).astype(np.float32)
    return st


def _field(rng, amp):
    X, Y = np.meshgrid(np.arange(NX) * 2 * np.pi / NX, np.arange(NY) * 2 * np.pi / NY)
    a, b, p = rng.uniform(-1, 1, 3)
    f = amp * (a * np.cos(X + p) + b * np.sin(2 * Y - p) + 0.2 * rng.standard_normal(X.shape))
    return f.astype(np.float32)


def _grid():
    return L.grid(NX, NY, NX * DX, NY * DY)


@pytest.mark.parametrize("per_w", [False, True])
def test_deposit_t_is_the_position_transpose_of_the_charge_deposit(gpu, oracle, per_w):
    rng = np.random.default_rng(5 + per_w)
    n, qm, wu = 48, -1.3, 0.7
    st = _interior(rng, n)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32) if per_w else None
    lrho = rng.standard_normal((NY, NX)).astype(np.float32)
    lx0 = rng.standard_normal(n).astype(np.float32)
    ly0 = rng.standard_normal(n).astype(np.float32)
    P = DevParticles(st, w=w, w_uniform=0.0 if per_w else wu)
    lx, ly, lr = dev(lx0), dev(ly0), dev(lrho.ravel())
    with Ctx(_grid()) as ctx:
        L.call("aw_adj_deposit_t", ctx.h, C.byref(P.struct()), C.c_void_p(lr.data_ptr()),
               C.c_double(qm), C.c_void_p(lx.data_ptr()), C.c_void_p(ly.data_ptr()), stream())
        got_x, got_y = host(lx) - lx0, host(ly) - ly0
    og = oracle.grid(NX, NY, DX, DY)
    exp = np.zeros((2, n))
    for p in range(n):
        wp = None if w is None else w[p:p + 1]
        for d, (ex, ey) in enumerate(((Q, 0.0), (0.0, Q))):
            rp = oracle.deposit_rho_f64(og, [st["x"][p] + ex], [st["y"][p] + ey], qm, w=wp,
                                        w_uniform=wu)
            rm = oracle.deposit_rho_f64(og, [st["x"][p] - ex], [st["y"][p] - ey], qm, w=wp,
                                        w_uniform=wu)
            exp[d, p] = np.sum(lrho.astype(np.float64) * (rp - rm)) / (2 * Q)
    scale = np.max(np.abs(exp))
    np.testing.assert_allclose(got_x, exp[0], rtol=1e-4, atol=2e-6 * scale)
    np.testing.assert_allclose(got_y, exp[1], rtol=1e-4, atol=2e-6 * scale)


@pytest.mark.parametrize("n", [1, 5_000, 300_000])
def test_gather_t_equals_the_deposit_of_the_particle_vectors(gpu, oracle, n):
    rng = np.random.default_rng(n)
    st = random_particles(rng, NX, NY, n)
    lex = (rng.standard_normal(n) * np.exp(rng.uniform(-6, 0, n))).astype(np.float32)
    ley = rng.standard_normal(n).astype(np.float32)
    P = DevParticles(st)
    a, b = dev(lex), dev(ley)
    base_x = rng.standard_normal(NX * NY).astype(np.float32)
    outs = []
    with Ctx(_grid()) as ctx:
        for start in (base_x, np.zeros(NX * NY, np.float32), np.zeros(NX * NY, np.float32)):
            lEx, lEy = dev(start), dev(np.zeros(NX * NY, np.float32))
            L.call("aw_adj_gather_t", ctx.h, C.byref(P.struct()), C.c_void_p(a.data_ptr()),
                   C.c_void_p(b.data_ptr()), C.c_void_p(lEx.data_ptr()),
                   C.c_void_p(lEy.data_ptr()), stream())
            outs.append((host(lEx), host(lEy)))
    og = oracle.grid(NX, NY, DX, DY)
    # the transpose of W: a deposit with weight le (coefficient area * le / area)
    ex = oracle.deposit_rho_f64(og, st["x"], st["y"], DX * DY, w=lex).ravel()
    ey = oracle.deposit_rho_f64(og, st["x"], st["y"], DX * DY, w=ley).ravel()
    assert rel_l2(outs[1][0], ex) < 1e-6
    assert rel_l2(outs[1][1], ey) < 1e-6
    # accumulates into the caller's array
    assert rel_l2(outs[0][0], base_x.astype(np.float64) + ex) < 1e-6
    # order-free integer sums: repeat calls agree bit for bit
    assert np.array_equal(outs[1][0], outs[2][0]) and np.array_equal(outs[1][1], outs[2][1])


def _push_reverse(ctx, P, Ex, Ey, rp, has_b0, half_back, lin):
    t = {k: dev(v) for k, v in lin.items()}
    n = len(lin["lx"])
    lex, ley = dev(np.zeros(n, np.float32)), dev(np.zeros(n, np.float32))
    L.call("aw_adj_push_reverse", ctx.h, C.byref(P.struct()), C.c_void_p(Ex.data_ptr()),
           C.c_void_p(Ey.data_ptr()), C.byref(rp), int(has_b0), int(half_back),
           *(C.c_void_p(t[k].data_ptr()) for k in ("lx", "ly", "lux", "luy")),
           C.c_void_p(lex.data_ptr()), C.c_void_p(ley.data_ptr()), stream())
    out = {k: host(v) for k, v in t.items()}
    out["lex"], out["ley"] = host(lex), host(ley)
    return out, lex, ley


@pytest.mark.parametrize("b0z", [0.0, 0.6])
@pytest.mark.parametrize("half_back", [0, 1])
def test_push_reverse_is_the_transpose_of_one_push(gpu, oracle, b0z, half_back):
    """The per-particle gradient of J = lx.x' + ly.y' + lux.ux' + luy.uy'
    with respect to (x, y, ux, uy) and to the gathered field (Ex, Ey), by
    central differences of the FP64 push; then gather_t of the field
    adjoint against differences over a random grid perturbation."""
    rng = np.random.default_rng(17 + 2 * half_back + (b0z > 0))
    n, qm, dt = 64, -1.0, 0.1
    st = _interior(rng, n)
    Ex, Ey = _field(rng, 0.8), _field(rng, 0.8)
    has_b0 = b0z != 0.0
    rp = L.run_params(dt=dt, qm=qm, B0=(0.0, 0.0, b0z))
    lin = {k: rng.standard_normal(n).astype(np.float32) for k in ("lx", "ly", "lux", "luy")}
    P = DevParticles(st)
    tEx, tEy = dev(Ex.ravel()), dev(Ey.ravel())
    with Ctx(_grid()) as ctx:
        got, lex_d, ley_d = _push_reverse(ctx, P, tEx, tEy, rp, has_b0, half_back, lin)
        dEx, dEy = rng.standard_normal((2, NY * NX)).astype(np.float32)
        lEx, lEy = dev(np.zeros(NX * NY, np.float32)), dev(np.zeros(NX * NY, np.float32))
        L.call("aw_adj_gather_t", ctx.h, C.byref(P.struct()), C.c_void_p(lex_d.data_ptr()),
               C.c_void_p(ley_d.data_ptr()), C.c_void_p(lEx.data_ptr()),
               C.c_void_p(lEy.data_ptr()), stream())
        lEx, lEy = host(lEx), host(lEy)

    og = oracle.grid(NX, NY, DX, DY)
    pp = oracle.push_params(qm, -0.5 * dt if half_back else dt, (0.0, 0.0, b0z), has_b0)
    s64 = {k: st[k].astype(np.float64) for k in st}
    l64 = {k: lin[k].astype(np.float64) for k in lin}

    def J(s, fx, fy):
        o = oracle.push_es_f64(og, pp, s, fx, fy, wrap=False)
        xo, yo = (s["x"], s["y"]) if half_back else (o["x"], o["y"])
        return l64["lx"] * xo + l64["ly"] * yo + l64["lux"] * o["ux"] + l64["luy"] * o["uy"]

    E64 = (Ex.astype(np.float64), Ey.astype(np.float64))
    h = 1e-6
    for key, coord in (("lx", "x"), ("ly", "y"), ("lux", "ux"), ("luy", "uy")):
        sp, sm = dict(s64), dict(s64)
        sp[coord] = s64[coord] + h
        sm[coord] = s64[coord] - h
        exp = (J(sp, *E64) - J(sm, *E64)) / (2 * h)
        np.testing.assert_allclose(got[key], exp, rtol=1e-4, atol=1e-5 * np.max(np.abs(exp)),
                                   err_msg=key)
    # a uniform field shift moves every gathered value by the same amount
    for key, fx, fy in (("lex", (E64[0] + h, E64[0] - h), (E64[1], E64[1])),
                        ("ley", (E64[0], E64[0]), (E64[1] + h, E64[1] - h))):
        exp = (J(s64, fx[0], fy[0]) - J(s64, fx[1], fy[1])) / (2 * h)
        np.testing.assert_allclose(got[key], exp, rtol=1e-4, atol=1e-5 * np.max(np.abs(exp)),
                                   err_msg=key)
    # <lE, dE> against the change of sum J under a random field perturbation
    d = (dEx.reshape(NY, NX).astype(np.float64), dEy.reshape(NY, NX).astype(np.float64))
    fd = (J(s64, E64[0] + h * d[0], E64[1] + h * d[1]).sum()
          - J(s64, E64[0] - h * d[0], E64[1] - h * d[1]).sum()) / (2 * h)
    ip = float(np.dot(lEx.astype(np.float64), dEx) + np.dot(lEy.astype(np.float64), dEy))
    assert ip == pytest.approx(fd, rel=1e-4, abs=1e-6 * np.abs(lEx).sum())


def test_adjoint_primitives_reject_bad_arguments(gpu):
    rng = np.random.default_rng(3)
    st = _interior(rng, 8)
    P = DevParticles(st)
    z = [dev(np.zeros(max(8, NX * NY), np.float32)) for _ in range(8)]
    ptr = [C.c_void_p(t.data_ptr()) for t in z]
    with Ctx(_grid()) as ctx:
        rp = L.run_params(dt=0.1, qm=-1.0, B0=(0.2, 0.0, 0.5))
        with pytest.raises(L.AwError, match="along z"):
            L.call("aw_adj_push_reverse", ctx.h, C.byref(P.struct()), ptr[0], ptr[1],
                   C.byref(rp), 1, 0, *ptr[2:8], stream())
        rp = L.run_params(dt=0.1, qm=-1.0, rel=1)
        with pytest.raises(L.AwError, match="non-relativistic"):
            L.call("aw_adj_push_reverse", ctx.h, C.byref(P.struct()), ptr[0], ptr[1],
                   C.byref(rp), 0, 0, *ptr[2:8], stream())
        with pytest.raises(L.AwError):
            L.call("aw_adj_gather_t", ctx.h, C.byref(P.struct()), None, ptr[1], ptr[2], ptr[3],
                   stream())
        with pytest.raises(L.AwError):
            L.call("aw_adj_deposit_t", ctx.h, C.byref(P.struct()), None, C.c_double(1.0),
                   ptr[0], ptr[1], stream())
        # no particles: nothing to do, nothing written
        E = DevParticles.empty(0)
        L.call("aw_adj_gather_t", ctx.h, C.byref(E.struct()), None, None, ptr[2], ptr[3],
               stream())
        torch.cuda.synchronize()
