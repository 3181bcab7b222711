"""GPU tests of the standalone Darwin operators (any particle order):
aw_deposit_rho_j_amu and aw_deposit_dcu against the FP64 oracle,
aw_push_boris_em bit for bit against the host FP32 restatement, and
aw_form_e against its closed form."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from arcwarden_amd import _lib as L

from gpu_util import Ctx, DevParticles, dev, host, random_particles, rel_l2, stream

pytestmark = pytest.mark.gpu


def _grids(n, k, nx, ny):
    out = [torch.zeros(nx * ny, dtype=torch.float32, device="cuda") for _ in range(k)]
    return out


def _sources(grids, names):
    s = L.AwSources()
    for name, t in zip(names, grids):
        setattr(s, name, t.data_ptr())
    return s


def _fields(E, B=None, EL=None, ET=None):
    f = L.AwFields()
    for name, t in zip(("Ex", "Ey", "Ez"), E):
        setattr(f, name, t.data_ptr())
    for name, t in zip(("Bx", "By", "Bz"), B or ()):
        setattr(f, name, t.data_ptr())
    for name, t in zip(("ELx", "ELy"), EL or ()):
        setattr(f, name, t.data_ptr())
    for name, t in zip(("ETx", "ETy", "ETz"), ET or ()):
        setattr(f, name, t.data_ptr())
    return f


def _smooth_fields(rng, nx, ny, amp):
    X, Y = np.meshgrid(np.arange(nx) * 2 * np.pi / nx, np.arange(ny) * 2 * np.pi / ny)
    out = []
    for _ in range(6):
        a, b, p = rng.uniform(-1, 1, 3)
        out.append((amp * (a * np.cos(X + p) + b * np.sin(2 * Y - p) + 0.3 * a * b))
                   .astype(np.float32))
    return out


@pytest.mark.parametrize("per_w", [False, True])
@pytest.mark.parametrize("nx,ny,n", [(64, 48, 200_000), (40, 24, 50_000), (16, 4, 3_000)])
def test_deposit_rho_j_amu_vs_fp64(gpu, oracle, nx, ny, n, per_w):
    rng = np.random.default_rng(nx * 3 + n + per_w)
    st = random_particles(rng, nx, ny, n, umax=0.4)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32) * 1e-3 if per_w else None
    wu = 0.0 if per_w else 1.7e-3
    dx, dy = 0.5, 0.7
    g = L.grid(nx, ny, nx * dx, ny * dy)
    og = oracle.grid(nx, ny, dx, dy)
    rho, jx, jy, jz = oracle.deposit_rho_j_f64(og, st, -1.0, w=w, w_uniform=wu)
    amu = oracle.deposit_amu_f64(og, st, -1.0, w=w, w_uniform=wu)
    with Ctx(g, 8, 8) as ctx:
        p = DevParticles(st, w=w, w_uniform=wu)
        out = _grids(n, 8, nx, ny)
        names = ("rho", "Jx", "Jy", "Jz", "amu0", "amu1", "amu2", "amu3")
        src = _sources(out, names)
        L.call("aw_deposit_rho_j_amu", ctx.h, C.byref(p.struct()), C.c_double(-1.0),
               C.byref(src), stream())
        got = [host(t).reshape(ny, nx) for t in out]
        # accumulates: a second call doubles every grid exactly
        L.call("aw_deposit_rho_j_amu", ctx.h, C.byref(p.struct()), C.c_double(-1.0),
               C.byref(src), stream())
        twice = [host(t).reshape(ny, nx) for t in out]
    for name, a, ref in zip(names, got, (rho, jx, jy, jz, *amu)):
        assert rel_l2(a, ref) < 1e-6, name
    for a, b in zip(got, twice):
        assert np.array_equal(2 * a, b)


def test_deposit_dcu_vs_fp64(gpu, oracle):
    nx, ny, n = 48, 32, 120_000
    rng = np.random.default_rng(5)
    st = random_particles(rng, nx, ny, n, umax=0.3)
    F = _smooth_fields(rng, nx, ny, 0.2)
    dx, dy, wu = 0.5, 0.5, 2e-3
    g = L.grid(nx, ny, nx * dx, ny * dy)
    og = oracle.grid(nx, ny, dx, dy)
    ref = oracle.deposit_dcu_f64(og, st, -1.0, F[:3], F[3:], w_uniform=wu)
    with Ctx(g) as ctx:
        p = DevParticles(st, w_uniform=wu)
        Fd = [dev(a.ravel()) for a in F]
        f = _fields(Fd[:3], Fd[3:])
        out = _grids(n, 3, nx, ny)
        src = _sources(out, ("dcux", "dcuy", "dcuz"))
        L.call("aw_deposit_dcu", ctx.h, C.byref(p.struct()), C.c_double(-1.0), C.byref(f),
               C.byref(src), stream())
        got = [host(t).reshape(ny, nx) for t in out]
    for q in range(3):
        assert rel_l2(got[q], ref[q]) < 1e-5, q


def test_push_boris_em_bit_exact(gpu, oracle):
    nx, ny, n = 40, 28, 100_000
    rng = np.random.default_rng(9)
    st = random_particles(rng, nx, ny, n, umax=0.5)
    F = _smooth_fields(rng, nx, ny, 0.4)
    dx, dy, dt = 0.5, 0.5, 0.1
    g = L.grid(nx, ny, nx * dx, ny * dy)
    rp = L.run_params(dt=dt, qm=-1.0)
    ref = oracle.push_em_f32(oracle.grid(nx, ny, dx, dy), -1.0, dt, st, F[:3], F[3:])
    with Ctx(g) as ctx:
        p = DevParticles(st, w_uniform=1.0)
        Fd = [dev(a.ravel()) for a in F]
        f = _fields(Fd[:3], Fd[3:])
        L.call("aw_push_boris_em", ctx.h, C.byref(p.struct()), C.byref(f), C.byref(rp),
               stream())
        out = p.numpy()
    for k in ("x", "y", "ux", "uy", "uz"):
        assert np.array_equal(out[k], ref[k]), k


@pytest.mark.parametrize("with_et", [False, True])
def test_form_e_closed_form(gpu, with_et):
    nx, ny = 32, 8
    dx = 0.5
    g = L.grid(nx, ny, nx * dx, ny * dx)
    rng = np.random.default_rng(3)
    EL = [rng.normal(size=nx * ny).astype(np.float32) for _ in range(2)]
    ET = [rng.normal(size=nx * ny).astype(np.float32) for _ in range(3)]
    rp = L.run_params(dt=0.1)
    rp.pump = 1
    rp.pump_ex, rp.pump_ey, rp.pump_ez = 0.3, 0.2, -0.1
    rp.pump_k0, rp.pump_w0, rp.pump_trmp, rp.pump_toff = 2 * math.pi / (nx * dx), 0.4, 2.0, 50.0
    t = 1.3
    with Ctx(g) as ctx:
        E = [torch.zeros(nx * ny, dtype=torch.float32, device="cuda") for _ in range(3)]
        ELd = [dev(a) for a in EL]
        ETd = [dev(a) for a in ET] if with_et else None
        f = _fields(E, EL=ELd, ET=ETd)
        L.call("aw_form_e", ctx.h, C.byref(f), C.byref(rp), C.c_double(t), stream())
        got = [host(a) for a in E]
    x = (np.arange(nx * ny) % nx) * dx
    th = rp.pump_k0 * x - rp.pump_w0 * t
    r = t / rp.pump_trmp
    pump = (0.3 * np.cos(th) * r, -0.2 * np.sin(th) * r, -0.1 * np.cos(th) * r)
    base = [EL[0], EL[1], np.zeros(nx * ny)]
    if with_et:
        base = [EL[0] + ET[0], EL[1] + ET[1], ET[2]]
    for q in range(3):
        assert np.max(np.abs(got[q] - (base[q] + pump[q]))) < 1e-5, q


def test_deposit_rho_j_amu_partial_outputs(gpu):
    """NULL members of the output are skipped (the facade's charge_current
    and deposit_amu): the grids that are given equal the full deposit's."""
    rng = np.random.default_rng(12)
    nx, ny, n = 48, 32, 100_000
    st = random_particles(rng, nx, ny, n, umax=0.4)
    g = L.grid(nx, ny, nx * 0.5, ny * 0.5)
    names = ("rho", "Jx", "Jy", "Jz", "amu0", "amu1", "amu2", "amu3")
    with Ctx(g) as ctx:
        p = DevParticles(st, w_uniform=1e-3)
        full = _grids(n, 8, nx, ny)
        L.call("aw_deposit_rho_j_amu", ctx.h, C.byref(p.struct()), C.c_double(-1.0),
               C.byref(_sources(full, names)), stream())
        rj = _grids(n, 4, nx, ny)
        L.call("aw_deposit_rho_j_amu", ctx.h, C.byref(p.struct()), C.c_double(-1.0),
               C.byref(_sources(rj, names[:4])), stream())
        am = _grids(n, 4, nx, ny)
        L.call("aw_deposit_rho_j_amu", ctx.h, C.byref(p.struct()), C.c_double(-1.0),
               C.byref(_sources(am, names[4:])), stream())
        for a, b in zip(rj + am, full):
            assert np.array_equal(host(a), host(b))
        with pytest.raises(L.AwError):
            L.call("aw_deposit_rho_j_amu", ctx.h, C.byref(p.struct()), C.c_double(-1.0),
                   C.byref(L.AwSources()), stream())
