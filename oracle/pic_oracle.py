"""numpy binding of liboracle_pic.so -- the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the
bench's cpu_baseline leg, as the checker.  The product library never uses it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_pic.so"


class OrcGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("dx", C.c_double), ("dy", C.c_double)]


class OrcPush(C.Structure):
    _fields_ = [("qm", C.c_double), ("dt", C.c_double), ("B0", C.c_float * 3),
                ("has_b0", C.c_int32)]


class OrcSpecies(C.Structure):
    _fields_ = [("ppc", C.c_int32), ("index", C.c_int32), ("base", C.c_int64),
                ("count", C.c_int64), ("weight", C.c_double), ("uth", C.c_double * 3),
                ("ufl", C.c_double * 3), ("noisy", C.c_int32), ("seed", C.c_uint64)]


class OrcYee(C.Structure):
    _fields_ = [("qm", C.c_double), ("dt", C.c_double), ("c", C.c_double),
                ("B0", C.c_float * 3),
                ("rel", C.c_int32), ("pump", C.c_int32), ("deltaf", C.c_int32),
                ("df_wprec", C.c_int32), ("df_dist", C.c_int32), ("b0_prof", C.c_int32),
                ("bnd_x", C.c_int32), ("bnd_nd", C.c_int32),
                ("pump_ex", C.c_double), ("pump_ey", C.c_double), ("pump_ez", C.c_double),
                ("pump_k0", C.c_double), ("pump_w0", C.c_double), ("pump_trmp", C.c_double),
                ("pump_toff", C.c_double),
                ("df_tpar", C.c_double), ("df_tperp", C.c_double), ("df_kappa", C.c_double),
                ("df_rho", C.c_double), ("df_taud", C.c_double),
                ("b0_a", C.c_double), ("b0_xc", C.c_double), ("b0_sref", C.c_double),
                ("b0_c", C.c_double * 6), ("bnd_numax", C.c_double)]


_L = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _L
    if _L is None:
        if not LIB.exists():
            build()
        _L = C.CDLL(str(LIB))
    return _L


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def grid(nx, ny, dx, dy) -> OrcGrid:
    return OrcGrid(nx, ny, dx, dy)


def push_params(qm, dt, B0=(0.0, 0.0, 0.0), has_b0=True) -> OrcPush:
    pp = OrcPush(qm, dt)
    for i in range(3):
        pp.B0[i] = B0[i]
    pp.has_b0 = int(has_b0)
    return pp


# ---- integer layer -------------------------------------------------------------

def cell_index(g, x, y):
    x, y = _f(x), _f(y)
    out = np.empty(x.size, np.int32)
    lib().orc_cell_index(C.byref(g), _p(x), _p(y), C.c_int64(x.size), _p(out))
    return out


def tile_ids(g, tx, ty, x, y):
    x, y = _f(x), _f(y)
    out = np.empty(x.size, np.uint32)
    lib().orc_tile_ids(C.byref(g), tx, ty, _p(x), _p(y), C.c_int64(x.size), _p(out))
    return out


def cell_keys(g, tx, ty, x, y):
    x, y = _f(x), _f(y)
    out = np.empty(x.size, np.uint32)
    lib().orc_cell_keys(C.byref(g), tx, ty, _p(x), _p(y), C.c_int64(x.size), _p(out))
    return out


def counting_sort(keys, nkeys):
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    perm = np.empty(k.size, np.int32)
    off = np.empty(nkeys + 1, np.int32)
    lib().orc_counting_sort(_p(k), C.c_int64(k.size), C.c_uint32(nkeys), _p(perm), _p(off))
    return perm, off


def migrate(g, x, y):
    x, y = _f(x).copy(), _f(y).copy()
    cell = np.empty(x.size, np.int32)
    lib().orc_migrate_f32(C.byref(g), _p(x), _p(y), _p(cell), C.c_int64(x.size))
    return x, y, cell


# ---- deposits ------------------------------------------------------------------

def deposit_rho_f64(g, x, y, qm, w=None, w_uniform=0.0):
    x, y = _f(x), _f(y)
    w = _f(w) if w is not None else None
    rho = np.zeros(g.nx * g.ny, np.float64)
    lib().orc_deposit_rho_f64(C.byref(g), _p(x), _p(y), _p(w), C.c_double(w_uniform),
                              C.c_int64(x.size), C.c_double(qm), _p(rho))
    return rho.reshape(g.ny, g.nx)


def deposit_rho_fix(g, x, y, qm, scale_log2, w=None, w_uniform=0.0):
    x, y = _f(x), _f(y)
    w = _f(w) if w is not None else None
    acc = np.zeros(g.nx * g.ny, np.int64)
    lib().orc_deposit_rho_fix(C.byref(g), _p(x), _p(y), _p(w), C.c_double(w_uniform),
                              C.c_int64(x.size), C.c_double(qm), C.c_int32(scale_log2), _p(acc))
    out = np.empty(g.nx * g.ny, np.float32)
    lib().orc_fix_to_float(_p(acc), C.c_int64(acc.size), C.c_int32(scale_log2), _p(out))
    return out.reshape(g.ny, g.nx), acc.reshape(g.ny, g.nx)


# ---- push ------------------------------------------------------------------------

def push_es_f32(g, pp, state, Ex, Ey, wrap):
    s = {k: _f(state[k]).copy() for k in ("x", "y", "ux", "uy", "uz")}
    Ex, Ey = _f(Ex).ravel(), _f(Ey).ravel()
    lib().orc_push_es_f32(C.byref(g), C.byref(pp), *(_p(s[k]) for k in ("x", "y", "ux", "uy", "uz")),
                          C.c_int64(s["x"].size), _p(Ex), _p(Ey), C.c_int32(int(wrap)))
    return s


def push_es_f64(g, pp, state, Ex, Ey, wrap):
    s = {k: _d(state[k]).copy() for k in ("x", "y", "ux", "uy", "uz")}
    Ex, Ey = _d(Ex).ravel(), _d(Ey).ravel()
    lib().orc_push_es_f64(C.byref(g), C.byref(pp), *(_p(s[k]) for k in ("x", "y", "ux", "uy", "uz")),
                          C.c_int64(s["x"].size), _p(Ex), _p(Ey), C.c_int32(int(wrap)))
    return s


def half_back_f32(g, pp, state, Ex, Ey):
    s = {k: _f(state[k]).copy() for k in ("x", "y", "ux", "uy", "uz")}
    Ex, Ey = _f(Ex).ravel(), _f(Ey).ravel()
    lib().orc_half_back_f32(C.byref(g), C.byref(pp), _p(s["x"]), _p(s["y"]), _p(s["ux"]),
                            _p(s["uy"]), _p(s["uz"]), C.c_int64(s["x"].size), _p(Ex), _p(Ey))
    return s


def push_em_f32(g, qm, dt, state, E3, B3):
    s = {k: _f(state[k]).copy() for k in ("x", "y", "ux", "uy", "uz")}
    f = [_f(a).ravel() for a in (*E3, *B3)]
    lib().orc_push_em_f32(C.byref(g), C.c_double(qm), C.c_double(dt),
                          *(_p(s[k]) for k in ("x", "y", "ux", "uy", "uz")),
                          C.c_int64(s["x"].size), *(_p(a) for a in f))
    return s


def push_em_f64(g, qm, dt, state, E3, B3, wrap=True):
    s = {k: _d(state[k]).copy() for k in ("x", "y", "ux", "uy", "uz")}
    f = [_d(a).ravel() for a in (*E3, *B3)]
    lib().orc_push_em_f64(C.byref(g), C.c_double(qm), C.c_double(dt),
                          *(_p(s[k]) for k in ("x", "y", "ux", "uy", "uz")),
                          C.c_int64(s["x"].size), *(_p(a) for a in f), C.c_int32(int(wrap)))
    return s


def deposit_rho_j_f64(g, state, qm, w=None, w_uniform=0.0):
    x, y, ux, uy, uz = (_f(state[k]) for k in ("x", "y", "ux", "uy", "uz"))
    w = _f(w) if w is not None else None
    out = [np.zeros(g.nx * g.ny) for _ in range(4)]
    lib().orc_deposit_rho_j_f64(C.byref(g), _p(x), _p(y), _p(ux), _p(uy), _p(uz), _p(w),
                                C.c_double(w_uniform), C.c_int64(x.size), C.c_double(qm),
                                *(_p(a) for a in out))
    return [a.reshape(g.ny, g.nx) for a in out]


def deposit_amu_f64(g, state, qm, w=None, w_uniform=0.0):
    x, y, ux, uy, uz = (_f(state[k]) for k in ("x", "y", "ux", "uy", "uz"))
    w = _f(w) if w is not None else None
    out = [np.zeros(g.nx * g.ny) for _ in range(4)]
    lib().orc_deposit_amu_f64(C.byref(g), _p(x), _p(y), _p(ux), _p(uy), _p(uz), _p(w),
                              C.c_double(w_uniform), C.c_int64(x.size), C.c_double(qm),
                              *(_p(a) for a in out))
    return [a.reshape(g.ny, g.nx) for a in out]


def deposit_dcu_f64(g, state, qm, E3, B3, w=None, w_uniform=0.0):
    x, y, ux, uy, uz = (_f(state[k]) for k in ("x", "y", "ux", "uy", "uz"))
    w = _f(w) if w is not None else None
    f = [_f(a).ravel() for a in (*E3, *B3)]
    out = [np.zeros(g.nx * g.ny) for _ in range(3)]
    lib().orc_deposit_dcu_f64(C.byref(g), _p(x), _p(y), _p(ux), _p(uy), _p(uz), _p(w),
                              C.c_double(w_uniform), C.c_int64(x.size), C.c_double(qm),
                              *(_p(a) for a in f), *(_p(a) for a in out))
    return [a.reshape(g.ny, g.nx) for a in out]


def esirkepov_f64(g, x0, y0, x1, y1, uz, qm, dt, w=None, w_uniform=0.0):
    """Charge-conserving current of moves (x0, y0) -> (x1, y1), |step| < 1 cell:
    staggered Jx (i+1/2, j), Jy (i, j+1/2), nodal Jz, each [ny][nx]."""
    x0, y0, x1, y1, uz = (_f(a) for a in (x0, y0, x1, y1, uz))
    w = _f(w) if w is not None else None
    out = [np.zeros(g.nx * g.ny) for _ in range(3)]
    lib().orc_esirkepov_f64(C.byref(g), _p(x0), _p(y0), _p(x1), _p(y1), _p(uz), _p(w),
                            C.c_double(w_uniform), C.c_int64(x0.size), C.c_double(qm),
                            C.c_double(dt), *(_p(a) for a in out))
    return [a.reshape(g.ny, g.nx) for a in out]


def rho_nodes_f64(g, x, y, qm, w=None, w_uniform=0.0):
    x, y = _f(x), _f(y)
    w = _f(w) if w is not None else None
    rho = np.zeros(g.nx * g.ny, np.float64)
    lib().orc_rho_nodes_f64(C.byref(g), _p(x), _p(y), _p(w), C.c_double(w_uniform),
                            C.c_int64(x.size), C.c_double(qm), _p(rho))
    return rho.reshape(g.ny, g.nx)


def solve_darwin_b_f64(g, jx, jy, jz, eps0=1.0, c=1.0, staggered=False):
    """B = curl^-1 of mu0 J (no B0).  staggered: jx at (i+1/2, j), jy at
    (i, j+1/2), as the Esirkepov deposit writes them."""
    out = [np.empty((g.ny, g.nx)) for _ in range(3)]
    lib().orc_solve_darwin_b_f64(C.byref(g), C.c_double(eps0), C.c_double(c), _p(_d(jx)),
                                 _p(_d(jy)), _p(_d(jz)), *(_p(a) for a in out),
                                 C.c_int32(1 if staggered else 0))
    return out


def solve_darwin_et_f64(g, dcu3, amu4, eps0=1.0, c=1.0, n0=1.0):
    out = [np.empty((g.ny, g.nx)) for _ in range(3)]
    src = [_d(a) for a in (*dcu3, *amu4)]
    lib().orc_solve_darwin_et_f64(C.byref(g), C.c_double(eps0), C.c_double(c), C.c_double(n0),
                                  *(_p(a) for a in src), *(_p(a) for a in out))
    return out


# ---- spectral ----------------------------------------------------------------------

def fft2_r2c(field):
    f = _d(field)
    ny, nx = f.shape
    out = np.empty((ny, nx // 2 + 1, 2), np.float64)
    lib().orc_fft2_r2c(nx, ny, _p(f), _p(out))
    return out[..., 0] + 1j * out[..., 1]


def fft2_c2r(spec, nx):
    ny = spec.shape[0]
    s = np.ascontiguousarray(np.stack([spec.real, spec.imag], axis=-1), dtype=np.float64)
    out = np.empty((ny, nx), np.float64)
    lib().orc_fft2_c2r(nx, ny, _p(s), _p(out))
    return out


def dft2_brute(field):
    f = _d(field)
    ny, nx = f.shape
    out = np.empty((ny, nx // 2 + 1, 2), np.float64)
    lib().orc_dft2_r2c_brute(nx, ny, _p(f), _p(out))
    return out[..., 0] + 1j * out[..., 1]


def solve_es_f64(g, rho, eps0=1.0):
    r = _d(rho)
    ex = np.empty((g.ny, g.nx), np.float64)
    ey = np.empty((g.ny, g.nx), np.float64)
    lib().orc_solve_es_f64(C.byref(g), C.c_double(eps0), _p(r), _p(ex), _p(ey))
    return ex, ey


# ---- loaders / diagnostics ---------------------------------------------------------

def erfinv(y):
    lib().orc_erfinv.restype = C.c_double
    return lib().orc_erfinv(C.c_double(y))


def quiet_u(t, base, uth=1.0, ufl=0.0):
    lib().orc_quiet_u.restype = C.c_float
    lib().orc_quiet_u.argtypes = [C.c_int64, C.c_uint32, C.c_double, C.c_double]
    return lib().orc_quiet_u(t, base, uth, ufl)


def apply_modes(g, x, mode_amp):
    """x += sum_m A_m sin(m k1 x) with A_m physical (divided by dx into cells),
    then the periodic wrap; float32 in, float32 out."""
    x = _f(x).copy()
    a = _d(np.asarray(mode_amp, np.float64) / g.dx)
    lib().orc_apply_modes_f32(C.byref(g), _p(x), C.c_int64(x.size), _p(a), C.c_int32(a.size))
    return x


def load_species(g, specs, noisy, seed, mode_amp=None):
    """specs: list of (density, ppc, uth3, ufl3).  Same draw + rank-major storage as
    the library's loader; mode_amp (physical amplitudes of modes 1..M) displaces x."""
    nc = g.nx * g.ny
    n = sum(p * nc for _, p, _, _ in specs)
    arrs = {k: np.empty(n, np.float32) for k in ("x", "y", "ux", "uy", "uz")}
    base, weights = 0, []
    for idx, (dens, ppc, uth, ufl) in enumerate(specs):
        si = OrcSpecies(ppc, idx, base, ppc * nc, dens * g.dx * g.dy / ppc)
        for i in range(3):
            si.uth[i], si.ufl[i] = uth[i], ufl[i]
        si.noisy, si.seed = int(noisy), seed
        lib().orc_load_species(C.byref(g), C.byref(si), *(_p(arrs[k]) for k in
                                                           ("x", "y", "ux", "uy", "uz")))
        weights.append(si.weight)
        base += ppc * nc
    if mode_amp is not None and len(mode_amp):
        arrs["x"] = apply_modes(g, arrs["x"], mode_amp)
    return arrs, weights


def phase_hist(g, state, nxb, nvb, bx, bz, vth, vlo, vhi, fp64):
    """uint32 counts [nvb, nxb] of f(x, v_par/v_th) (see orc_phase_hist)."""
    x, ux, uz = _f(state["x"]), _f(state["ux"]), _f(state["uz"])
    out = np.zeros(nvb * nxb, np.uint32)
    lib().orc_phase_hist(C.byref(g), _p(x), _p(ux), _p(uz), C.c_int64(x.size), C.c_int32(nxb),
                         C.c_int32(nvb), C.c_double(bx), C.c_double(bz), C.c_double(vth),
                         C.c_double(vlo), C.c_double(vhi), C.c_int32(int(fp64)), _p(out))
    return out.reshape(nvb, nxb)


def energies(g, state, Ex, Ey, rho, w_uniform, eps0=1.0, w=None):
    out = np.empty(5, np.float64)
    ux, uy, uz = _f(state["ux"]), _f(state["uy"]), _f(state["uz"])
    w = _f(w) if w is not None else None
    lib().orc_energies(C.byref(g), _p(ux), _p(uy), _p(uz), _p(w), C.c_double(w_uniform),
                       C.c_int64(ux.size), _p(_f(Ex).ravel()), _p(_f(Ey).ravel()),
                       _p(_f(rho).ravel()) if rho is not None else None, C.c_double(eps0),
                       _p(out))
    return dict(zip(("ke", "ee", "total", "charge", "max_e"), out))


def es_step_f32(g, pp, state, w_uniform, scale_log2, rho, eps0=1.0, nthreads=1):
    """One FP32 same-algorithm ES step in place on `state` (dict of float32 arrays)
    and `rho` (float32 [ny,nx]); returns (Ex, Ey)."""
    ex = np.empty((g.ny, g.nx), np.float32)
    ey = np.empty((g.ny, g.nx), np.float32)
    lib().orc_es_step_f32(C.byref(g), C.byref(pp), C.c_double(eps0),
                          *(_p(state[k]) for k in ("x", "y", "ux", "uy", "uz")),
                          C.c_int64(state["x"].size), C.c_double(w_uniform),
                          C.c_int32(scale_log2), _p(rho), _p(ex), _p(ey), C.c_int32(nthreads))
    return ex, ey


def max_threads() -> int:
    return lib().orc_max_threads()


def es_adjoint_f64(g, state, w_uniform, nsteps, qm=-1.0, dt=0.1, eps0=1.0, grad=True,
                   nthreads=None, b0z=0.0):
    """FP64 electrostatic run (B0 = b0z along z) from `state` (x, y in cells, ux, uy at
    t = 0 before the half step back) for nsteps steps.  Returns the loss
    EE(nsteps) and, with grad, dict(x, y, ux, uy) of dL/d(initial state)."""
    x, y, ux, uy = (_d(state[k]).copy() for k in ("x", "y", "ux", "uy"))
    n = x.size
    out = {k: np.zeros(n) for k in ("x", "y", "ux", "uy")} if grad else None
    nt = nthreads or max_threads()
    ptrs = [_p(out[k]) for k in ("x", "y", "ux", "uy")] if grad else [None] * 4
    f = lib().orc_es_adjoint_f64
    f.restype = C.c_double
    loss = f(C.byref(g), C.c_double(qm), C.c_double(dt),
                                    C.c_double(eps0), C.c_double(b0z), C.c_double(w_uniform),
                                    _p(x), _p(y),
                                    _p(ux), _p(uy), C.c_int64(n), C.c_int32(nsteps), *ptrs,
                                    C.c_int32(nt))
    return (loss, out) if grad else loss


# ---- full-Maxwell (Yee) branch -------------------------------------------------

def yee_params(**kw) -> OrcYee:
    """OrcYee from keyword values (RunParams member names); B0 a 3-sequence."""
    p = OrcYee()
    p.df_kappa, p.df_tpar, p.df_tperp, p.pump_toff = 1.0, 1.0, 1.0, 1e30
    for k, v in kw.items():
        if k in ("B0", "b0_c"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


def yee_faraday(g, F, h):
    F = _f(F).copy()
    lib().orc_yee_faraday_f32(C.byref(g), _p(F), C.c_float(h))
    return F


def yee_ampere(g, F, c2, dt):
    F = _f(F).copy()
    lib().orc_yee_ampere_f32(C.byref(g), _p(F), C.c_float(c2), C.c_float(dt))
    return F


def yee_binomial(g, a):
    a = _f(a).ravel()
    out = np.empty_like(a)
    lib().orc_yee_binomial_f32(C.byref(g), _p(a), _p(out))
    return out


def yee_antenna(g, F, amp, x0, sigma, w0, trmp, toff, t):
    """orc_yee_antenna on a copy of the nine mesh grids F."""
    F = _f(F).copy()
    lib().orc_yee_antenna(C.byref(g), C.c_double(amp), C.c_double(x0), C.c_double(sigma),
                          C.c_double(w0), C.c_double(trmp), C.c_double(toff), C.c_double(t),
                          _p(F))
    return F


def yee_cold_fluid(g, yp, cold_nc, F, vcy, vcz):
    """orc_yee_cold_fluid_f32 on copies; returns (F, vcy, vcz)."""
    F, vcy, vcz = _f(F).copy(), _f(vcy).copy(), _f(vcz).copy()
    lib().orc_yee_cold_fluid_f32(C.byref(g), C.byref(yp), C.c_double(cold_nc), _p(F), _p(vcy),
                                 _p(vcz))
    return F, vcy, vcz


def yee_damp(g, nd, numax, dt, F, vcy=None, vcz=None):
    """orc_yee_damp_f32 on copies; returns (F, vcy, vcz) (None without the fluid)."""
    F = _f(F).copy()
    if vcy is not None:
        vcy, vcz = _f(vcy).copy(), _f(vcz).copy()
    lib().orc_yee_damp_f32(C.byref(g), int(nd), C.c_double(numax), C.c_double(dt), _p(F),
                           _p(vcy) if vcy is not None else None,
                           _p(vcz) if vcz is not None else None)
    return F, vcy, vcz


def yee_push(g, yp, tnow, F, state, w=None, w_uniform=0.0, wd=None, wc=None, wdd=None):
    """One Yee push (orc_yee_push_f32) of copies of the state; returns
    (new state, x0, y0, vz, wdep)."""
    st = {k: _f(state[k]).copy() for k in ("x", "y", "ux", "uy", "uz")}
    n = st["x"].size
    wd = _f(wd).copy() if wd is not None else None
    wc = _f(wc).copy() if wc is not None else None
    wdd = _d(wdd).copy() if wdd is not None else None
    x0, y0, vz = (np.empty(n, np.float32) for _ in range(3))
    wdep = np.empty(n, np.float32) if yp.deltaf else None
    lib().orc_yee_push_f32(C.byref(g), C.byref(yp), C.c_double(tnow), _p(_f(F)),
                           *(_p(st[k]) for k in ("x", "y", "ux", "uy", "uz")),
                           _p(_f(w) if w is not None else None), C.c_float(w_uniform),
                           _p(wd), _p(wc), _p(wdd), _p(x0), _p(y0), _p(vz), _p(wdep),
                           C.c_int64(n))
    if wd is not None:
        st["wd"] = wd
    if wc is not None:
        st["wc"] = wc
    if wdd is not None:
        st["wdd"] = wdd
    return st, x0, y0, vz, wdep


def yee_energy(g, F, c2):
    out = np.empty(2, np.float64)
    lib().orc_yee_energy_f64(C.byref(g), _p(_f(F)), C.c_double(c2), _p(out))
    return out


def yee_divb_rms(g, F):
    lib().orc_yee_divb_rms.restype = C.c_double
    return lib().orc_yee_divb_rms(C.byref(g), _p(_f(F)))


def yee_div_e(g, F):
    out = np.empty(g.nx * g.ny, np.float64)
    lib().orc_yee_div_e(C.byref(g), _p(_f(F)), _p(out))
    return out.reshape(g.ny, g.nx)


def b0_profile(yp, xph):
    lib().orc_b0_profile.restype = C.c_float
    return lib().orc_b0_profile(C.byref(yp), C.c_float(xph))


def mirror_offsets(g, yp, ppc, uth_par, uth_perp, dist=0, lc_rho=0.0, lc_kappa=0.5):
    off = np.empty(g.nx + 1, np.int32)
    lib().orc_mirror_offsets.restype = C.c_int64
    lib().orc_mirror_offsets(C.byref(g), C.byref(yp), C.c_int32(ppc), C.c_double(uth_par),
                             C.c_double(uth_perp), C.c_int32(dist), C.c_double(lc_rho),
                             C.c_double(lc_kappa), _p(off))
    return off
