"""Named workloads shared by the bench and the tests: the Darwin
whistler-pump scan run and the differentiable two-stream setup."""
from __future__ import annotations

import math

from . import _lib as L


def whistler_run(vr_over_vth: float, nx: int = 512, ny: int = 128, ppc: int = 256,
                 dump_every: int = 100):
    """One run of the Darwin whistler-pump scan: 512x128, dx 0.5, c 10,
    w_ce/w_pe 0.5 with B0 at 30 deg to x in the x-z plane, v_r = 0.04c = 0.4,
    v_th = v_r / (v_r/v_th), pump w0 = 0.3 w_ce = 0.15 at mode M = 18
    (k0 = 2 pi 18 / 256) with dB_pump = 1e-3 B0 converted to the pump's E
    amplitude by Faraday (|E_perp| = (w0/k0) dB), circular (ey, ez = +-E),
    dt 0.05, noisy Maxwellian load.  Returns (grid, run params, species)."""
    dx = 0.5
    wce, th = 0.5, math.radians(30.0)
    B0 = (wce * math.cos(th), 0.0, wce * math.sin(th))
    vr = 0.4
    vth = vr / vr_over_vth
    k0 = 2 * math.pi * 18 / (nx * dx)
    w0 = 0.3 * wce
    e_perp = (w0 / k0) * 1e-3 * wce
    g = L.grid(nx, ny, nx * dx, ny * dx)
    rp = L.run_params(dt=0.05, qm=-1.0, eps0=1.0, c=10.0, n0=1.0, noisy_load=1,
                      rng_seed=20260631, B0=B0, wce=wce, dump_every=dump_every)
    rp.pump = 1
    rp.pump_ex, rp.pump_ey, rp.pump_ez = 0.0, e_perp, -e_perp
    rp.pump_k0, rp.pump_w0 = k0, w0
    rp.pump_trmp, rp.pump_toff = 2 * math.pi / w0, 1e30
    sp = [L.species("electrons", 1.0, ppc, (vth,) * 3, (0.0, 0.0, 0.0))]
    return g, rp, sp


WHISTLER_RATIOS = (3.2, 2.1, 1.0)
