"""Closed-form dispersion roots and time-series estimators shared by the CPU
(oracle) and GPU physics tests.

Reference tests these follow: tests/test_langmuir.cu:1-12,58-73 (cold
Langmuir, omega = omega_pe), tests/test_two_stream.cu:3-7,57-75 (cold
symmetric two-stream), tests/test_weibel.cu:9-11 (cold Weibel).  The
north-star gates are tighter than the reference's tolerances, so the fits
below are pinned to the analytic roots, not to the reference's bands.
"""
from __future__ import annotations

import math

import numpy as np


def two_stream_growth(k: float, v0: float, wp2_total: float = 1.0) -> float:
    """Cold symmetric two-stream (each beam omega_p^2 = wp2_total/2):
    omega^2 = a + w/2 - sqrt(2 a w + w^2/4) with a = (k v0)^2 (w = wp2_total);
    returns gamma = sqrt(-omega^2) (0 if stable)."""
    a = (k * v0) ** 2
    w = wp2_total
    om2 = a + 0.5 * w - math.sqrt(2.0 * a * w + 0.25 * w * w)
    return math.sqrt(-om2) if om2 < 0 else 0.0


def langmuir_leapfrog_omega(dt: float, wpe: float = 1.0) -> float:
    """Discrete leapfrog oscillator frequency: sin(omega dt/2) = wpe dt/2."""
    return 2.0 / dt * math.asin(0.5 * wpe * dt)


def weibel_growth(k: float, v0: float, c: float) -> float:
    """Cold counter-streaming filamentation: gamma = k v0 / sqrt(1 + k^2 c^2)."""
    return k * v0 / math.sqrt(1.0 + (k * c) ** 2)


def weibel_growth_warm(k: float, v0: float, c: float, vth: float) -> float:
    """Filamentation of beams with thermal spread vth, fluid closure (the
    beams' pressure along k opposes the bunching):
    gamma^2 = k^2 v0^2 / (1 + k^2 c^2) - k^2 vth^2 (0 if stable)."""
    g2 = (k * v0) ** 2 / (1.0 + (k * c) ** 2) - (k * vth) ** 2
    return math.sqrt(g2) if g2 > 0 else 0.0


def peak_times(t: np.ndarray, y: np.ndarray, thresh_frac: float = 0.25) -> np.ndarray:
    """Local maxima above thresh_frac*max, refined by a parabola through the
    three samples around each maximum."""
    y = np.asarray(y, dtype=np.float64)
    thr = thresh_frac * y.max()
    out = []
    for i in range(1, len(y) - 1):
        if y[i] > thr and y[i] >= y[i - 1] and y[i] > y[i + 1]:
            d = y[i - 1] - 2 * y[i] + y[i + 1]
            off = 0.5 * (y[i - 1] - y[i + 1]) / d if d != 0 else 0.0
            out.append(t[i] + off * (t[1] - t[0]))
    return np.array(out)


def langmuir_omega_from_ee(t: np.ndarray, ee: np.ndarray) -> float:
    """EE ~ cos^2(omega t): peaks every pi/omega (tests/test_langmuir.cu:58-73),
    skipping the first transient peak."""
    pk = peak_times(t, ee)
    if len(pk) < 4:
        raise AssertionError(f"only {len(pk)} EE peaks")
    spacing = np.diff(pk[1:]).mean()
    return math.pi / spacing


def growth_rate(t: np.ndarray, energy: np.ndarray, lo_efold: float = 1.0,
                hi_frac: float = 0.8) -> float:
    """gamma from a least-squares fit of ln(energy) = 2 gamma t + c over the
    exponential window: from lo_efold e-folds above the early minimum up to
    hi_frac of the way to the peak (tests/test_weibel.cu:84-107 window)."""
    ln = np.log(np.maximum(np.asarray(energy, np.float64), 1e-300))
    imax = int(np.argmax(ln))
    lnmin = ln[: imax + 1].min()
    lo, hi = lnmin + lo_efold, lnmin + hi_frac * (ln[imax] - lnmin)
    imin = int(np.argmin(ln[: imax + 1]))
    sel = [i for i in range(imin, imax + 1) if lo <= ln[i] <= hi]
    if len(sel) < 5:
        raise AssertionError(f"exponential window too short ({len(sel)} samples)")
    slope = np.polyfit(np.asarray(t)[sel], ln[sel], 1)[0]
    return 0.5 * slope
