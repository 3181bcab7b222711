"""The rho fixed-point scale rule and its overflow guard, on the CPU (a host
function of the library: no GPU needed).  The scale puts |qm| * density at
2^40; enough particles of a large weight lower it so that no int64 cell sum
can wrap; a configuration no scale >= 2^8 can hold is an error, never a
silent wrap."""
import ctypes as C
import math

import pytest

from arcwarden_amd import _lib as L


def _scale(g, qm, density, wmax, n):
    out = C.c_int32(-1)
    st = L.lib().aw_rho_fixed_scale(C.byref(g), C.c_double(qm), C.c_double(density),
                                   C.c_double(wmax), C.c_int64(n), C.byref(out))
    return st, out.value


def test_scale_of_unit_density():
    g = L.grid(512, 512, 256.0, 256.0)
    w = 1.0 * 0.25 / 256  # density 1, 256 ppc, dx = dy = 0.5
    st, s = _scale(g, -1.0, 1.0, w, 512 * 512 * 256)
    assert st == L.AW_OK and s == 40
    # every particle in one cell: n |qm| w / (dx dy) 2^40 = 2^26 * 2^40 / 2^8 < 2^62
    assert 512 * 512 * 256 * w / 0.25 * 2.0 ** 40 < 2.0 ** 62


def test_scale_follows_density_and_charge():
    g = L.grid(64, 64, 32.0, 32.0)
    assert _scale(g, -1.0, 8.0, 1e-3, 1000) == (L.AW_OK, 37)
    assert _scale(g, 2.0, 0.25, 1e-3, 1000) == (L.AW_OK, 41)


def test_many_heavy_particles_lower_the_scale():
    g = L.grid(64, 64, 32.0, 32.0)
    n, w = 10 ** 9, 1.0
    st, s = _scale(g, -1.0, 1.0, w, n)
    assert st == L.AW_OK and s < 40
    total = n * w / 0.25
    assert total * 2.0 ** s < 2.0 ** 62 <= total * 2.0 ** (s + 1)


def test_unholdable_configuration_is_an_error():
    g = L.grid(64, 64, 32.0, 32.0)
    st, _ = _scale(g, -1.0, 1.0, 1e30, 10 ** 9)
    assert st == L.AW_ERR_INVALID
    assert "cannot wrap" in L.last_error()
    st, _ = _scale(g, -1.0, 1.0, math.inf, 10)
    assert st == L.AW_ERR_INVALID
