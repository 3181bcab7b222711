"""Deck loader (aw_deck_load, host only) against the reference's decks.

The fixtures under tests/golden/decks are input decks from the reference's
decks/ folder (data files).  Expected values are read off the deck text and
the documented finalize rules of load_deck (include/pic/deck.hpp:238-272):
c = (Lx/nx)/dx_wpe_c, B0 = wce (cos th, 0, sin th), pump amplitudes = Table I
x amp x dx/1e4, k0 = 2 pi M / Lx.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from arcwarden_amd import _lib as L
from arcwarden_amd.deck import Deck, load_deck

DECKS = Path(__file__).resolve().parent / "golden" / "decks"


def test_two_stream_deck():
    d = load_deck(DECKS / "two_stream.ini")
    assert (d.nx, d.ny, d.Lx, d.Ly) == (64, 8, 10.2633, 1.0)
    assert d.rp.dt == 0.05 and d.rp.nsteps == 600 and d.dump_every == 10
    assert d.rp.dump_every == 0  # the runner's cadence lives in Deck, not RunParams
    assert d.rp.qm == -1.0 and d.rp.eps0 == 1.0 and d.rp.wpe == 1.0
    assert d.rp.noisy_load == 1 and d.rp.rng_seed == 20260627
    assert [s.name.decode() for s in d.species] == ["beam_plus", "beam_minus"]
    for s, sign in zip(d.species, (1.0, -1.0)):
        assert s.density == 0.5 and s.ppc == 2048
        assert list(s.uth) == [0.1, 0.1, 0.1] and list(s.ufl) == [sign, 0.0, 0.0]
    assert not d.darwin and not d.yee and d.rp.pump == 0
    assert d.outdir == "deck_frames"
    assert d.rp.c == 1.0 and list(d.rp.B0) == [0.0, 0.0, 0.0]


@pytest.mark.parametrize("name", ["bump_on_tail", "profile"])
def test_bump_on_tail_decks(name):
    d = load_deck(DECKS / f"{name}.ini")
    assert (d.nx, d.ny, d.Lx, d.Ly) == (512, 8, 50.265482, 1.0)
    assert d.rp.dt == 0.05
    assert d.rp.nsteps == (4000 if name == "bump_on_tail" else 400)
    assert d.dump_every == (40 if name == "bump_on_tail" else 0)
    bulk, beam = d.species
    assert (bulk.name.decode(), bulk.density, bulk.ppc) == ("bulk", 0.92, 920)
    assert (beam.name.decode(), beam.density, beam.ppc) == ("beam", 0.08, 80)
    assert list(beam.ufl) == [1.0, 0.0, 0.0] and list(beam.uth) == [0.15, 0.15, 0.15]
    # equal macro-particle weights (density / ppc)
    assert bulk.density / bulk.ppc == pytest.approx(beam.density / beam.ppc)


@pytest.mark.parametrize("name,nx,Lx,dx_wpe_c,mode,w0,ex0,toff,tsnap,band", [
    ("an2019_sim1", 4096, 8192.0, 0.0267, 10, 0.0215, 1.39, 1100.0, 1540.0, (300, 400)),
    ("an2019_sim2", 2048, 4096.0, 0.04, 7, 0.0194, 1.35, 750.0, 2500.0, (20, 250)),
])
def test_an2019_darwin_pump_decks(name, nx, Lx, dx_wpe_c, mode, w0, ex0, toff, tsnap, band):
    d = load_deck(DECKS / f"{name}.ini")
    dx = Lx / nx
    assert (d.nx, d.ny, d.Lx, d.Ly) == (nx, 1, Lx, 2.0)
    assert d.darwin and not d.yee and d.rp.ndc == 2
    assert d.rp.c == pytest.approx(dx / dx_wpe_c, rel=1e-15)
    th = math.radians(30.0)
    assert list(d.rp.B0) == [np.float32(0.1 * math.cos(th)), 0.0, np.float32(0.1 * math.sin(th))]
    assert d.rp.wce == 0.1
    s = 6.0 * dx / 1e4
    assert d.rp.pump == 1
    assert d.rp.pump_k0 == pytest.approx(2 * math.pi * mode / Lx, rel=1e-15)
    assert d.rp.pump_w0 == w0
    assert d.rp.pump_ex == pytest.approx(ex0 * s, rel=1e-15)
    assert d.rp.pump_ey == pytest.approx(1.81 * s, rel=1e-15)
    assert d.rp.pump_ez == pytest.approx(-1.81 * s, rel=1e-15)
    assert d.rp.pump_trmp == 62.83 and d.rp.pump_toff == toff
    assert d.diag_enable == ["spectrum", "kt", "phase_video", "snapshot"]
    assert d.tsnap == tsnap and (d.band_lo, d.band_hi) == band
    assert d.prefix == name
    assert d.kt_stride == 5 and d.n_frames == 150  # defaults
    (e,) = d.species
    assert (e.name.decode(), e.density, e.ppc) == ("electrons", 1.0, 2000)
    assert list(e.uth) == [1.0, 1.0, 1.0]


def _write(tmp_path, text):
    p = tmp_path / "d.ini"
    p.write_text(text)
    return p


GRID = "[grid]\nnx = 8\nny = 4\nLx = 8\nLy = 4\n"
SPECIES = "[species e]\nppc = 4\n"


@pytest.mark.parametrize("text,msg", [
    (GRID, "at least one [species"),
    (SPECIES, "[grid] nx/ny/Lx/Ly"),
    (GRID + "[species\n", "bad section at line 6"),
    (GRID + SPECIES + "density 2\n", "expected key=value at line 8"),
    (GRID + SPECIES + "dist = maxwell\n", "dist must be bimax|losscone"),
    (GRID + SPECIES + "[background]\nprofile = ramp\n", "profile must be"),
    (GRID + SPECIES + "[background]\nprofile = parabolic\ntheta_deg = 30\nwce = 1\n",
     "requires B0 along x"),
    (GRID + SPECIES + "[background]\nprofile = dipole\n", "requires lre"),
])
def test_deck_errors(tmp_path, text, msg):
    with pytest.raises(L.AwError, match=None) as ei:
        load_deck(_write(tmp_path, text))
    assert msg in str(ei.value)


def test_deck_missing_file(tmp_path):
    with pytest.raises(L.AwError) as ei:
        load_deck(tmp_path / "nope.ini")
    assert "deck: cannot open" in str(ei.value)


def test_deck_comments_defaults_and_forward_compat(tmp_path):
    text = ("! header comment\n" + GRID +
            "[time]\ndt = 0.1 # trailing comment\nnsteps = 1e3\n"
            "[unknown]\nwhatever = 1\n"
            "[plasma]\nseed = 7 ! bang comment\nnoisy = .true.\n"
            "[species]\nppc = 3\nuth = 0.5\n"
            "[field]\nmodel = darwin\nc = 12.5\ndx_wpe_c = 0.1\n"
            "[background]\nB0 = 0 0 0.25\nwce = 3\ntheta_deg = 45\n"
            "[background]\nprofile = parabolic\n")
    with pytest.raises(L.AwError):  # parabolic profile needs B0 along x (B0 given in z)
        load_deck(_write(tmp_path, text))
    d = load_deck(_write(tmp_path, text.replace("profile = parabolic", "a = 0.5")))
    assert d.rp.dt == 0.1 and d.rp.nsteps == 1  # std::stol("1e3") reads the leading 1
    assert d.rp.rng_seed == 7 and d.rp.noisy_load == 1
    (s,) = d.species
    assert s.name.decode() == "electrons" and s.ppc == 3 and list(s.uth) == [0.5, 0.0, 0.0]
    assert s.density == 1.0 and s.lc_kappa == 0.3
    assert d.darwin and d.rp.c == 12.5  # a direct c wins over dx_wpe_c
    assert list(d.rp.B0) == [0.0, 0.0, 0.25]  # B0 given verbatim: wce/theta ignored
    assert d.rp.wce == 3.0 and d.rp.b0_a == 0.5


def test_deck_profile_centre(tmp_path):
    text = GRID + SPECIES + "[background]\nwce = 2\nprofile = parabolic\na = 0.1\n"
    d = load_deck(_write(tmp_path, text))
    assert d.rp.b0_prof == 1 and d.rp.b0_xc == 4.0  # defaults to Lx / 2
    assert list(d.rp.B0) == [2.0, 0.0, 0.0]
    assert isinstance(d, Deck)


def test_deck_dipole_profile_fit(tmp_path):
    """profile = dipole fits the even polynomial of B/B_eq along a dipole field
    line (include/pic/background_b0.hpp:57-146): s_ref = 1.01 max(xc, Lx - xc),
    and the fit reproduces the closed form -- B/B_eq = sqrt(1 + 3 sin^2 l)/cos^6 l
    at arc length s = L R_E (sin l sqrt(1 + 3 sin^2 l) + asinh(sqrt(3) sin l)/sqrt(3))/2
    -- within 2e-4 over the whole span."""
    lre = 40.0
    text = GRID + SPECIES + f"[background]\nwce = 1\nprofile = dipole\nlre = {lre}\n"
    d = load_deck(_write(tmp_path, text))
    assert d.rp.b0_prof == 2
    assert d.rp.b0_sref == pytest.approx(1.01 * max(d.rp.b0_xc, d.Lx - d.rp.b0_xc), rel=1e-15)
    lam = np.linspace(0.0, 0.9, 2000)
    t = np.sin(lam)
    s = lre * 0.5 * (t * np.sqrt(1 + 3 * t * t) + np.arcsinh(np.sqrt(3) * t) / np.sqrt(3))
    b = np.sqrt(1 + 3 * t * t) / np.cos(lam) ** 6
    keep = s <= d.rp.b0_sref
    u = (s[keep] / d.rp.b0_sref) ** 2
    poly = np.polyval(list(d.rp.b0_c)[::-1], u)
    assert np.max(np.abs(poly - b[keep]) / b[keep]) < 2e-4
    assert d.rp.b0_c[0] == pytest.approx(1.0, abs=2e-4)
    # the profile needs lre
    bad = GRID + SPECIES + "[background]\nwce = 1\nprofile = dipole\n"
    with pytest.raises(Exception, match="lre"):
        load_deck(_write(tmp_path, bad))
