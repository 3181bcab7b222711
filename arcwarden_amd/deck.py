"""Input decks: ``load_deck`` over aw_deck_load (the library's C++ parser of
the reference's INI decks, include/pic/deck.hpp:31-274) and a Simulation
built from a deck, as tools/run_deck.cu does."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from os import PathLike

from . import _lib as L


@dataclass
class Deck:
    """Deck (include/pic/deck.hpp:31-78).  ``rp`` is the finalized RunParams
    (AwRunParams), ``species`` a list of AwSpecies."""
    nx: int
    ny: int
    Lx: float
    Ly: float
    rp: L.AwRunParams
    species: list
    dump_every: int
    outdir: str
    darwin: bool
    yee: bool
    dx_wpe_c: float
    c_direct: float
    wce: float
    theta_deg: float
    bnoise: float
    pump_enable: bool
    pump_M: int
    diag_enable: list
    probes: list
    tsnap: float
    band_lo: int
    band_hi: int
    kt_stride: int
    n_frames: int
    prefix: str

    def grid(self) -> L.AwGrid:
        return L.grid(self.nx, self.ny, self.Lx, self.Ly)


def load_deck(path: str | PathLike) -> Deck:
    d = L.AwDeck()
    L.call("aw_deck_load", str(path).encode(), C.byref(d))
    species = []
    for k in range(d.nspecies):
        s = L.AwSpecies()
        C.memmove(C.byref(s), C.byref(d.species[k]), C.sizeof(s))
        species.append(s)
    rp = L.AwRunParams()
    C.memmove(C.byref(rp), C.byref(d.rp), C.sizeof(rp))
    return Deck(nx=d.nx, ny=d.ny, Lx=d.Lx, Ly=d.Ly, rp=rp, species=species,
                dump_every=d.dump_every, outdir=d.outdir.decode(), darwin=bool(d.darwin),
                yee=bool(d.yee), dx_wpe_c=d.dx_wpe_c, c_direct=d.c_direct, wce=d.wce,
                theta_deg=d.theta_deg, bnoise=d.bnoise, pump_enable=bool(d.pump_enable),
                pump_M=d.pump_M, diag_enable=d.diag_enable.decode().split(),
                probes=list(d.probes[:d.nprobes]), tsnap=d.tsnap, band_lo=d.band_lo,
                band_hi=d.band_hi, kt_stride=d.kt_stride, n_frames=d.n_frames,
                prefix=d.prefix.decode())


def simulation_from_deck(deck: Deck, config=None):
    """Simulation for a deck: the Darwin model when ``[field] model = darwin``
    (run_deck's --em), electrostatic otherwise; rp.dump_every from the deck's
    frame cadence."""
    from .sim import SimConfig, Simulation
    if deck.yee:
        raise L.AwError(5, "deck: [field] model = yee is outside this build (ES and Darwin only)")
    rp = L.AwRunParams()
    C.memmove(C.byref(rp), C.byref(deck.rp), C.sizeof(rp))
    rp.dump_every = deck.dump_every if deck.dump_every > 0 else (1 << 30)
    if config is None:
        b0 = any(v != 0.0 for v in rp.B0)
        config = SimConfig(has_b0=b0,
                           field_model=L.AW_MODEL_DARWIN if deck.darwin else L.AW_MODEL_ES)
    return Simulation(deck.grid(), rp, deck.species, config)
