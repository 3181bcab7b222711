// pic/deck.hpp -- text input decks.
//
// Reference: Deck / load_deck (include/pic/deck.hpp:31-274).  Parsing and the
// finalize step run in the library (aw_deck_load); errors are rethrown as
// std::runtime_error with the reference's messages.
#ifndef ARC_PIC_DECK_HPP
#define ARC_PIC_DECK_HPP

#include "pic/config.hpp"
#include "pic/hip_utils.hpp"
#include "pic/species.hpp"

#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace arc {

struct Deck {
    int nx = 0, ny = 0;
    double Lx = 0.0, Ly = 0.0;
    RunParams rp;
    SpeciesList species;
    long dump_every = 0;
    std::string outdir = "deck_frames";

    bool darwin = false;
    bool yee = false;
    double dx_wpe_c = 0.0;
    double c_direct = 0.0;
    double wce = 0.0;
    double theta_deg = 0.0;
    bool b0_direct = false;
    bool b0_xc_set = false;
    double bnoise = 0.0;

    bool pump_enable = false;
    int pump_M = 0;
    double pump_w0 = 0.0;
    double pump_ex0 = 1.39;
    double pump_ey0 = 1.81;
    double pump_ez0 = -1.81;
    double pump_amp = 1.0;
    double pump_trmp = 0.0;
    double pump_toff = 0.0;

    std::vector<std::string> diag_enable;
    std::vector<double> probes;
    double tsnap = 0.0;
    int band_lo = 0, band_hi = 0;
    int kt_stride = 5;
    int n_frames = 150;
    std::string prefix;
};

inline Deck load_deck(const std::string& path) {
    auto c = std::make_unique<aw_deck>();
    if (aw_deck_load(path.c_str(), c.get()) != AW_OK) throw std::runtime_error(last_error());
    Deck d;
    d.nx = c->nx;
    d.ny = c->ny;
    d.Lx = c->Lx;
    d.Ly = c->Ly;
    static_cast<aw_run_params&>(d.rp) = c->rp;
    for (int k = 0; k < c->nspecies; ++k) d.species.push_back(Species::from_c(c->species[k]));
    d.dump_every = (long)c->dump_every;
    d.outdir = c->outdir;
    d.darwin = c->darwin != 0;
    d.yee = c->yee != 0;
    d.dx_wpe_c = c->dx_wpe_c;
    d.c_direct = c->c_direct;
    d.wce = c->wce;
    d.theta_deg = c->theta_deg;
    d.b0_direct = c->b0_direct != 0;
    d.b0_xc_set = c->b0_xc_set != 0;
    d.bnoise = c->bnoise;
    d.pump_enable = c->pump_enable != 0;
    d.pump_M = c->pump_M;
    d.pump_w0 = c->pump_w0;
    d.pump_ex0 = c->pump_ex0;
    d.pump_ey0 = c->pump_ey0;
    d.pump_ez0 = c->pump_ez0;
    d.pump_amp = c->pump_amp;
    d.pump_trmp = c->pump_trmp;
    d.pump_toff = c->pump_toff;
    std::istringstream mods(c->diag_enable);
    for (std::string m; mods >> m;) d.diag_enable.push_back(m);
    d.probes.assign(c->probes, c->probes + c->nprobes);
    d.tsnap = c->tsnap;
    d.band_lo = c->band_lo;
    d.band_hi = c->band_hi;
    d.kt_stride = c->kt_stride;
    d.n_frames = c->n_frames;
    d.prefix = c->prefix;
    return d;
}

}  // namespace arc

#endif  // ARC_PIC_DECK_HPP
