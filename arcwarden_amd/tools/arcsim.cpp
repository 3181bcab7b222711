// arcsim -- the deck-driven runner: one program for every experiment, with
// the physics and the diagnostics chosen by the deck.
//
// Reference: tools/arcsim.cu:41-223 (command line, output prefix, deck
// overrides, default diagnostics, RUN_META) and tools/run_deck.cu:137-166
// (the older runner's command line).  Built twice from this file:
//
//   arcsim   <deck.ini> [outdir] [--em] [--ppc=N] [--amp=X] [--nsteps=N]
//   run_deck <deck.ini> [outdir] [--em] [--c=C] [--ndc=N]
//
// arcsim writes <outdir>/RUN_META.txt, deck_snapshot.ini, energy.csv and
// whatever [diagnostics] enable = ... selects (default: phase_frames, plus
// em_energy for Darwin and Yee runs), with output files prefixed
// "<outdir>/<deck prefix or stem>_".  run_deck writes energy.csv and the
// phase frames every [time] dump_every steps into outdir (default: the deck's
// outdir), and with --em runs the Darwin model with c from --c (default 20).
// The field model comes from [field] model = electrostatic | darwin | yee
// (--em forces Darwin).
#include "pic/config.hpp"
#include "pic/deck.hpp"
#include "pic/diag/manager.hpp"
#include "pic/grid.hpp"
#include "pic/run_meta.hpp"
#include "pic/simulation.hpp"
#include "pic/simulation_maxwell.hpp"
#include "pic/species.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <filesystem>
#include <fstream>
#include <string>

namespace {

#ifdef ARC_RUNNER_RUN_DECK
constexpr bool kRunDeckCli = true;
#else
constexpr bool kRunDeckCli = false;
#endif

struct Options {
    std::string deck, outdir;
    bool em = false;
    int ppc = 0;          // arcsim: species[0].ppc override
    double amp = 0.0;     // arcsim: pump amplitude override
    long nsteps = 0;      // arcsim: step count override
    double c = 20.0;      // run_deck --em: light speed
    int ndc = 2;          // run_deck --em: transverse iterations (kept in RunParams)
};

bool starts_with(const std::string& s, const char* p) { return s.rfind(p, 0) == 0; }

// Flags may come in any order; the first bare argument is the deck, the
// second the output directory.
bool parse(int argc, char** argv, Options& o) {
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--em") o.em = true;
        else if (!kRunDeckCli && starts_with(a, "--ppc=")) o.ppc = std::atoi(a.c_str() + 6);
        else if (!kRunDeckCli && starts_with(a, "--amp=")) o.amp = std::atof(a.c_str() + 6);
        else if (!kRunDeckCli && starts_with(a, "--nsteps=")) o.nsteps = std::atol(a.c_str() + 9);
        else if (kRunDeckCli && starts_with(a, "--c=")) o.c = std::atof(a.c_str() + 4);
        else if (kRunDeckCli && starts_with(a, "--ndc=")) o.ndc = std::atoi(a.c_str() + 6);
        else if (starts_with(a, "--")) continue;  // unknown flags are ignored
        else if (o.deck.empty()) o.deck = a;
        else if (o.outdir.empty()) o.outdir = a;
    }
    return !o.deck.empty();
}

std::string stem_of(const std::string& path) {
    return std::filesystem::path(path).stem().string();
}

const char* model_name(const arc::Deck& d) {
    return d.yee ? "yee" : (d.darwin ? "darwin" : "electrostatic");
}

// Samples every step n = 0..nsteps (n > 0 after step n - 1), as the
// reference's runner loop does; returns the wall time of the stepping.
template <class Sim>
double drive(Sim& sim, arc::diag::DiagManager& mgr, const arc::Grid& g, const arc::RunParams& rp,
             long nsteps, bool per_step_sample) {
    using clock = std::chrono::steady_clock;
    double stepping = 0.0;
    for (long n = 0; n <= nsteps; ++n) {
        if (n > 0) {
            const auto t0 = clock::now();
            sim.step(n - 1);
            if (!per_step_sample) sim.synchronize();
            stepping += std::chrono::duration<double>(clock::now() - t0).count();
        }
        mgr.sample(sim.fields(), sim.particles(), g, rp, n, n * rp.dt);
    }
    sim.synchronize();
    return stepping;
}

void report(const char* what, long nsteps, std::size_t n, double secs) {
    if (nsteps <= 0 || secs <= 0) return;
    std::printf("%s: %ld steps, %.3f s, %.3f ms/step, %.3g particle-steps/s\n", what, nsteps,
                secs, 1e3 * secs / nsteps, static_cast<double>(n) * nsteps / secs);
}

// Electrostatic and spectral Darwin runs on Simulation<Cfg>.
template <class Cfg>
int run_spectral(arc::Deck& d, const Options& o, const std::string& pref,
                 const std::string& outdir, int argc, char** argv) {
    const arc::Grid g(d.nx, d.ny, d.Lx, d.Ly);
    std::string energy_csv;
    if (kRunDeckCli) {
        if (d.dump_every > 0) {
            std::filesystem::create_directories(outdir);
            energy_csv = outdir + "/energy.csv";
        }
        d.rp.dump_every = d.dump_every > 0 ? d.dump_every : (1L << 30);
    } else {
        if (!outdir.empty()) {
            std::filesystem::create_directories(outdir);
            arc::write_run_meta(outdir, o.deck, argc, argv);
            energy_csv = outdir + "/energy.csv";
        }
        const long every = d.dump_every > 0 ? d.dump_every : std::max(1L, d.rp.nsteps / 100);
        d.rp.dump_every = energy_csv.empty() ? (1L << 30) : every;
    }
    arc::Simulation<Cfg> sim(g, d.rp, d.species, energy_csv);
    sim.init();
    arc::diag::DiagManager mgr(d, d.rp, pref, outdir);
    const double secs = drive(sim, mgr, g, d.rp, d.rp.nsteps, true);
    mgr.finalize();
    const std::size_t n = sim.particles().n;
    report(model_name(d), d.rp.nsteps, n, secs);
    if (kRunDeckCli && Cfg::aw_model == AW_MODEL_DARWIN) {
        const arc::Fields& f = sim.fields();
        const double half_cell = 0.5 * g.dx * g.dy;
        using arc::diag::sum_sq;
        const double we = half_cell * (sum_sq(f.Ex) + sum_sq(f.Ey) + sum_sq(f.Ez));
        const double wb = half_cell * (sum_sq(f.Bx) + sum_sq(f.By) + sum_sq(f.Bz));
        const double wt = half_cell * (sum_sq(f.ETx) + sum_sq(f.ETy) + sum_sq(f.ETz));
        std::printf("field energy at the end: E %.4e, B %.4e (B/E %.2e), E_T %.4e (E_T/E "
                    "%.2e)\n",
                    we, wb, we > 0 ? wb / we : 0.0, wt, we > 0 ? wt / we : 0.0);
    }
    return 0;
}

// The full-Maxwell (Yee) branch (tools/arcsim.cu:69-168): the spectral
// diagnostics read a Fields view of the Yee arrays with B0 and the pump added
// (MaxwellSimulation::diag_fields), and <pref>maxwell.csv logs the field
// energies and the conservation residuals (and the delta-f weight statistics)
// every dump_every steps.
int run_yee(arc::Deck& d, const Options& o, const std::string& pref, const std::string& outdir,
            int argc, char** argv) {
    const arc::Grid g(d.nx, d.ny, d.Lx, d.Ly);
    arc::RunParams& rp = d.rp;
    const double kmax =
        std::sqrt(1.0 / (g.dx * g.dx) + (d.ny > 1 ? 1.0 / (g.dy * g.dy) : 0.0));
    if (rp.c * rp.dt * kmax >= 1.0) {
        std::fprintf(stderr,
                     "arcsim: model=yee CFL violated: c*dt*kmax = %.3f >= 1 (max stable dt = "
                     "%.5g)\n",
                     rp.c * rp.dt * kmax, 1.0 / (rp.c * kmax));
        return 1;
    }
    if (!outdir.empty()) {
        std::filesystem::create_directories(outdir);
        arc::write_run_meta(outdir, o.deck, argc, argv);
    }
    rp.dump_every = d.dump_every > 0 ? d.dump_every : std::max(1L, rp.nsteps / 100);
    bool deltaf = false;
    for (const auto& s : d.species) deltaf = deltaf || s.deltaf;
    if (deltaf) {
        if (d.species.size() != 1 || !d.species[0].deltaf) {
            std::fprintf(stderr, "arcsim: rep=deltaf requires exactly one species\n");
            return 1;
        }
        const arc::Species& s = d.species[0];
        if (rp.B0[1] != 0.f || rp.B0[2] != 0.f || s.uth[1] != s.uth[2]) {
            std::fprintf(stderr, "arcsim: deltaf needs B0 along x and uth_y == uth_z\n");
            return 1;
        }
        rp.deltaf = 1;
        rp.df_tpar = s.uth[0] * s.uth[0];
        rp.df_tperp = s.uth[1] * s.uth[1];
    }
    if (rp.b0_prof && (g.ny != 1 || d.species.size() != 1)) {
        std::fprintf(stderr,
                     "arcsim: [background] profile=parabolic requires ny=1 and exactly one "
                     "species\n");
        return 1;
    }
    arc::MaxwellSimulation sim(g, rp);
    if (rp.b0_prof) sim.particles().initialize_mirror(d.species[0], g, rp, sim.stream());
    else sim.particles().initialize(d.species, g, rp, sim.stream());
    if (rp.deltaf) sim.particles().enable_deltaf(sim.stream(), rp.df_wprec);
    sim.stream().synchronize();
    arc::diag::DiagManager mgr(d, rp, pref, outdir);
    std::ofstream log(pref + "maxwell.csv");
    if (log)
        log << (rp.deltaf ? "step,time,WE,WB,divB_rms,gauss_drift_rms,wd_sum,wd_rms,wd_max\n"
                          : "step,time,WE,WB,divB_rms,gauss_drift_rms\n");
    sim.residuals();  // the Gauss reference at t = 0
    using clock = std::chrono::steady_clock;
    double secs = 0.0;
    for (long n = 0; n <= rp.nsteps; ++n) {
        if (n > 0) {
            const auto t0 = clock::now();
            sim.step();
            secs += std::chrono::duration<double>(clock::now() - t0).count();
        }
        const double t = n * rp.dt;
        mgr.sample(sim.diag_fields(t), sim.particles(), g, rp, n, t);
        if (log && (n % rp.dump_every == 0 || n == rp.nsteps)) {
            const auto e = sim.field_energy();
            const auto r = sim.residuals();
            log << n << ',' << t << ',' << e.we << ',' << e.wb << ',' << r.divb_rms << ','
                << r.gauss_drift_rms;
            if (rp.deltaf) {
                const auto w = sim.wd_stats();
                log << ',' << w.sum << ',' << w.rms << ',' << w.max;
            }
            log << '\n';
        }
    }
    sim.synchronize();
    mgr.finalize();
    report("yee", rp.nsteps, sim.particles().n, secs);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    std::setvbuf(stdout, nullptr, _IOLBF, 0);
    Options o;
    if (!parse(argc, argv, o)) {
        if (kRunDeckCli)
            std::fprintf(stderr, "usage: %s <deck-file> [outdir] [--em] [--c=20] [--ndc=2]\n",
                         argv[0]);
        else
            std::fprintf(stderr,
                         "usage: %s <deck.ini> [outdir] [--em] [--ppc=N] [--amp=X] "
                         "[--nsteps=N]\n",
                         argv[0]);
        return 2;
    }
    arc::Deck d;
    try {
        d = arc::load_deck(o.deck);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "deck error: %s\n", e.what());
        return 1;
    }
    std::string outdir = o.outdir;
    if (kRunDeckCli) {
        if (outdir.empty()) outdir = d.outdir;
        if (o.em) {
            d.darwin = true;
            d.yee = false;
            d.rp.c = o.c;
            d.rp.ndc = o.ndc;
        }
        d.diag_enable.clear();
        if (d.dump_every > 0) d.diag_enable.push_back("phase_frames");
    } else {
        if (o.em) d.darwin = true;
        if (o.ppc > 0 && !d.species.empty()) d.species[0].ppc = o.ppc;
        if (o.nsteps > 0) d.rp.nsteps = o.nsteps;
        if (d.pump_enable) {  // the pump amplitudes follow an --amp override
            const double a = o.amp > 0 ? o.amp : d.pump_amp;
            const double s = a * (d.Lx / d.nx) / 1e4;
            d.rp.pump_ex = d.pump_ex0 * s;
            d.rp.pump_ey = d.pump_ey0 * s;
            d.rp.pump_ez = d.pump_ez0 * s;
        }
        if (d.diag_enable.empty()) {
            d.diag_enable.push_back("phase_frames");
            if (d.darwin || d.yee) d.diag_enable.push_back("em_energy");
        }
    }
    const std::string name = d.prefix.empty() ? stem_of(o.deck) : d.prefix;
    const std::string pref = (outdir.empty() ? std::string() : outdir + "/") + name + "_";
    const long long particles =
        static_cast<long long>(d.nx) * d.ny * (d.species.empty() ? 0 : d.species[0].ppc);
    std::printf("%s: deck %s, %s model, %d x %d cells, %lld particles (species 0), %ld steps, "
                "diagnostics:",
                kRunDeckCli ? "run_deck" : "arcsim", o.deck.c_str(), model_name(d), d.nx, d.ny,
                particles, d.rp.nsteps);
    for (const auto& k : d.diag_enable) std::printf(" %s", k.c_str());
    std::printf(", output '%s'\n", outdir.empty() ? "." : outdir.c_str());
    try {
        if (d.yee) return run_yee(d, o, pref, outdir, argc, argv);
        if (d.darwin) return run_spectral<arc::CfgDarwin>(d, o, pref, outdir, argc, argv);
        return run_spectral<arc::Cfg>(d, o, pref, outdir, argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s: %s\n", kRunDeckCli ? "run_deck" : "arcsim", e.what());
        return 1;
    }
}
