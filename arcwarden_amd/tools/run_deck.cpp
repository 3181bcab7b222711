// run_deck -- run a simulation from a text input deck.
//
//   run_deck <deck-file> [outdir] [--em] [--c=20] [--ndc=2] [--autotune]
//
// The reference's tools/run_deck.cu on the facade: reads the deck, runs the
// PIC loop (electrostatic, or the spectral Darwin model with --em or
// [field] model = darwin), and when the deck sets dump_every writes
// energy.csv, (x, vx) phase-space frames frame_%04d.csv (at most 100000
// sampled particles) and manifest.csv into outdir.
#include "pic/deck.hpp"
#include "pic/simulation.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <numeric>
#include <random>
#include <string>
#include <vector>

using namespace arc;

static constexpr int kDumpCap = 100000;

template <class C>
static int run_sim(Deck& d, const std::string& outdir, const std::string& deckname,
                   bool autotune) {
    namespace fs = std::filesystem;
    const bool frames = d.dump_every > 0;
    if (frames) {
        fs::create_directories(outdir);
        d.rp.dump_every = d.dump_every;
    } else {
        d.rp.dump_every = 1L << 30;
    }
    Grid g(d.nx, d.ny, d.Lx, d.Ly);
    Simulation<C> sim(g, d.rp, d.species, frames ? outdir + "/energy.csv" : "");
    sim.exec().autotune = autotune;
    sim.init();
    const long long N = static_cast<long long>(sim.particles().n);
    const bool em = C::field_model == FieldModel::Darwin;
    std::printf("deck '%s': %lld particles, %dx%d grid, %s load, %s, %s\n", deckname.c_str(), N,
                g.nx, g.ny, d.rp.noisy_load ? "noisy" : "quiet",
                em ? "Darwin EM" : "electrostatic", aw_version());
    if (em) std::printf("  Darwin: c=%.3g ndc=%d\n", d.rp.c, d.rp.ndc);
    for (const auto& s : d.species)
        std::printf("  species %-10s density=%.3g ppc=%d ufl=(%.3g,%.3g,%.3g) "
                    "uth=(%.3g,%.3g,%.3g) weight=%.3g\n",
                    s.name.c_str(), s.density, s.ppc, s.ufl[0], s.ufl[1], s.ufl[2], s.uth[0],
                    s.uth[1], s.uth[2], s.density * g.dx * g.dy / s.ppc);

    // a fixed sample of particle slots (the engine re-sorts, so the sampled
    // slots are a random sample of the population at each frame)
    std::vector<int> sample(static_cast<std::size_t>(N));
    std::iota(sample.begin(), sample.end(), 0);
    if (N > kDumpCap) {
        std::mt19937 rng(12345);
        std::shuffle(sample.begin(), sample.end(), rng);
        sample.resize(kDumpCap);
        std::sort(sample.begin(), sample.end());
    }
    int frame = 0;
    std::ofstream man;
    if (frames) {
        man.open(outdir + "/manifest.csv");
        man << "frame,step,time,file\n";
    }
    auto write_frame = [&](long step) {
        if (!frames) return;
        sim.synchronize();
        Particles& P = sim.particles();
        const std::vector<float> x = P.x.to_host(), vx = P.ux.to_host();
        char name[64];
        std::snprintf(name, sizeof(name), "frame_%04d.csv", frame);
        std::ofstream f(outdir + "/" + name);
        f << "x,vx\n";
        for (int i : sample) f << x[i] * g.dx << ',' << vx[i] << '\n';
        man << frame << ',' << step << ',' << step * d.rp.dt << ',' << name << '\n';
        man.flush();
        ++frame;
    };

    write_frame(0);
    const auto t0 = std::chrono::steady_clock::now();
    for (long n = 0; n < d.rp.nsteps; ++n) {
        sim.step(n);
        if (frames && (n + 1) % d.dump_every == 0) write_frame(n + 1);
    }
    sim.synchronize();
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    (void)sim.diagnostics();  // flushes energy.csv
    if (frames) std::printf("wrote %d frames to %s/ (<= %d pts/frame)\n", frame, outdir.c_str(),
                            kDumpCap);
    std::printf("ran %ld steps in %.3f s (%.3f ms/step, %.1f M particle-updates/s)\n",
                (long)d.rp.nsteps, secs, 1e3 * secs / std::max<long>(1, d.rp.nsteps),
                double(N) * d.rp.nsteps / std::max(secs, 1e-12) / 1e6);
    if (em) {
        const Fields& f = sim.fields();
        auto sq = [](const DeviceArray<float>& a) {
            double s = 0;
            for (float v : a.to_host()) s += double(v) * v;
            return s;
        };
        const double cell = 0.5 * g.dx * g.dy;
        const double eE = cell * (sq(f.Ex) + sq(f.Ey) + sq(f.Ez));
        const double eB = cell * (sq(f.Bx) + sq(f.By) + sq(f.Bz));
        const double eT = cell * (sq(f.ETx) + sq(f.ETy) + sq(f.ETz));
        std::printf("  final field energy: electric(E_total)=%.4e  magnetic=%.4e (%.2e of E)"
                    "  transverse-E=%.4e (%.2e of E)\n",
                    eE, eB, eE > 0 ? eB / eE : 0.0, eT, eE > 0 ? eT / eE : 0.0);
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <deck-file> [outdir] [--em] [--c=20] [--ndc=2] "
                             "[--autotune]\n", argv[0]);
        return 2;
    }
    bool em = false, autotune = false;
    double c = 20.0;
    int ndc = 2;
    std::string deckfile, outdir;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--em") em = true;
        else if (a == "--autotune") autotune = true;
        else if (a.rfind("--c=", 0) == 0) c = std::atof(a.c_str() + 4);
        else if (a.rfind("--ndc=", 0) == 0) ndc = std::atoi(a.c_str() + 6);
        else if (deckfile.empty()) deckfile = a;
        else if (outdir.empty()) outdir = a;
    }
    try {
        Deck d = load_deck(deckfile);
        if (d.yee) {
            std::fprintf(stderr, "run_deck: [field] model = yee is not part of this build\n");
            return 2;
        }
        if (outdir.empty()) outdir = d.outdir;
        const std::string name = std::filesystem::path(deckfile).filename().string();
        if (em && !d.darwin) {
            d.rp.c = c;
            d.rp.ndc = ndc;
        }
        const bool b0 = d.rp.B0[0] != 0.f || d.rp.B0[1] != 0.f || d.rp.B0[2] != 0.f;
        if (em || d.darwin) return run_sim<CfgDarwin>(d, outdir, name, autotune);
        if (b0) return run_sim<SimConfig<2, 3, ShapeOrder::CIC, float, true, TiledBinnedDeposit>>(
            d, outdir, name, autotune);
        return run_sim<SimConfig<2, 3, ShapeOrder::CIC, float, false, TiledBinnedDeposit>>(
            d, outdir, name, autotune);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "run_deck: %s\n", e.what());
        return 1;
    }
}
