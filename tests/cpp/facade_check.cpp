// facade_check.cpp -- the C++ facade (arcwarden_amd/include/pic) used the way
// the reference's tools and tests use arc:: (include/pic/*.hpp): a deck-driven
// Simulation<Cfg>, the standalone Depositor / Pusher / solver operators on
// user-owned arrays, and the direct array access tools rely on
// (P.ux.data(), .bytes(), f.Ex.size(), sim.particles(), sim.fields()).
//
// Usage: facade_check <two_stream.ini>.  Prints one line per check and exits
// non-zero on the first failure.
#include "pic/deck.hpp"
#include "pic/depositor.hpp"
#include "pic/diagnostics.hpp"
#include "pic/pusher.hpp"
#include "pic/simulation.hpp"
#include "pic/spectral.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <string>
#include <vector>

using namespace arc;

static int g_fail = 0;

static void expect(bool ok, const char* what, double a = 0.0, double b = 0.0) {
    std::printf("%s %-58s %.6g %.6g\n", ok ? "ok  " : "FAIL", what, a, b);
    std::fflush(stdout);
    if (!ok) ++g_fail;
}

static double sum(const std::vector<float>& v) {
    return std::accumulate(v.begin(), v.end(), 0.0, [](double s, float x) { return s + x; });
}

static double rel_l2(const std::vector<float>& a, const std::vector<float>& b) {
    double num = 0, den = 0;
    for (std::size_t i = 0; i < a.size(); ++i) {
        num += (double(a[i]) - b[i]) * (double(a[i]) - b[i]);
        den += double(b[i]) * b[i];
    }
    return std::sqrt(num / (den > 0 ? den : 1.0));
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <two_stream.ini>\n", argv[0]);
        return 2;
    }
    try {
        // ---- deck -> Simulation<Cfg> (ES), as tools/run_deck.cu --------------------
        Deck d = load_deck(argv[1]);
        expect(d.nx == 64 && d.species.size() == 2, "deck: grid and species");
        d.rp.dump_every = 5;
        d.rp.nsteps = 40;
        Grid g(d.nx, d.ny, d.Lx, d.Ly);
        for (auto& s : d.species) s.ppc = 256;  // keep the check small
        using CfgES = SimConfig<2, 3, ShapeOrder::CIC, float, false, TiledBinnedDeposit>;
        Simulation<CfgES> sim(g, d.rp, d.species, "facade_energy.csv");
        sim.init();
        Particles& P = sim.particles();
        expect(P.n == std::size_t(2 * 256 * 64 * 8), "particles().n", double(P.n));
        expect(P.ux.bytes() == P.n * sizeof(float), "P.ux.bytes()");
        sim.run();
        const auto& h = sim.diagnostics().history();
        expect(h.size() == 8, "diagnostics every 5 steps", double(h.size()));
        const double q0 = h.front().charge, q1 = h.back().charge;
        // the deposit's int64 total is exact; the diagnostic sums the per-cell
        // FP32 rho, each rounded by at most 2^-24 relative (one sign: electrons)
        expect(std::fabs(q1 - q0) <= 2.0 * 0x1p-24 * std::fabs(q0), "charge conserved", q0, q1);
        expect(std::fabs(q0 + d.Lx * d.Ly) < 1e-4 * d.Lx * d.Ly, "total charge = -n0 L^2", q0);
        const double e0 = h.front().total, e1 = h.back().total;
        expect(std::fabs(e1 - e0) < 0.02 * e0, "energy within 2% over 40 steps", e0, e1);
        const Fields& F = sim.fields();
        expect(F.Ex.size() == std::size_t(g.real_size()) && F.Bx.size() == 0,
               "fields(): Ex sized, no B in ES");

        // ---- standalone operators on user-owned arrays ----------------------------
        // ES step by hand: charge -> solve -> push -> migrate, against the engine's
        // rho of the same particle state.
        Particles& Q = sim.particles();
        std::vector<float> x = Q.x.to_host(), y = Q.y.to_host(), ux = Q.ux.to_host(),
                           uy = Q.uy.to_host(), uz = Q.uz.to_host();
        Particles mine;
        mine.allocate_n(Q.n);
        mine.x.from_host(x);
        mine.y.from_host(y);
        mine.ux.from_host(ux);
        mine.uy.from_host(uy);
        mine.uz.from_host(uz);
        expect(Q.w.size() == 0 && Q.w_uniform > 0, "equal weights: w empty, w_uniform set",
               Q.w_uniform);
        mine.w.reset();  // the same uniform weight as the engine
        mine.w_uniform = Q.w_uniform;
        Sources src;
        src.allocate(g);
        src.zero();
        Fields fl;
        fl.allocate(g);
        Depositor<CfgES>::charge(mine, src, g, d.rp, nullptr);
        Depositor<SimConfig<2, 3, ShapeOrder::CIC, float, false, AtomicGlobalDeposit>>::charge(
            mine, src, g, d.rp, nullptr);
        std::vector<float> rho2 = src.rho.to_host();
        const double qsum = sum(rho2) * g.dx * g.dy;
        expect(std::fabs(qsum - 2 * q1) < 1e-4 * std::fabs(q1), "two deposits = 2x charge", qsum,
               2 * q1);
        src.zero();
        Depositor<CfgES>::charge(mine, src, g, d.rp, nullptr);
        SpectralEngine eng(g);
        ElectrostaticSpectralSolver es;
        es.solve(src, fl, eng, d.rp, nullptr);
        std::vector<float> ex_mine = fl.Ex.to_host();
        sim.step(d.rp.nsteps);  // the engine's next step solves the same rho
        std::vector<float> ex_sim = sim.fields().Ex.to_host();
        // the engine takes the mean charge density off its transform input in
        // the integer domain; the standalone solve transforms the FP32 grid
        expect(rel_l2(ex_mine, ex_sim) < 1e-5, "standalone solve = engine field (FP32)",
               rel_l2(ex_mine, ex_sim));
        // with the engine's field the standalone push is the engine's step
        fl.Ex.from_host(ex_sim);
        fl.Ey.from_host(sim.fields().Ey.to_host());
        Pusher<CfgES>::boris(mine, fl, g, d.rp, nullptr);
        mine.migrate(g, nullptr);
        // the engine's step moved the same particles with the same field (it may
        // also have re-sorted them: compare the sorted coordinate multisets)
        std::vector<float> xs = mine.x.to_host(), xe = sim.particles().x.to_host();
        std::sort(xs.begin(), xs.end());
        std::sort(xe.begin(), xe.end());
        expect(xs == xe, "standalone push+migrate = engine step (bit-exact)");

        // ---- Darwin engine through the facade -------------------------------------
        RunParams rp;
        rp.dt = 0.05;
        rp.c = 5.0;
        rp.B0[2] = 0.3f;
        rp.dump_every = 1;
        rp.nsteps = 4;
        rp.noisy_load = 1;
        rp.rng_seed = 5;
        SpeciesList sp(1);
        sp[0].ppc = 16;
        sp[0].uth[0] = sp[0].uth[1] = sp[0].uth[2] = 0.05;
        sp[0].ufl[0] = 0.2;
        Grid gd(32, 16, 16.0, 8.0);
        Simulation<CfgDarwin> dsim(gd, rp, sp);
        dsim.init();
        dsim.run();
        const Fields& DF = dsim.fields();
        expect(DF.Bz.size() == std::size_t(gd.real_size()), "Darwin fields(): B present");
        std::vector<float> bz = DF.Bz.to_host();
        expect(std::fabs(sum(bz) / bz.size() - 0.3) < 1e-3, "mean Bz = B0", sum(bz) / bz.size());
        const Sources& DS = dsim.sources();
        std::vector<float> jx = DS.Jx.to_host();
        expect(std::fabs(sum(jx) / jx.size() + 0.2) < 0.02, "mean Jx = q n0 v_drift",
               sum(jx) / jx.size());
    } catch (const std::exception& e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 1;
    }
    std::printf("%s\n", g_fail ? "FAILED" : "PASSED");
    return g_fail ? 1 : 0;
}
