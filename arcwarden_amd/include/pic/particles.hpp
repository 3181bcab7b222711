// pic/particles.hpp -- particle SoA.
//
// Reference: Particles (include/pic/particles.hpp:453-726).  x, y in cell
// units, u = gamma v with gamma = 1, per-particle weight w.  initialize()
// runs the library's loaders (stratified quiet or hashed noisy positions,
// Maxwellian velocities through erfinv, per-species weight
// density*dx*dy/ppc); migrate() and sort_by_tile() are the library's
// periodic wrap and stable tile sort.  The facade always stores w; the
// library reads it per particle.  Particles borrowed from a Simulation whose
// species share one weight leave w empty and carry w_uniform instead.
#ifndef ARC_PIC_PARTICLES_HPP
#define ARC_PIC_PARTICLES_HPP

#include "pic/config.hpp"
#include "pic/device_array.hpp"
#include "pic/grid.hpp"
#include "pic/species.hpp"

#include <vector>

namespace arc {

struct Particles {
    DeviceArray<float> x, y;
    DeviceArray<float> ux, uy, uz;
    DeviceArray<float> w;
    DeviceArray<int> cell;
    std::size_t n = 0;
    double w_uniform = 0.0;  // weight of every particle when w is empty
    bool sorted = false;
    DeviceArray<int> tile_off;  // tile offsets after sort_by_tile (ntiles + 1)
    // delta-f marker weight and its precision-study accumulators (Yee branch)
    DeviceArray<float> wd, wc;
    DeviceArray<double> wdd;
    bool has_wd = false;
    int wprec = 0;

    Particles() = default;

    void allocate_n(std::size_t n_total) {
        n = n_total;
        x = DeviceArray<float>(n);
        y = DeviceArray<float>(n);
        ux = DeviceArray<float>(n);
        uy = DeviceArray<float>(n);
        uz = DeviceArray<float>(n);
        w = DeviceArray<float>(n);
        cell = DeviceArray<int>(n);
    }
    void allocate(const Grid& g, int ppc) {
        allocate_n(static_cast<std::size_t>(g.real_size()) * static_cast<std::size_t>(ppc));
    }

    aw_particles views() {
        aw_particles p{};
        p.x = x.data();
        p.y = y.data();
        p.ux = ux.data();
        p.uy = uy.data();
        p.uz = uz.data();
        p.w = w.size() ? w.data() : nullptr;
        p.n = static_cast<int64_t>(n);
        p.w_uniform = w_uniform;
        return p;
    }

    // Particles::initialize(species, grid, rp, stream) (:521-559): one block
    // per species, per-species weights, golden-ratio phase per species.
    void initialize(const SpeciesList& species, const Grid& g, const RunParams& rp,
                    hipStream_t s) {
        std::size_t total = 0;
        for (const auto& sp : species)
            total += static_cast<std::size_t>(sp.ppc) * static_cast<std::size_t>(g.real_size());
        allocate_n(total);
        const std::vector<aw_species> c = to_c(species);
        aw_particles p = views();
        std::vector<double> wts(c.size());
        check(aw_load_species(g.ctx(), c.data(), static_cast<int32_t>(c.size()), &rp, nullptr, 0,
                              &p, wts.data(), to_aw(s)));
        sorted = false;
    }

    // Particles::initialize(rp, grid, stream) (:318-379): the single-Maxwellian
    // legacy loader (perturb_amp, two_stream, bump_on_tail from RunParams).
    void initialize(const RunParams& rp, const Grid& g, hipStream_t s) {
        if (n == 0) allocate(g, rp.ppc);
        aw_particles p = views();
        check(aw_load_legacy(g.ctx(), &rp, &p, to_aw(s)));  // w = rp.weight
        sorted = false;
    }

    // Particles::enable_deltaf (:484-491): wd = 0, plus the Kahan compensation
    // (wprec 1) or the FP64 accumulator (wprec 2).
    void enable_deltaf(hipStream_t s = nullptr, int wprec_ = 0) {
        wd = DeviceArray<float>(n);
        wd.zero(s);
        has_wd = true;
        wprec = wprec_;
        if (wprec == 1) {
            wc = DeviceArray<float>(n);
            wc.zero(s);
        }
        if (wprec == 2) {
            wdd = DeviceArray<double>(n);
            wdd.zero(s);
        }
    }

    aw_deltaf_state deltaf_state() {
        aw_deltaf_state d{};
        d.wd = wd.size() ? wd.data() : nullptr;
        d.wc = wc.size() ? wc.data() : nullptr;
        d.wdd = wdd.size() ? wdd.data() : nullptr;
        return d;
    }

    // Particles::initialize_mirror (:573-612): the mirror-equilibrium load of
    // one species in the rp.b0_prof background (ny = 1); the marker count
    // follows the profile.
    void initialize_mirror(const Species& q, const Grid& g, const RunParams& rp,
                           hipStream_t s) {
        const aw_grid cg = g.to_c();
        const aw_species cq = q.to_c();
        int64_t count = 0;
        check(aw_mirror_count(&cg, &rp, &cq, &count));
        allocate_n(static_cast<std::size_t>(count));
        aw_particles p = views();
        check(aw_load_mirror(&cg, &rp, &cq, &p, to_aw(s)));
        sorted = false;
    }

    // Particles::migrate (:615-621): periodic wrap + owning cell.
    void migrate(const Grid& g, hipStream_t s) {
        aw_particles p = views();
        check(aw_migrate(g.ctx(), &p, cell.data(), to_aw(s)));
    }

    // Particles::sort_by_tile (:678-726): stable physical reorder into tile
    // order (the reference's order within a tile is atomic-order; this one is
    // the stable order).
    void sort_by_tile(const Grid& g, hipStream_t s) {
        aw_ctx* c = g.ctx();
        int32_t ntiles = 0, ntx = 0, tx = 0, ty = 0;
        check(aw_ctx_tiles(c, &ntiles, &ntx, &tx, &ty));
        Particles dst;
        dst.allocate_n(n);
        if (tile_off.size() != static_cast<std::size_t>(ntiles) + 1)
            tile_off = DeviceArray<int>(static_cast<std::size_t>(ntiles) + 1);
        aw_particles a = views(), b = dst.views();
        check(aw_sort_by_tile(c, &a, &b, tile_off.data(), to_aw(s)));
        stream_synchronize(s);
        x = std::move(dst.x);
        y = std::move(dst.y);
        ux = std::move(dst.ux);
        uy = std::move(dst.uy);
        uz = std::move(dst.uz);
        if (w.size()) w = std::move(dst.w);
        sorted = true;
    }
};

}  // namespace arc

#endif  // ARC_PIC_PARTICLES_HPP
