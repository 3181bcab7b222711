// pic/simulation.hpp -- the PIC main loop.
//
// Reference: Simulation<Cfg> (include/pic/simulation.hpp:38-140).  The facade
// keeps its constructors, init / step / run and accessors; the loop itself
// runs in the library's engine (aw_sim_*), one HIP stream per simulation:
//   ES:     [rho^n] -> FFT Poisson -> fused gather + Boris + move + wrap +
//           deposit of rho^{n+1}, with a stable tile re-sort on a cadence;
//   Darwin: rho, J, amu -> E_L, B (+B0) -> dcu with E_L + pump -> E_T ->
//           E = E_L + E_T + pump -> fused EM push + next-state moments.
// particles(), fields() and sources() return non-owning arrays over the
// engine's buffers (re-read on every call: a re-sort swaps the particle
// buffers).  Diagnostics are reduced on the device every rp.dump_every
// steps and appear in diagnostics() (and the CSV) when it is read, at the
// end of run() and at destruction.
#ifndef ARC_PIC_SIMULATION_HPP
#define ARC_PIC_SIMULATION_HPP

#include "pic/config.hpp"
#include "pic/diagnostics.hpp"
#include "pic/fields.hpp"
#include "pic/grid.hpp"
#include "pic/particles.hpp"
#include "pic/sources.hpp"
#include "pic/species.hpp"

#include <string>
#include <vector>

namespace arc {

// MI355X execution knobs of the engine (set before init()).
struct ExecConfig {
    int tile_x = 0, tile_y = 0;  // 0: chosen from the particle density
    int blocks_per_tile = 0;     // 0: heuristic
    int sort_every = 20;         // stable tile re-sort cadence (0: adaptive only)
    double stray_frac = 0.0;     // adaptive re-sort threshold (0: 2%)
    bool autotune = false;       // time tile / workgroup candidates at init
    int current_deposit = AW_CURRENT_CIC;  // Darwin: AW_CURRENT_ESIRKEPOV for the
                                           // charge-conserving current
};

template <class Cfg = arc::Cfg>
class Simulation {
public:
    Simulation(const Grid& g, const RunParams& rp, const std::string& csv = "")
        : p_(rp), grid_(g), diag_(csv) {
        diag_.set_grid(g);
    }
    Simulation(const Grid& g, const RunParams& rp, const SpeciesList& species,
               const std::string& csv = "")
        : p_(rp), grid_(g), species_(species), diag_(csv) {
        diag_.set_grid(g);
    }
    ~Simulation() {
        if (sim_) {
            try {
                sync_diag();
            } catch (...) {
            }
            (void)aw_sim_destroy(sim_);
        }
    }
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    ExecConfig& exec() { return exec_; }

    // Load particles, compute the t=0 field, roll velocities back half a step
    // (ES); Darwin deposits the t=0 moments.
    void init() {
        create();
        check(aw_sim_init(sim_));
    }

    // Writes a restart file (ARCW v1 with the state arrays) after the steps
    // run so far; deck is stored verbatim in the header.
    void save_checkpoint(const std::string& path, long step, const std::string& deck = "",
                         const std::string& git_hash = "") {
        check(aw_sim_checkpoint_save(sim_, path.c_str(), step, step * p_.dt, deck.c_str(),
                                     git_hash.c_str()));
    }

    // Instead of init(): the state of a file written by save_checkpoint for
    // the same grid, species and configuration; continue with step(saved step).
    void restore_checkpoint(const std::string& path) {
        create();
        check(aw_sim_checkpoint_restore(sim_, path.c_str()));
    }

private:
    void create() {
        aw_sim_config c{};
        c.has_b0 = Cfg::has_b0 ? 1 : 0;
        c.deposit_policy = Cfg::aw_deposit;
        c.field_model = Cfg::aw_model;
        c.tile_x = exec_.tile_x;
        c.tile_y = exec_.tile_y;
        c.blocks_per_tile = exec_.blocks_per_tile;
        c.sort_every = exec_.sort_every;
        c.current_deposit = exec_.current_deposit;
        c.stray_frac = exec_.stray_frac;
        c.autotune = exec_.autotune ? 1 : 0;
        c.legacy_load = species_.empty() ? 1 : 0;
        const aw_grid g = grid_.to_c();
        const std::vector<aw_species> sp = to_c(species_);
        if (sim_) {
            (void)aw_sim_destroy(sim_);
            sim_ = nullptr;
        }
        check(aw_sim_create(&g, &p_, sp.empty() ? nullptr : sp.data(),
                            static_cast<int32_t>(sp.size()), &c, &sim_));
    }

public:
    void step(long n) { check(aw_sim_step(sim_, n)); }

    void run() {
        for (long n = 0; n < p_.nsteps; ++n) step(n);
        check(aw_sim_synchronize(sim_));
        sync_diag();
    }

    void synchronize() { check(aw_sim_synchronize(sim_)); }

    const Diagnostics& diagnostics() {
        sync_diag();
        return diag_;
    }
    const Fields& fields() {
        aw_fields f{};
        check(aw_sim_fields(sim_, &f));
        const std::size_t n = static_cast<std::size_t>(grid_.real_size());
        auto b = [&](float* p) { return DeviceArray<float>::borrow(p, p ? n : 0); };
        fields_.Ex = b(f.Ex);
        fields_.Ey = b(f.Ey);
        fields_.Ez = b(f.Ez);
        fields_.Bx = b(f.Bx);
        fields_.By = b(f.By);
        fields_.Bz = b(f.Bz);
        fields_.ELx = b(f.ELx);
        fields_.ELy = b(f.ELy);
        fields_.ETx = b(f.ETx);
        fields_.ETy = b(f.ETy);
        fields_.ETz = b(f.ETz);
        return fields_;
    }
    const Sources& sources() {
        aw_sources s{};
        check(aw_sim_sources(sim_, &s));
        const std::size_t n = static_cast<std::size_t>(grid_.real_size());
        auto b = [&](float* p) { return DeviceArray<float>::borrow(p, p ? n : 0); };
        sources_.rho = b(s.rho);
        sources_.Jx = b(s.Jx);
        sources_.Jy = b(s.Jy);
        sources_.Jz = b(s.Jz);
        sources_.dcux = b(s.dcux);
        sources_.dcuy = b(s.dcuy);
        sources_.dcuz = b(s.dcuz);
        sources_.amu0 = b(s.amu0);
        sources_.amu1 = b(s.amu1);
        sources_.amu2 = b(s.amu2);
        sources_.amu3 = b(s.amu3);
        return sources_;
    }
    Particles& particles() {
        aw_particles p{};
        check(aw_sim_particles(sim_, &p));
        const std::size_t n = static_cast<std::size_t>(p.n);
        particles_.n = n;
        particles_.x = DeviceArray<float>::borrow(p.x, n);
        particles_.y = DeviceArray<float>::borrow(p.y, n);
        particles_.ux = DeviceArray<float>::borrow(p.ux, n);
        particles_.uy = DeviceArray<float>::borrow(p.uy, n);
        particles_.uz = DeviceArray<float>::borrow(p.uz, n);
        particles_.w = DeviceArray<float>::borrow(p.w, p.w ? n : 0);
        particles_.w_uniform = p.w_uniform;
        particles_.sorted = true;
        return particles_;
    }
    const RunParams& params() const { return p_; }
    const Grid& grid() const { return grid_; }
    aw_sim* handle() const { return sim_; }

private:
    void sync_diag() {
        if (!sim_) return;
        int64_t count = 0;
        check(aw_sim_history(sim_, nullptr, 0, &count));
        const std::size_t have = diag_.history().size();
        if (static_cast<std::size_t>(count) <= have) return;
        std::vector<aw_diag_sample> h(static_cast<std::size_t>(count));
        check(aw_sim_history(sim_, h.data(), count, &count));
        for (std::size_t k = have; k < h.size(); ++k) diag_.record(DiagSample::from_c(h[k]));
    }

    RunParams p_;
    Grid grid_;
    SpeciesList species_;
    ExecConfig exec_;
    aw_sim* sim_ = nullptr;
    Particles particles_;
    Sources sources_;
    Fields fields_;
    Diagnostics diag_;
};

}  // namespace arc

#endif  // ARC_PIC_SIMULATION_HPP
