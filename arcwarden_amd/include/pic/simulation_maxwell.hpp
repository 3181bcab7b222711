// pic/simulation_maxwell.hpp -- the full-Maxwell (Yee) main loop.
//
// Reference: MaxwellSimulation (include/pic/simulation_maxwell.hpp:24-227)
// and YeeFields (include/pic/yee2d.hpp:684-709).  The facade keeps the
// reference's names: the simulation owns the Particles (callers load them
// through particles().initialize / initialize_mirror / enable_deltaf, as the
// reference's tools do) and the YeeFields arrays ex_ .. jz_, which borrow the
// engine's mesh (aw_maxwell_*).  Each step is B half step, push + charge-
// conserving current (+ wrap), smoothing, antenna / cold fluid, B half step,
// E full step, damping (see aw_maxwell_step); the particle order never
// changes.
#ifndef ARC_PIC_SIMULATION_MAXWELL_HPP
#define ARC_PIC_SIMULATION_MAXWELL_HPP

#include "pic/config.hpp"
#include "pic/fields.hpp"
#include "pic/grid.hpp"
#include "pic/particles.hpp"

#include <cstddef>

namespace arc {

struct YeeMaxwell {};  // field-model tag of the third branch

// The engine's stream: converts to hipStream_t and synchronises.
class SimStream {
public:
    explicit SimStream(hipStream_t s = nullptr) : s_(s) {}
    operator hipStream_t() const { return s_; }
    hipStream_t get() const { return s_; }
    void synchronize() const { stream_synchronize(s_); }

private:
    hipStream_t s_;
};

class YeeFields {
public:
    DeviceArray<float> ex_, ey_, ez_, bx_, by_, bz_, jx_, jy_, jz_;

    void bind(const aw_yee_mesh& m, const Grid& g) {
        nx_ = g.nx;
        ny_ = g.ny;
        const std::size_t n = static_cast<std::size_t>(g.real_size());
        ex_ = DeviceArray<float>::borrow(m.ex, n);
        ey_ = DeviceArray<float>::borrow(m.ey, n);
        ez_ = DeviceArray<float>::borrow(m.ez, n);
        bx_ = DeviceArray<float>::borrow(m.bx, n);
        by_ = DeviceArray<float>::borrow(m.by, n);
        bz_ = DeviceArray<float>::borrow(m.bz, n);
        jx_ = DeviceArray<float>::borrow(m.jx, n);
        jy_ = DeviceArray<float>::borrow(m.jy, n);
        jz_ = DeviceArray<float>::borrow(m.jz, n);
    }
    void zero(hipStream_t s = nullptr) {
        for (DeviceArray<float>* a : {&ex_, &ey_, &ez_, &bx_, &by_, &bz_}) a->zero(s);
        zero_j(s);
    }
    void zero_j(hipStream_t s = nullptr) {
        jx_.zero(s);
        jy_.zero(s);
        jz_.zero(s);
    }
    int nx() const { return nx_; }
    int ny() const { return ny_; }

private:
    int nx_ = 0, ny_ = 0;
};

class MaxwellSimulation {
public:
    MaxwellSimulation(const Grid& g, const RunParams& rp) : g_(g), rp_(rp) {
        const aw_grid cg = g_.to_c();
        check(aw_maxwell_create(&cg, &rp_, &m_));
        aw_stream s = nullptr;
        check(aw_maxwell_stream(m_, &s));
        s_ = SimStream(reinterpret_cast<hipStream_t>(s));
        aw_yee_mesh mesh{};
        check(aw_maxwell_mesh(m_, &mesh));
        flds_.bind(mesh, g_);
        const std::size_t n = static_cast<std::size_t>(g_.real_size());
        vcy_ = DeviceArray<float>::borrow(mesh.vcy, mesh.vcy ? n : 0);
        vcz_ = DeviceArray<float>::borrow(mesh.vcz, mesh.vcz ? n : 0);
    }
    ~MaxwellSimulation() {
        if (m_) (void)aw_maxwell_destroy(m_);
    }
    MaxwellSimulation(const MaxwellSimulation&) = delete;
    MaxwellSimulation& operator=(const MaxwellSimulation&) = delete;

    Grid& grid() { return g_; }
    RunParams& params() { return rp_; }
    Particles& particles() { return parts_; }
    YeeFields& fields() { return flds_; }
    SimStream& stream() { return s_; }
    DeviceArray<float>& vcy() { return vcy_; }
    DeviceArray<float>& vcz() { return vcz_; }

    void step() {
        step_at(static_cast<double>(nstep_) * rp_.dt);
        ++nstep_;
    }

    void step_at(double tnow) {
        aw_particles p = parts_.views();
        aw_deltaf_state df = parts_.deltaf_state();
        check(aw_maxwell_step_at(m_, &p, rp_.deltaf ? &df : nullptr, tnow));
    }

    void synchronize() { check(aw_maxwell_synchronize(m_)); }

    struct Energies {
        double we, wb;
    };
    Energies field_energy() {
        Energies e{};
        check(aw_maxwell_field_energy(m_, &e.we, &e.wb));
        return e;
    }

    struct Residuals {
        double divb_rms, gauss_drift_rms;
    };
    Residuals residuals() {
        Residuals r{};
        aw_particles p = parts_.views();
        check(aw_maxwell_residuals(m_, &p, rp_.deltaf && parts_.wd.size() ? parts_.wd.data()
                                                                          : nullptr,
                                   &r.divb_rms, &r.gauss_drift_rms));
        return r;
    }

    struct WdStats {
        double sum, rms, max;
    };
    WdStats wd_stats() {
        WdStats w{0.0, 0.0, 0.0};
        if (!rp_.deltaf || parts_.wd.size() == 0) return w;
        check(aw_wd_stats(parts_.wd.data(), static_cast<int64_t>(parts_.n), &w.sum, &w.rms,
                          &w.max, to_aw(s_.get())));
        return w;
    }

    // The Fields the diagnostics modules read at time t: wave E + pump, wave
    // B + B0, ELx = wave Ex (the reference's runner builds the same mirror).
    const Fields& diag_fields(double t) {
        aw_fields f{};
        check(aw_maxwell_diag_fields(m_, t, &f));
        const std::size_t n = static_cast<std::size_t>(g_.real_size());
        auto b = [&](float* p) { return DeviceArray<float>::borrow(p, p ? n : 0); };
        diag_.Ex = b(f.Ex);
        diag_.Ey = b(f.Ey);
        diag_.Ez = b(f.Ez);
        diag_.Bx = b(f.Bx);
        diag_.By = b(f.By);
        diag_.Bz = b(f.Bz);
        diag_.ELx = b(f.ELx);
        return diag_;
    }

    aw_maxwell* handle() const { return m_; }

private:
    Grid g_;
    RunParams rp_;
    aw_maxwell* m_ = nullptr;
    SimStream s_;
    Particles parts_;
    YeeFields flds_;
    Fields diag_;
    DeviceArray<float> vcy_, vcz_;
    long nstep_ = 0;
};

}  // namespace arc

#endif  // ARC_PIC_SIMULATION_MAXWELL_HPP
