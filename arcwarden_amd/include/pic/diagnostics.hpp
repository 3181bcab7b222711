// pic/diagnostics.hpp -- scalar diagnostics and the energy CSV.
//
// Reference: DiagSample / Diagnostics (include/pic/diagnostics.hpp:36-130):
// KE = 1/2 sum w u^2, EE = 1/2 eps0 sum (Ex^2 + Ey^2) dx dy, charge =
// sum rho dx dy, max|E|, all accumulated in FP64; a CSV row
// "step,time,ke,ee,total,charge,max_e" per dump step.  The sums run on the
// device (aw_reduce_energies) instead of after a full copy to the host.
#ifndef ARC_PIC_DIAGNOSTICS_HPP
#define ARC_PIC_DIAGNOSTICS_HPP

#include "pic/config.hpp"
#include "pic/fields.hpp"
#include "pic/grid.hpp"
#include "pic/particles.hpp"
#include "pic/sources.hpp"

#include <fstream>
#include <string>
#include <vector>

namespace arc {

struct DiagSample {
    long step = 0;
    double time = 0.0;
    double ke = 0.0;
    double ee = 0.0;
    double total = 0.0;
    double charge = 0.0;
    double max_e = 0.0;

    static DiagSample from_c(const aw_diag_sample& c) {
        return DiagSample{static_cast<long>(c.step), c.time, c.ke, c.ee, c.total, c.charge,
                          c.max_e};
    }
};

class Diagnostics {
public:
    Diagnostics() = default;
    explicit Diagnostics(const std::string& csv_path) {
        if (!csv_path.empty()) {
            csv_.open(csv_path);
            csv_ << "step,time,ke,ee,total,charge,max_e\n";
        }
    }

    // The grid whose context runs the reductions (the reference only needs
    // dx, dy: set_geometry).
    void set_grid(const Grid& g) { g_ = g; }
    void set_geometry(double dx, double dy) {
        g_.dx = dx;
        g_.dy = dy;
    }

    DiagSample compute(long step, Particles& p, const Sources& src, const Fields& f,
                       const RunParams& rp, hipStream_t s) {
        aw_particles v = p.views();
        aw_diag_sample d{};
        check(aw_reduce_energies(g_.ctx(), &v, f.Ex.data(), f.Ey.data(), src.rho.data(),
                                 rp.eps0, &d, to_aw(s)));
        DiagSample smp = DiagSample::from_c(d);
        smp.step = step;
        smp.time = step * rp.dt;
        hist_.push_back(smp);
        return smp;
    }

    void maybe_compute(long step, Particles& p, const Sources& src, const Fields& f,
                       const RunParams& rp, hipStream_t s) {
        if (rp.dump_every > 0 && (step % rp.dump_every) != 0) return;
        record(compute(step, p, src, f, rp, s), false);
    }

    // Appends a sample computed elsewhere (the Simulation engine) and writes
    // its CSV row.
    void record(const DiagSample& smp, bool push = true) {
        if (push) hist_.push_back(smp);
        if (csv_.is_open()) {
            csv_ << smp.step << ',' << smp.time << ',' << smp.ke << ',' << smp.ee << ','
                 << smp.total << ',' << smp.charge << ',' << smp.max_e << '\n';
            csv_.flush();
        }
    }

    const std::vector<DiagSample>& history() const { return hist_; }
    const DiagSample& last() const { return hist_.back(); }

private:
    std::ofstream csv_;
    std::vector<DiagSample> hist_;
    Grid g_;
};

}  // namespace arc

#endif  // ARC_PIC_DIAGNOSTICS_HPP
