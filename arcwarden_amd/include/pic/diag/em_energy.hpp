// pic/diag/em_energy.hpp -- electromagnetic field energies of a Darwin (or
// Yee) run.
//
// Reference: EMEnergy (include/pic/diag/em_energy.hpp:14-38).  Every
// nsteps/40 steps (at least 1) it appends step, time and
// 1/2 dx dy sum |F|^2 of E (total), B and E_T to <pref>emenergy.csv; the sums
// are FP64 reductions on the device.  Runs without Ez (electrostatic) write
// only the header.
#ifndef ARC_PIC_DIAG_EM_ENERGY_HPP
#define ARC_PIC_DIAG_EM_ENERGY_HPP

#include "pic/deck.hpp"
#include "pic/diag/module.hpp"

#include <cstdio>
#include <fstream>
#include <string>

namespace arc {
namespace diag {

class EMEnergy : public IDiag {
public:
    EMEnergy(const Deck& d, const std::string& pref)
        : every_(d.rp.nsteps / 40 > 0 ? d.rp.nsteps / 40 : 1) {
        csv_.open(pref + "emenergy.csv");
        if (csv_) csv_ << "step,time,E_total,B,E_T\n";
    }

    void sample(const DiagFrame& fr) override {
        if (fr.step % every_ != 0 || fr.f.Ez.size() == 0) return;
        device_sync();
        const double half_cell = 0.5 * fr.g.dx * fr.g.dy;
        auto energy = [&](const DeviceArray<float>& a, const DeviceArray<float>& b,
                          const DeviceArray<float>& c) {
            return half_cell * (sum_sq(a) + sum_sq(b) + sum_sq(c));
        };
        we_ = energy(fr.f.Ex, fr.f.Ey, fr.f.Ez);
        wb_ = energy(fr.f.Bx, fr.f.By, fr.f.Bz);
        wt_ = fr.f.ETx.size() ? energy(fr.f.ETx, fr.f.ETy, fr.f.ETz) : 0.0;
        if (csv_) csv_ << fr.step << ',' << fr.t << ',' << we_ << ',' << wb_ << ',' << wt_ << '\n';
    }

    void finalize() override {
        if (csv_) csv_.flush();
        std::printf("  em energy at the last sample: E %.4e, B %.4e (B/E %.2e), E_T %.4e "
                    "(E_T/E %.2e)\n",
                    we_, wb_, we_ > 0 ? wb_ / we_ : 0.0, wt_, we_ > 0 ? wt_ / we_ : 0.0);
    }

    double electric() const { return we_; }
    double magnetic() const { return wb_; }
    double transverse() const { return wt_; }

private:
    long every_;
    std::ofstream csv_;
    double we_ = 0.0, wb_ = 0.0, wt_ = 0.0;
};

}  // namespace diag
}  // namespace arc

#endif  // ARC_PIC_DIAG_EM_ENERGY_HPP
