// pic/diag/spectrum.hpp -- a console line with the pump-mode powers of Ex
// (row 0) and the rms magnetic perturbation.
//
// Reference: SpectrumDiag (include/pic/diag/spectrum.hpp:13-47).  Every
// nsteps/40 steps: |E_M|^2, |E_2M|^2, |E_3M|^2 (M the pump mode, or modes
// 1, 2, 3 without a pump), the summed power of modes band_lo..band_hi, and
// dB_rms / |B0| with dB^2 = sum_q (B_q - B0_q)^2 reduced on the device.
#ifndef ARC_PIC_DIAG_SPECTRUM_HPP
#define ARC_PIC_DIAG_SPECTRUM_HPP

#include "pic/deck.hpp"
#include "pic/diag/module.hpp"

#include <cmath>
#include <cstdio>

namespace arc {
namespace diag {

class SpectrumDiag : public IDiag {
public:
    explicit SpectrumDiag(const Deck& d)
        : nx_(d.nx), M_(d.pump_M), band_lo_(d.band_lo), band_hi_(d.band_hi),
          every_(d.rp.nsteps / 40 > 0 ? d.rp.nsteps / 40 : 1) {}

    void sample(const DiagFrame& fr) override {
        if (fr.step % every_ != 0) return;
        device_sync();
        const std::vector<float> ex = first_row(fr.f.Ex, nx_);
        const int m1 = M_ > 0 ? M_ : 1;
        const double p1 = mode_power(ex, nx_, m1), p2 = mode_power(ex, nx_, 2 * m1),
                     p3 = mode_power(ex, nx_, 3 * m1);
        double band = 0.0;
        for (int m = band_lo_; m <= band_hi_; ++m) band += mode_power(ex, nx_, m);
        double db_rel = 0.0;
        if (fr.f.Bx.size() > 0) {
            const double db2 = sum_sq(fr.f.Bx, fr.rp.B0[0]) + sum_sq(fr.f.By, fr.rp.B0[1]) +
                               sum_sq(fr.f.Bz, fr.rp.B0[2]);
            db_rel = std::sqrt(db2 / fr.g.real_size()) / (fr.bmag > 0 ? fr.bmag : 1.0);
        }
        std::printf("t=%7.1f  |E_M|2=%.3e |E_2M|2=%.3e |E_3M|2=%.3e  band[%d-%d]=%.3e  "
                    "dB/B0=%.4f\n",
                    fr.t, p1, p2, p3, band_lo_, band_hi_, band, db_rel);
    }

private:
    int nx_, M_, band_lo_, band_hi_;
    long every_;
};

}  // namespace diag
}  // namespace arc

#endif  // ARC_PIC_DIAG_SPECTRUM_HPP
