// pic/diag/kt.hpp -- the longitudinal field E_Lx(x, t) along row 0, for k-t
// and omega-k plots.
//
// Reference: KTDump (include/pic/diag/kt.hpp:14-44).  Every kt_stride steps
// (default 5) the nx values of ELx on row 0 are appended to <pref>kt.bin as
// FP32; finalize() writes <pref>kt.meta (key value per line).
#ifndef ARC_PIC_DIAG_KT_HPP
#define ARC_PIC_DIAG_KT_HPP

#include "pic/deck.hpp"
#include "pic/diag/module.hpp"

#include <cstdio>
#include <fstream>
#include <string>

namespace arc {
namespace diag {

class KTDump : public IDiag {
public:
    KTDump(const Deck& d, const std::string& pref)
        : nx_(d.nx), every_(d.kt_stride > 0 ? d.kt_stride : 5), M_(d.pump_M),
          band_lo_(d.band_lo), band_hi_(d.band_hi), w0_(d.rp.pump_w0), k0_(d.rp.pump_k0),
          toff_(d.rp.pump_toff), Lx_(d.Lx), dt_(d.rp.dt), meta_path_(pref + "kt.meta") {
        bin_ = std::fopen((pref + "kt.bin").c_str(), "wb");
    }
    ~KTDump() override { close(); }

    void sample(const DiagFrame& fr) override {
        if (!bin_ || fr.step % every_ != 0 || fr.f.ELx.size() == 0) return;
        device_sync();
        const std::vector<float> row = first_row(fr.f.ELx, nx_);
        std::fwrite(row.data(), sizeof(float), row.size(), bin_);
        ++samples_;
    }

    void finalize() override {
        close();
        std::ofstream m(meta_path_);
        m << "nx " << nx_ << "\nnsamp " << samples_ << "\ndt_sample " << every_ * dt_
          << "\ntoff " << toff_ << "\nLx " << Lx_ << "\nM " << M_ << "\nw0 " << w0_
          << "\nband_lo " << band_lo_ << "\nband_hi " << band_hi_ << "\nvphx "
          << (k0_ > 0 ? w0_ / k0_ : 0.0) << "\n";
    }

private:
    void close() {
        if (bin_) std::fclose(bin_);
        bin_ = nullptr;
    }

    int nx_, every_, M_, band_lo_, band_hi_;
    double w0_, k0_, toff_, Lx_, dt_;
    std::string meta_path_;
    std::FILE* bin_ = nullptr;
    int samples_ = 0;
};

}  // namespace diag
}  // namespace arc

#endif  // ARC_PIC_DIAG_KT_HPP
