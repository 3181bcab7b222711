// pic/diag/phase_video.hpp -- movie frames: f(x, v_par/v_th) and three field
// rows per frame.
//
// Reference: PhaseVideo (include/pic/diag/phase_video.hpp:14-70).  About
// n_frames (default 150) frames over the run: <pref>vid_phase.bin gets the
// 200 x 256 histogram (v_par/v_th in [-6, 6) by x) as FP32 counts, binned on
// the device; <pref>vid_field.bin gets E_par = Ex b_x + Ez b_z, E_perp = Ey
// and dB_y = By - B0_y along row 0; <pref>vid.meta is written at finalize().
#ifndef ARC_PIC_DIAG_PHASE_VIDEO_HPP
#define ARC_PIC_DIAG_PHASE_VIDEO_HPP

#include "pic/deck.hpp"
#include "pic/diag/module.hpp"

#include <cstdio>
#include <fstream>
#include <string>

namespace arc {
namespace diag {

class PhaseVideo : public IDiag {
public:
    static constexpr int kXBins = 256, kVBins = 200;
    static constexpr double kVLo = -6.0, kVHi = 6.0;

    PhaseVideo(const Deck& d, const std::string& pref)
        : nx_(d.nx), toff_(d.rp.pump_toff), Lx_(d.Lx), dt_(d.rp.dt),
          meta_path_(pref + "vid.meta") {
        const long frames = d.n_frames > 0 ? d.n_frames : 150;
        every_ = d.rp.nsteps / frames > 0 ? static_cast<int>(d.rp.nsteps / frames) : 1;
        phase_ = std::fopen((pref + "vid_phase.bin").c_str(), "wb");
        field_ = std::fopen((pref + "vid_field.bin").c_str(), "wb");
    }
    ~PhaseVideo() override { close(); }

    void sample(const DiagFrame& fr) override {
        if (!phase_ || !field_ || fr.step % every_ != 0) return;
        vr_ = fr.vr;
        device_sync();
        const std::vector<float> h = phase_histogram(fr, kXBins, kVBins, kVLo, kVHi, false);
        std::fwrite(h.data(), sizeof(float), h.size(), phase_);
        const bool em = fr.f.Ez.size() > 0;
        const std::vector<float> ex = first_row(fr.f.Ex, nx_), ey = first_row(fr.f.Ey, nx_);
        const std::vector<float> ez = em ? first_row(fr.f.Ez, nx_) : std::vector<float>(nx_, 0.f);
        const std::vector<float> by = em ? first_row(fr.f.By, nx_) : std::vector<float>(nx_, 0.f);
        std::vector<float> rows(3 * static_cast<std::size_t>(nx_));
        for (int i = 0; i < nx_; ++i) {
            const double bfield = em ? by[i] : 0.0;
            rows[i] = static_cast<float>(ex[i] * fr.cost + ez[i] * fr.sint);
            rows[nx_ + i] = ey[i];
            rows[2 * nx_ + i] = static_cast<float>(bfield - fr.rp.B0[1]);
        }
        std::fwrite(rows.data(), sizeof(float), rows.size(), field_);
        ++frames_;
    }

    void finalize() override {
        close();
        std::ofstream m(meta_path_);
        m << "nframe " << frames_ << "\nnxb " << kXBins << "\nnvb " << kVBins << "\nnx " << nx_
          << "\nLx " << Lx_ << "\nvlo " << kVLo << "\nvhi " << kVHi << "\ntoff " << toff_
          << "\ndt_frame " << every_ * dt_ << "\nvr " << vr_ << "\n";
    }

private:
    void close() {
        if (phase_) std::fclose(phase_);
        if (field_) std::fclose(field_);
        phase_ = field_ = nullptr;
    }

    int nx_, every_ = 1;
    double toff_, Lx_, dt_;
    std::string meta_path_;
    std::FILE* phase_ = nullptr;
    std::FILE* field_ = nullptr;
    int frames_ = 0;
    double vr_ = 0.0;
};

}  // namespace diag
}  // namespace arc

#endif  // ARC_PIC_DIAG_PHASE_VIDEO_HPP
