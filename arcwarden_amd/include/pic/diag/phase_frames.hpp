// pic/diag/phase_frames.hpp -- (x, vx) phase-space CSV frames and their
// manifest.
//
// Reference: PhaseFrames (include/pic/diag/phase_frames.hpp:20-62).  Every
// dump_every steps (default nsteps/100) it writes <outdir>/frame_%04d.csv
// ("x,vx", x in physical units) and appends "frame,step,time,file" to
// <outdir>/manifest.csv.  Runs with more than 100000 particles write a fixed
// sample: the indices 0..N-1 shuffled by std::mt19937(12345), the first
// 100000 kept and sorted -- the reference's sampling rule.  The sample is
// gathered on the device, so only 8 bytes per sampled particle are copied.
#ifndef ARC_PIC_DIAG_PHASE_FRAMES_HPP
#define ARC_PIC_DIAG_PHASE_FRAMES_HPP

#include "pic/deck.hpp"
#include "pic/diag/module.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <numeric>
#include <random>
#include <string>
#include <vector>

namespace arc {
namespace diag {

// The fixed particle sample of a frame: at most cap indices of [0, n).
inline std::vector<int32_t> frame_sample(std::size_t n, std::size_t cap) {
    std::vector<int32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    if (n > cap) {
        std::mt19937 gen(12345);
        std::shuffle(idx.begin(), idx.end(), gen);
        idx.resize(cap);
        std::sort(idx.begin(), idx.end());
    }
    return idx;
}

class PhaseFrames : public IDiag {
public:
    static constexpr std::size_t kCap = 100000;

    PhaseFrames(const Deck& d, const std::string& outdir)
        : every_(d.dump_every > 0 ? d.dump_every
                                  : (d.rp.nsteps / 100 > 0 ? d.rp.nsteps / 100 : 1)),
          dir_(outdir.empty() ? "." : outdir) {
        std::filesystem::create_directories(dir_);
        manifest_.open(dir_ + "/manifest.csv");
        if (manifest_) manifest_ << "frame,step,time,file\n";
    }

    void sample(const DiagFrame& fr) override {
        if (fr.step % every_ != 0) return;
        const std::size_t n = fr.p.n;
        if (idx_host_.empty() && n > 0) {
            idx_host_ = frame_sample(n, kCap);
            idx_dev_ = DeviceArray<int32_t>(idx_host_.size());
            idx_dev_.from_host(idx_host_);
            gx_ = DeviceArray<float>(idx_host_.size());
            gv_ = DeviceArray<float>(idx_host_.size());
        }
        device_sync();
        const int64_t m = static_cast<int64_t>(idx_host_.size());
        std::vector<float> x(m), vx(m);
        if (m > 0) {
            check(aw_gather_f32(fr.p.x.data(), idx_dev_.data(), m, gx_.data(), nullptr));
            check(aw_gather_f32(fr.p.ux.data(), idx_dev_.data(), m, gv_.data(), nullptr));
            check(aw_copy(x.data(), gx_.data(), m * sizeof(float)));
            check(aw_copy(vx.data(), gv_.data(), m * sizeof(float)));
        }
        char name[32];
        std::snprintf(name, sizeof name, "frame_%04d.csv", frame_);
        {
            std::ofstream o(dir_ + "/" + name);
            o << "x,vx\n";
            const float fnx = static_cast<float>(fr.g.nx);
            for (int64_t k = 0; k < m; ++k) {
                float xc = x[k];
                // a tiny negative x + nx rounds to nx: the second test maps it to 0
                if (xc < 0.0f) xc += fnx;
                if (xc >= fnx) xc -= fnx;
                o << xc * fr.g.dx << ',' << vx[k] << '\n';
            }
        }
        if (manifest_) {
            manifest_ << frame_ << ',' << fr.step << ',' << fr.t << ',' << name << '\n';
            manifest_.flush();
        }
        ++frame_;
    }

    int frames() const { return frame_; }

private:
    long every_;
    std::string dir_;
    std::ofstream manifest_;
    int frame_ = 0;
    std::vector<int32_t> idx_host_;
    DeviceArray<int32_t> idx_dev_;
    DeviceArray<float> gx_, gv_;
};

}  // namespace diag
}  // namespace arc

#endif  // ARC_PIC_DIAG_PHASE_FRAMES_HPP
