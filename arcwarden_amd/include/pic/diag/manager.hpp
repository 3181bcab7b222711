// pic/diag/manager.hpp -- builds the diagnostics a deck enables and fans
// every sample and the final flush out to them.
//
// Reference: DiagManager (include/pic/diag/manager.hpp:24-61).  The keys of
// [diagnostics] enable = ... are spectrum, kt, phase_video, snapshot,
// em_energy and phase_frames; an unknown key is reported on stderr and
// skipped.  The manager derives, once, the B0 direction in the x-z plane, the
// reference thermal speed (species[0].uth[0]) and the pump's resonant speed
// along B0 in units of it.
#ifndef ARC_PIC_DIAG_MANAGER_HPP
#define ARC_PIC_DIAG_MANAGER_HPP

#include "pic/deck.hpp"
#include "pic/diag/em_energy.hpp"
#include "pic/diag/kt.hpp"
#include "pic/diag/module.hpp"
#include "pic/diag/phase_frames.hpp"
#include "pic/diag/phase_video.hpp"
#include "pic/diag/snapshot.hpp"
#include "pic/diag/spectrum.hpp"

#include <cmath>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

namespace arc {
namespace diag {

class DiagManager {
public:
    // pref: output file prefix ("<outdir>/<name>_" or "<name>_"); outdir: the
    // directory of the phase frames and manifest ("" = the working directory).
    DiagManager(const Deck& d, const RunParams& rp, const std::string& pref,
                const std::string& outdir) {
        const double b0x = rp.B0[0], b0y = rp.B0[1], b0z = rp.B0[2];
        bmag_ = std::sqrt(b0x * b0x + b0y * b0y + b0z * b0z);
        cost_ = bmag_ > 0 ? b0x / bmag_ : 1.0;
        sint_ = bmag_ > 0 ? b0z / bmag_ : 0.0;
        vth_ = d.species.empty() ? 1.0 : d.species[0].uth[0];
        vr_ = (rp.pump_k0 > 0 && cost_ > 0) ? rp.pump_w0 / (rp.pump_k0 * cost_) / vth_ : 0.0;
        for (const std::string& key : d.diag_enable) {
            std::unique_ptr<IDiag> m = make(key, d, pref, outdir);
            if (m) modules_.push_back(std::move(m));
            else std::fprintf(stderr, "DiagManager: unknown diagnostic '%s' (ignored)\n",
                              key.c_str());
        }
    }

    bool empty() const { return modules_.empty(); }
    std::size_t size() const { return modules_.size(); }

    void sample(const Fields& f, const Particles& p, const Grid& g, const RunParams& rp,
                long step, double t) {
        if (modules_.empty()) return;
        const DiagFrame fr{f, p, g, rp, step, t, bmag_, cost_, sint_, vth_, vr_};
        for (auto& m : modules_) m->sample(fr);
    }

    void finalize() {
        for (auto& m : modules_) m->finalize();
    }

private:
    static std::unique_ptr<IDiag> make(const std::string& key, const Deck& d,
                                       const std::string& pref, const std::string& outdir) {
        if (key == "spectrum") return std::make_unique<SpectrumDiag>(d);
        if (key == "kt") return std::make_unique<KTDump>(d, pref);
        if (key == "phase_video") return std::make_unique<PhaseVideo>(d, pref);
        if (key == "snapshot") return std::make_unique<Snapshot>(d, pref);
        if (key == "em_energy") return std::make_unique<EMEnergy>(d, pref);
        if (key == "phase_frames") return std::make_unique<PhaseFrames>(d, outdir);
        return nullptr;
    }

    std::vector<std::unique_ptr<IDiag>> modules_;
    double bmag_ = 0.0, cost_ = 1.0, sint_ = 0.0, vth_ = 1.0, vr_ = 0.0;
};

}  // namespace diag
}  // namespace arc

#endif  // ARC_PIC_DIAG_MANAGER_HPP
