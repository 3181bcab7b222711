// pic/diag/snapshot.hpp -- one field + phase-space snapshot at t_snap.
//
// Reference: Snapshot (include/pic/diag/snapshot.hpp:14-74).  At the first
// sample with t >= t_snap (t_snap defaults to the end of the run; the last
// step also fires it) it writes <pref>fields.csv (x, E_par, E_perp, Ey, dB_y,
// |dB| along row 0), <pref>fhist.bin (the 480 x 512 histogram of
// v_par/v_th in [-6, 6) by x over every particle, FP64 binning on the
// device, FP32 counts) and <pref>fhist.meta.
#ifndef ARC_PIC_DIAG_SNAPSHOT_HPP
#define ARC_PIC_DIAG_SNAPSHOT_HPP

#include "pic/deck.hpp"
#include "pic/diag/module.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

namespace arc {
namespace diag {

class Snapshot : public IDiag {
public:
    static constexpr int kXBins = 512, kVBins = 480;
    static constexpr double kVLo = -6.0, kVHi = 6.0;

    Snapshot(const Deck& d, const std::string& pref)
        : nx_(d.nx), last_step_(d.rp.nsteps), dx_(d.Lx / d.nx), Lx_(d.Lx), pref_(pref) {
        const double t_snap = d.tsnap > 0 ? d.tsnap : d.rp.nsteps * d.rp.dt;
        snap_step_ = std::lround(t_snap / d.rp.dt);
    }

    void sample(const DiagFrame& fr) override {
        if (done_ || (fr.step < snap_step_ && fr.step < last_step_)) return;
        done_ = true;
        device_sync();
        const bool em = fr.f.Ez.size() > 0;
        const std::vector<float> zero(nx_, 0.f);
        const std::vector<float> ex = first_row(fr.f.Ex, nx_), ey = first_row(fr.f.Ey, nx_);
        const std::vector<float> ez = em ? first_row(fr.f.Ez, nx_) : zero;
        const std::vector<float> bx = em ? first_row(fr.f.Bx, nx_) : zero;
        const std::vector<float> by = em ? first_row(fr.f.By, nx_) : zero;
        const std::vector<float> bz = em ? first_row(fr.f.Bz, nx_) : zero;
        {
            std::ofstream o(pref_ + "fields.csv");
            o << "x,Epar,Eperp,Ey,dBy,dBmag\n";
            for (int i = 0; i < nx_; ++i) {
                const double e_par = ex[i] * fr.cost + ez[i] * fr.sint;
                const double d0 = bx[i] - fr.rp.B0[0], d1 = by[i] - fr.rp.B0[1],
                             d2 = bz[i] - fr.rp.B0[2];
                o << i * dx_ << ',' << e_par << ',' << static_cast<double>(ey[i]) << ','
                  << ey[i] << ',' << d1 << ',' << std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) << '\n';
            }
        }
        const std::vector<float> h = phase_histogram(fr, kXBins, kVBins, kVLo, kVHi, true);
        {
            std::ofstream b(pref_ + "fhist.bin", std::ios::binary);
            b.write(reinterpret_cast<const char*>(h.data()),
                    static_cast<std::streamsize>(h.size() * sizeof(float)));
        }
        {
            std::ofstream m(pref_ + "fhist.meta");
            m << "nxb " << kXBins << "\nnvb " << kVBins << "\nLx " << Lx_ << "\nvlo " << kVLo
              << "\nvhi " << kVHi << "\nvr " << fr.vr << "\ntsnap " << fr.t << "\n";
        }
        std::printf("  snapshot at t=%.1f: %sfields.csv, %sfhist.bin (%d x %d bins, %zu "
                    "particles)\n",
                    fr.t, pref_.c_str(), pref_.c_str(), kXBins, kVBins, fr.p.n);
    }

private:
    int nx_;
    long snap_step_ = 0, last_step_;
    double dx_, Lx_;
    std::string pref_;
    bool done_ = false;
};

}  // namespace diag
}  // namespace arc

#endif  // ARC_PIC_DIAG_SNAPSHOT_HPP
