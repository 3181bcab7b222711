// pic/diag/module.hpp -- deck-selected diagnostics: what every module sees
// each sampled step, the module interface and the shared reductions.
//
// Reference: include/pic/diag/module.hpp:22-77.  Modules are plain classes
// (not templated on the configuration), so one manager serves electrostatic,
// Darwin and Yee runs.  The reference copies whole grids and particle arrays
// to the host to sum or histogram them; here those reductions run on the
// device (aw_sum_sq, aw_phase_hist, aw_gather_f32) and only their results
// cross PCIe.
#ifndef ARC_PIC_DIAG_MODULE_HPP
#define ARC_PIC_DIAG_MODULE_HPP

#include "pic/config.hpp"
#include "pic/fields.hpp"
#include "pic/grid.hpp"
#include "pic/particles.hpp"

#include <cmath>
#include <cstdint>
#include <vector>

namespace arc {
namespace diag {

// The state of one sampled step, with the B0 direction (b_x, b_z in the x-z
// plane), the reference thermal speed and the pump's resonant speed
// precomputed by the manager.
struct DiagFrame {
    const Fields& f;
    const Particles& p;
    const Grid& g;
    const RunParams& rp;
    long step;
    double t;
    double bmag;       // |B0| (0: unmagnetised)
    double cost, sint; // B0 direction: b_x, b_z
    double vth;        // species[0].uth[0]
    double vr;         // pump phase speed along B0 over vth (0 without a pump)
};

class IDiag {
public:
    virtual void sample(const DiagFrame& fr) = 0;
    virtual void finalize() {}
    virtual ~IDiag() = default;
};

// Every module reads state that the simulation's stream produced.
inline void device_sync() { check(aw_device_synchronize()); }

// FP64 sum of (a[i] - offset)^2, reduced on the device.
inline double sum_sq(const DeviceArray<float>& a, double offset = 0.0) {
    double r = 0.0;
    check(aw_sum_sq(a.data(), static_cast<int64_t>(a.size()), offset, &r, nullptr));
    return r;
}

// The first nx values of a grid (row j = 0) on the host.
inline std::vector<float> first_row(const DeviceArray<float>& a, int nx) {
    std::vector<float> h(static_cast<std::size_t>(nx), 0.0f);
    if (a.size() >= static_cast<std::size_t>(nx))
        check(aw_copy(h.data(), a.data(), h.size() * sizeof(float)));
    return h;
}

// |F_m|^2 / nx^2 of a row, F_m = sum_i row_i exp(-2 pi i m i / nx) (direct DFT).
inline double mode_power(const std::vector<float>& row, int nx, int m) {
    const double k = 2.0 * M_PI * m / nx;
    double re = 0.0, im = 0.0;
    for (int i = 0; i < nx; ++i) {
        const double ph = k * i;
        re += row[i] * std::cos(ph);
        im -= row[i] * std::sin(ph);
    }
    return (re * re + im * im) / (static_cast<double>(nx) * nx);
}

// f(x, v_par/v_th) of all particles as float counts (row-major [nvb][nxb]).
// fp64 selects the double-precision binning of the snapshot module.
inline std::vector<float> phase_histogram(const DiagFrame& fr, int nxb, int nvb, double vlo,
                                          double vhi, bool fp64) {
    const std::size_t nb = static_cast<std::size_t>(nxb) * nvb;
    DeviceArray<uint32_t> counts(nb);
    counts.zero();
    aw_particles p{};
    p.x = const_cast<float*>(fr.p.x.data());
    p.y = const_cast<float*>(fr.p.y.data());
    p.ux = const_cast<float*>(fr.p.ux.data());
    p.uy = const_cast<float*>(fr.p.uy.data());
    p.uz = const_cast<float*>(fr.p.uz.data());
    p.n = static_cast<int64_t>(fr.p.n);
    const aw_grid g = fr.g.to_c();
    check(aw_phase_hist(&p, &g, nxb, nvb, fr.cost, fr.sint, fr.vth, vlo, vhi, fp64 ? 1 : 0,
                        counts.data(), nullptr));
    std::vector<uint32_t> h(nb);
    check(aw_copy(h.data(), counts.data(), nb * sizeof(uint32_t)));
    return std::vector<float>(h.begin(), h.end());
}

}  // namespace diag
}  // namespace arc

#endif  // ARC_PIC_DIAG_MODULE_HPP
