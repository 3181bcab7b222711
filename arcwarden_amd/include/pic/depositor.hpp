// pic/depositor.hpp -- charge and current deposits.
//
// Reference: Depositor<Cfg> (include/pic/depositor.hpp:34-496): static
// operators with the signature (Particles&, Sources&, const Grid&,
// const RunParams&, stream).  Every deposit accumulates (callers zero the
// grids, as in the reference) and sums int64 fixed-point quanta, so the
// result does not depend on the policy or on thread order.
#ifndef ARC_PIC_DEPOSITOR_HPP
#define ARC_PIC_DEPOSITOR_HPP

#include "pic/config.hpp"
#include "pic/fields.hpp"
#include "pic/grid.hpp"
#include "pic/particles.hpp"
#include "pic/sources.hpp"

namespace arc {

template <class Cfg = arc::Cfg>
struct Depositor {
    // rho += qm w/(dx dy) W_cic with Cfg's deposit policy (:440-496).
    static void charge(Particles& p, Sources& src, const Grid& g, const RunParams& rp,
                       hipStream_t s) {
        aw_particles v = p.views();
        check(aw_deposit_rho(g.ctx(), &v, rp.qm, Cfg::aw_deposit, src.rho.data(), to_aw(s)));
    }

    // rho, J, amu of the current state in one pass (charge_current +
    // deposit_amu, :160-270).  src must be allocate_em()'d.
    static void charge_current_amu(Particles& p, Sources& src, const Grid& g,
                                   const RunParams& rp, hipStream_t s) {
        aw_particles v = p.views();
        aw_sources o = src.c_view();
        check(aw_deposit_rho_j_amu(g.ctx(), &v, rp.qm, &o, to_aw(s)));
    }

    // rho and J only (:369-408); amu goes to a scratch block.
    static void charge_current(Particles& p, Sources& src, const Grid& g, const RunParams& rp,
                               hipStream_t s) {
        DeviceArray<float> scratch(4 * static_cast<std::size_t>(g.real_size()));
        scratch.zero(s);
        aw_particles v = p.views();
        aw_sources o = src.c_view();
        const std::size_t n = static_cast<std::size_t>(g.real_size());
        o.amu0 = scratch.data();
        o.amu1 = scratch.data() + n;
        o.amu2 = scratch.data() + 2 * n;
        o.amu3 = scratch.data() + 3 * n;
        check(aw_deposit_rho_j_amu(g.ctx(), &v, rp.qm, &o, to_aw(s)));
        stream_synchronize(s);
    }
    static void charge_current_sorted(Particles& p, Sources& src, const Grid& g,
                                      const RunParams& rp, hipStream_t s) {
        charge_current(p, src, g, rp, s);
    }

    // amu only (:412-422); rho and J go to a scratch block.
    static void deposit_amu(Particles& p, Sources& src, const Grid& g, const RunParams& rp,
                            hipStream_t s) {
        DeviceArray<float> scratch(4 * static_cast<std::size_t>(g.real_size()));
        scratch.zero(s);
        aw_particles v = p.views();
        aw_sources o = src.c_view();
        const std::size_t n = static_cast<std::size_t>(g.real_size());
        o.rho = scratch.data();
        o.Jx = scratch.data() + n;
        o.Jy = scratch.data() + 2 * n;
        o.Jz = scratch.data() + 3 * n;
        check(aw_deposit_rho_j_amu(g.ctx(), &v, rp.qm, &o, to_aw(s)));
        stream_synchronize(s);
    }

    // dcu += qm^2 w/(dx dy) (E + u x B) W with E = f.Ex, Ey, Ez (E_L + pump)
    // and B = f.Bx, By, Bz (:275-324, 426-437).
    static void deposit_dcu(Particles& p, Sources& src, const Fields& f, const Grid& g,
                            const RunParams& rp, hipStream_t s) {
        aw_particles v = p.views();
        aw_fields fv = const_cast<Fields&>(f).c_view();
        aw_sources o = src.c_view();
        check(aw_deposit_dcu(g.ctx(), &v, rp.qm, &fv, &o, to_aw(s)));
    }

    // Esirkepov charge-conserving current of the moves (x0, y0) -> p's
    // positions into src.Jx (staggered i+1/2), Jy (j+1/2), Jz (nodal)
    // (include/pic/yee2d.hpp:310-469, the deposit half).
    static void esirkepov(const float* x0, const float* y0, Particles& p, Sources& src,
                          const Grid& g, const RunParams& rp, hipStream_t s) {
        aw_particles v = p.views();
        check(aw_deposit_esirkepov(g.ctx(), x0, y0, &v, rp.qm, rp.dt, Cfg::aw_deposit,
                                   src.Jx.data(), src.Jy.data(), src.Jz.data(), to_aw(s)));
    }
};

}  // namespace arc

#endif  // ARC_PIC_DEPOSITOR_HPP
