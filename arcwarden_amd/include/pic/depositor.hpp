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

    // rho and J only (:369-408): the amu outputs are left out.  Stream-ordered,
    // no allocation or synchronisation per call.
    static void charge_current(Particles& p, Sources& src, const Grid& g, const RunParams& rp,
                               hipStream_t s) {
        aw_particles v = p.views();
        aw_sources o = src.c_view();
        o.amu0 = o.amu1 = o.amu2 = o.amu3 = nullptr;
        o.dcux = o.dcuy = o.dcuz = nullptr;
        check(aw_deposit_rho_j_amu(g.ctx(), &v, rp.qm, &o, to_aw(s)));
    }
    static void charge_current_sorted(Particles& p, Sources& src, const Grid& g,
                                      const RunParams& rp, hipStream_t s) {
        charge_current(p, src, g, rp, s);
    }

    // amu only (:412-422): rho and J are left out.
    static void deposit_amu(Particles& p, Sources& src, const Grid& g, const RunParams& rp,
                            hipStream_t s) {
        aw_particles v = p.views();
        aw_sources o = src.c_view();
        o.rho = o.Jx = o.Jy = o.Jz = nullptr;
        o.dcux = o.dcuy = o.dcuz = nullptr;
        check(aw_deposit_rho_j_amu(g.ctx(), &v, rp.qm, &o, to_aw(s)));
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
