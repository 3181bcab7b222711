// pic/pusher.hpp -- Boris pushers.
//
// Reference: Pusher<Cfg> (include/pic/pusher.hpp:112-330).  boris and
// half_step_back gather Ex, Ey at the particle and apply the uniform-B0
// Boris step (Cfg::has_b0); boris_em gathers E and B and applies the full
// rotation, then moves and wraps (the reference wraps in a separate
// Particles::migrate, which stays valid to call afterwards).
#ifndef ARC_PIC_PUSHER_HPP
#define ARC_PIC_PUSHER_HPP

#include "pic/config.hpp"
#include "pic/fields.hpp"
#include "pic/grid.hpp"
#include "pic/particles.hpp"

namespace arc {

template <class Cfg = arc::Cfg>
struct Pusher {
    // kick, rotate, kick, x += u dt/dx (no wrap; Particles::migrate wraps)
    static void boris(Particles& p, const Fields& f, const Grid& g, const RunParams& rp,
                      hipStream_t s) {
        aw_particles v = p.views();
        check(aw_push_es(g.ctx(), &v, f.Ex.data(), f.Ey.data(), &rp, Cfg::has_b0 ? 1 : 0,
                         to_aw(s)));
    }
    static void boris_tiled(Particles& p, const Fields& f, const Grid& g, const RunParams& rp,
                            hipStream_t s) {
        boris(p, f, g, rp, s);
    }
    // the same update with dt -> -dt/2 and no move (leapfrog start)
    static void half_step_back(Particles& p, const Fields& f, const Grid& g,
                               const RunParams& rp, hipStream_t s) {
        aw_particles v = p.views();
        check(aw_half_step_back(g.ctx(), &v, f.Ex.data(), f.Ey.data(), &rp,
                                Cfg::has_b0 ? 1 : 0, to_aw(s)));
    }
    // full Boris with gathered E and B, move, periodic wrap
    static void boris_em(Particles& p, const Fields& f, const Grid& g, const RunParams& rp,
                         hipStream_t s) {
        aw_particles v = p.views();
        aw_fields fv = const_cast<Fields&>(f).c_view();
        check(aw_push_boris_em(g.ctx(), &v, &fv, &rp, to_aw(s)));
    }
};

}  // namespace arc

#endif  // ARC_PIC_PUSHER_HPP
