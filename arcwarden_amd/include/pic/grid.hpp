// pic/grid.hpp -- real-space geometry.
//
// Reference: Grid (include/pic/grid.hpp:44-91): nx, ny, dx = Lx/nx, dy = Ly/ny,
// row-major [ny][nx] with x contiguous, periodic wraps.  The facade's Grid
// also owns (shared between copies) the library context of its shape:
// rocFFT plans, k-space workspace and deposit / sort scratch, created on
// first use.  Operations on one grid from several streams at once share
// that scratch and must not overlap.
#ifndef ARC_PIC_GRID_HPP
#define ARC_PIC_GRID_HPP

#include "pic/hip_utils.hpp"

#include <memory>

namespace arc {

inline constexpr double kTwoPi = 6.28318530717958647692;

struct Grid {
    int nx = 0, ny = 0;
    double dx = 0.0, dy = 0.0;
    double Lx = 0.0, Ly = 0.0;

    Grid() = default;
    Grid(int nx_, int ny_, double Lx_, double Ly_)
        : nx(nx_), ny(ny_), dx(Lx_ / nx_), dy(Ly_ / ny_), Lx(Lx_), Ly(Ly_) {}

    int idx(int i, int j) const { return j * nx + i; }
    int real_size() const { return nx * ny; }
    int wrap_x(int i) const { return i >= nx ? i - nx : (i < 0 ? i + nx : i); }
    int wrap_y(int j) const { return j >= ny ? j - ny : (j < 0 ? j + ny : j); }
    int idx_periodic(int i, int j) const { return idx(wrap_x(i), wrap_y(j)); }

    aw_grid to_c() const { return aw_grid{nx, ny, dx, dy, Lx, Ly}; }

    // The library context of this shape (created on first use, tile 16x16).
    aw_ctx* ctx() const {
        if (!ctx_) {
            const aw_grid g = to_c();
            aw_ctx* c = nullptr;
            check(aw_ctx_create(&g, 0, 0, &c));
            ctx_ = std::shared_ptr<aw_ctx>(c, [](aw_ctx* p) { aw_ctx_destroy(p); });
        }
        return ctx_.get();
    }

private:
    mutable std::shared_ptr<aw_ctx> ctx_;
};

}  // namespace arc

#endif  // ARC_PIC_GRID_HPP
