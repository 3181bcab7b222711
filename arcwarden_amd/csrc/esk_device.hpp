// esk_device.hpp -- the Esirkepov current's per-particle device code and
// its host set-up, shared by the deposit kernels (esirkepov.hip) and by
// kernels that deposit the current of the moves they make (yee.hip).
//
// Reference: s1_shape4, esirkepov_scatter (include/pic/yee2d.hpp:124-132,
// 310-389).  Contributions are int64 fixed point with one scale exponent per
// component, taken from a bound on the largest single contribution.
#pragma once
#include "aw_internal.hpp"

namespace aw::esk {

constexpr int kFixBits = 44;
constexpr int kSortedApron = 4;  // particles in the tile order of an earlier sort

struct EskArgs {
    const float *x0, *y0, *x1, *y1, *uz, *w;
    float *xw, *yw;  // optional: x1, y1 wrapped into the box, written back
    int n, nx, ny;
    int per_w;       // per-particle deposit weights (w, or the caller's own)
    double qm, w_uniform, dt, dx, dy;
    const unsigned int* bound;  // [0] max|w| (or 0), [1] max|w uz| (|uz| for uniform w)
};

// Scale exponents of the three components (identical in deposit and convert).
__device__ __forceinline__ void esk_exponents(const EskArgs& a, int S[3]) {
    const double wmax = a.per_w ? (double)__uint_as_float(a.bound[0]) : fabs(a.w_uniform);
    const double wuz = a.per_w ? (double)__uint_as_float(a.bound[1])
                               : fabs(a.w_uniform) * (double)__uint_as_float(a.bound[1]);
    const double b[3] = {fabs(a.qm) * wmax / (a.dt * a.dy), fabs(a.qm) * wmax / (a.dt * a.dx),
                         fabs(a.qm) * wuz / (a.dx * a.dy)};
    for (int q = 0; q < 3; ++q) {
        int e = 0;
        (void)frexp(b[q], &e);  // b < 2^e
        S[q] = kFixBits - e;
    }
}

// the current of accumulated quanta v of component q: v 2^-S_q
__device__ __forceinline__ float esk_value(long long v, int Sq) {
    return (float)ldexp((double)v, -Sq);
}

__device__ __forceinline__ void esk_scales(const EskArgs& a, double scale[3]) {
    int S[3];
    esk_exponents(a, S);
    for (int q = 0; q < 3; ++q) scale[q] = ldexp(1.0, S[q]);
}

// One particle's move: the shared stencil geometry plus its fixed-point
// scaled coefficients.
struct EskMove {
    dev::EskShape sh;
    double cJx, cJy, cJz;
};

// The coefficients' constant factors, once per thread: kc = (2^Sx/(dt dy),
// 2^Sy/(dt dx), 2^Sz/(dx dy)), so a particle's coefficients are products.
__device__ __forceinline__ void esk_coefs(const EskArgs& a, double kc[3]) {
    double scale[3];
    esk_scales(a, scale);
    kc[0] = scale[0] / (a.dt * a.dy);
    kc[1] = scale[1] / (a.dt * a.dx);
    kc[2] = scale[2] / (a.dx * a.dy);
}

// the move (x0, y0) -> (x1, y1) with deposit weight w and velocity uz:
// cJx = -q w kc0, cJy = -q w kc1, cJz = q w uz kc2
__device__ __forceinline__ void esk_prepare_v(const EskArgs& a, float x0, float y0, float x1,
                                              float y1, float uz, double w, const double kc[3],
                                              EskMove& m) {
    dev::esk_shape(x0, y0, x1, y1, a.nx, a.ny, m.sh);
    const double qw = a.qm * w;
    m.cJx = -qw * kc[0];
    m.cJy = -qw * kc[1];
    m.cJz = qw * (double)uz * kc[2];
}

__device__ __forceinline__ void esk_prepare(const EskArgs& a, int i, const double kc[3],
                                            EskMove& m) {
    esk_prepare_v(a, a.x0[i], a.y0[i], a.x1[i], a.y1[i], a.uz[i],
                  a.w ? (double)a.w[i] : a.w_uniform, kc, m);
}

// The contributions of one move.  On a one-row grid (ny = 1) the three
// stencil rows are the same row: each (component, column) sum of the
// fixed-point contributions is formed in registers (exact, int64) and added
// once, at row 1, instead of once per row.
template <bool ONE_ROW, class Add>
__device__ __forceinline__ void esk_emit(const EskMove& m, Add&& add) {
    if constexpr (!ONE_ROW) {
        dev::esk_emit(m.sh, m.cJx, m.cJy, m.cJz, add);
    } else {
        unsigned long long col[3][4] = {};
        dev::esk_emit(m.sh, m.cJx, m.cJy, m.cJz,
                      [&](int q, int k, int, unsigned long long v) { col[q][k] += v; });
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int k = 1; k < 4; ++k)
                if (col[q][k]) add(q, k, 1, col[q][k]);
    }
}

__device__ __forceinline__ int wrap_local(int l, int n) {
    if (l < 0) l += n;
    else if (l >= n) l -= n;
    return l;
}

// Host side.  args fills the kernels' arguments; setup also clears the
// accumulators and fixes the scale bound (from a bound pass, or the caller's
// fixed values); convert adds the accumulated current to jx, jy, jz.
EskArgs args(aw_ctx& c, const float* x0, const float* y0, const aw_particles& p, double qm,
             double dt, const EskExtra& ex);
EskArgs setup(aw_ctx& c, const float* x0, const float* y0, const aw_particles& p, double qm,
              double dt, hipStream_t s, const EskExtra& ex);
void convert(aw_ctx& c, const EskArgs& a, float* jx, float* jy, float* jz, hipStream_t s);

}  // namespace aw::esk
