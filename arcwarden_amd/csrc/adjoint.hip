// adjoint.hip -- discrete adjoint of the electrostatic run (unmagnetised or
// with B0 along z) with checkpointing.
//
// No reference counterpart (the reference has no gradients).  The forward
// run is the Simulation step of this library (fused gather + push + move +
// wrap + fixed-point deposit, spectral Poisson solve); the loss is the field
// energy EE = eps0/2 sum |E^N|^2 dx dy at the last step.  The reverse sweep
// is the transpose of every linear piece, step by step:
//   x^{n+1} = x^n + c u^{n+1/2}        ->  lu += c lx
//   u^{n+1/2} = R (u^{n-1/2} + h e) + h e   (Boris, e = e(x^n), h = qm dt/2,
//                                          R the rotation of B0 z; R = I without B0)
//                                      ->  le = h (I + R^T) lu;  lu = R^T lu;
//                                          lx += (de/dx)^T le;
//                                          lE^n = gather^T le (a deposit)
//   E^n = S rho^n,  rho^n = D(x^n)     ->  lrho = S^T lE;  lx += D_x^T lrho
//                                          (a gather with CIC derivative weights)
// with c = dt/dx, S^T = IFFT(conj(H) FFT(.)), H = -i k/(eps0 k^2); the
// leapfrog half step back (the same Boris form with dt -> -dt/2) is
// transposed after step 0.  B0 along z leaves uz out of the run.
//
// Checkpoints every K steps hold (x, y, ux, uy) and the tile order: the
// particles are re-sorted by tile at every checkpoint, each reverse segment
// recomputes its K forward steps from the checkpoint (keeping x^n and E^n),
// and the adjoint state follows the sort permutations back.  Gradients with
// respect to the x-displacement mode amplitudes and the species drifts are
// reductions of the adjoint of the initial state.
//
// The adjoint field lE^n = gather^T le is summed in int64 fixed point, like
// rho, so the sweep returns the same bits on every run.  Its scale 2^e is a
// pure function of order-free maxima: the largest |component| of the adjoint
// particle state entering the step (each kernel that writes the state
// atomicMax-es the FP32 bit pattern of its outputs' maximum) and the largest
// |Re|+|Im| of the lrho spectrum.  From them the step bounds |le| by B, and e
// is chosen so that n * B * 2^e < 2^62: no cell sum can wrap.
#include "aw_internal.hpp"

#include <cmath>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <utility>
#include <vector>

namespace aw {

void init_fft(aw_ctx& c);
void solve_es_contig(aw_ctx& c, const float* rho, float* E2, double eps0, hipStream_t s);
int push_deposit_es(aw_ctx& c, const aw_particles& p, const int32_t* off, const float* Ex,
                    const float* Ey, const aw_run_params& rp, bool has_b0,
                    unsigned long long* strays, double* ke_part, int bpt, hipStream_t s);
int push_deposit_es_save(aw_ctx& c, const aw_particles& p, const int32_t* off, const float* Ex,
                         const float* Ey, const aw_run_params& rp, bool has_b0,
                         unsigned long long* strays, double* ke_part, int bpt, float* save_xy,
                         hipStream_t s);

namespace {

constexpr int kThreads = 256;

inline int blocks_for(int64_t n) { return (int)((n + kThreads - 1) / kThreads); }

struct KG {
    int nx, ny, nkx;
    double dkx, dky, eps0;
};

// S^T in k-space: lrho_k = i G (kx lEx_k + ky lEy_k), G = 1/(eps0 k^2);
// DC and Nyquist zero (the transpose of es_kspace_kernel).
// Order-free maxima (FP32 bit patterns of non-negative values, unsigned order
// = value order) are kept in kBoundSlots slots, one 128-byte line each: a
// wave takes the max with atomicMax on its block's slot, and readers take the
// max over the slots.  One shared address would serialise the thousands of
// wave atomics of a 1 M-particle launch in L2 (about 0.2 ms).
constexpr int kBoundSlots = 64, kSlotStride = 32;
constexpr int kBoundWords = kBoundSlots * kSlotStride;

__device__ __forceinline__ void wave_max_to(float v, unsigned int* slots) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0 && v > 0.f)
        atomicMax(slots + (blockIdx.x % kBoundSlots) * kSlotStride, __float_as_uint(v));
}

__device__ __forceinline__ float slots_max(const unsigned int* slots) {
    unsigned int m = 0u;
    for (int k = 0; k < kBoundSlots; ++k) m = max(m, slots[k * kSlotStride]);
    return __uint_as_float(m);
}

__global__ void es_adj_kspace_kernel(const float2* __restrict__ ak, float2* __restrict__ rk,
                                     KG k, unsigned int* __restrict__ rmax) {
    const int64_t nk = (int64_t)k.nkx * k.ny;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t0 < nk;
    const int64_t t = live ? t0 : 0;  // lanes past the end recompute mode 0 and store nothing
    const int i = (int)(t % k.nkx), j = (int)(t / k.nkx);
    const int m = (j <= k.ny / 2) ? j : j - k.ny;
    const double kx = i * k.dkx, ky = m * k.dky;
    const double k2 = kx * kx + ky * ky;
    const double g = k2 > 0.0 ? 1.0 / (k.eps0 * k2) : 0.0;
    const bool nyq = ((k.nx % 2 == 0) && i == k.nx / 2) || ((k.ny % 2 == 0) && j == k.ny / 2);
    const float2 ax = ak[t], ay = ak[nk + t];
    float2 r = make_float2(0.f, 0.f);
    if (!nyq) {
        r.x = (float)(-g * (kx * (double)ax.y + ky * (double)ay.y));
        r.y = (float)(g * (kx * (double)ax.x + ky * (double)ay.x));
    }
    if (live) rk[t] = r;
    wave_max_to(live ? fabsf(r.x) + fabsf(r.y) : 0.f, rmax);
}

// Order-free bounds of one reverse step (slotted maxima, as above): smax[2]
// alternate between steps (the state entering step st is bounded by
// smax[(st+1)&1], the step's outputs go to smax[st&1]); rmax bounds the lrho
// spectrum; e is the step's exponent.
struct AdjBound {
    unsigned int smax[2][kBoundWords];
    unsigned int rmax[kBoundWords];
    int e;
};

// int64 adjoint-field accumulators (scale 2^e) -> float FFT input; clears the
// accumulators, the state slot the finished step read, and the spectrum bound
// that the following k-space pass rewrites
__global__ void lE_convert_kernel(unsigned long long* __restrict__ acc, float* __restrict__ out,
                                  int64_t n, AdjBound* __restrict__ bd, int free_slot) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < kBoundSlots) {  // the grid has at least one block of kThreads >= kBoundSlots
        bd->smax[free_slot][c * kSlotStride] = 0u;
        bd->rmax[c * kSlotStride] = 0u;
    }
    if (c >= n) return;
    const double inv = ldexp(1.0, -bd->e);
    out[c] = (float)((double)(long long)acc[c] * inv);
    acc[c] = 0ull;
}

// lE = eps0 dx dy E (the loss seed, as the FFT's FP32 input) and the
// per-block partial sums of |E|^2
__global__ void loss_kernel(const float* __restrict__ E2, int64_t nc, double w,
                            float* __restrict__ lE, double* __restrict__ part) {
    __shared__ double sh[kThreads];
    double s = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x; c < 2 * nc;
         c += (int64_t)gridDim.x * kThreads) {
        const double v = E2[c];
        s += v * v;
        lE[c] = (float)(w * v);
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

struct AdjArgs {
    const float* xn;   // x^n (this step's gather / deposit position)
    const float* yn;
    const float* xb;   // x^{n+1} for the fused D_x^T part (nullptr: skip)
    const float* yb;
    float *lx, *ly, *lux, *luy;
    const float* Ex;   // E^n
    const float* Ey;
    const float* lrho;  // lrho^{n+1} (with xb)
    unsigned long long* lE;  // [lEx | lEy] int64 accumulators of step n
    AdjBound* bd;       // bounds: state entering the step, lrho spectrum; e out
    int slot_in;        // smax slot bounding the entry state; outputs go to slot_in ^ 1
    float rfac;         // |lrho(x)| <= rfac * max(|Re|+|Im|) of its spectrum
    int log2n;          // ceil(log2(n))
    const int32_t* off;
    int ntx, tx, ty, W, H, bpt;
    int nx, ny, n;
    float cx, cy, coef;
    float h, rc, rs;     // the step's Boris: u' = R (u + h e) + h e, R = [[rc, rs], [-rs, rc]]
    float h0, rc0, rs0;  // the half step back, transposed after step 0 (h0 = 0: none)
};

struct Cic {
    int i0, j0;
    float fx, fy;
};

__device__ __forceinline__ Cic cic_of(float x, float y) {
    Cic c;
    const float fx0 = floorf(x), fy0 = floorf(y);
    c.i0 = (int)fx0;
    c.j0 = (int)fy0;
    c.fx = x - fx0;
    c.fy = y - fy0;
    return c;
}

__device__ __forceinline__ int local(int i, int org, int n) {
    int l = i - org;
    if (l < 0) l += n;
    else if (l >= n) l -= n;
    return l;
}

// the four corner values of a grid f around (i0, j0): tile when inside, else global
template <class T, class TileT>
__device__ __forceinline__ void corners(const T* g, const TileT* tile, int W, int H, int orgx,
                                        int orgy, int nx, int ny, const Cic& c, float v[4]) {
    const int li = local(c.i0, orgx, nx), lj = local(c.j0, orgy, ny);
    if (dev::in_tile(li, W - 2) && dev::in_tile(lj, H - 2)) {
        const int b = lj * W + li;
        v[0] = (float)tile[b];
        v[1] = (float)tile[b + 1];
        v[2] = (float)tile[b + W];
        v[3] = (float)tile[b + W + 1];
    } else {
        int i0, i1, j0, j1;
        dev::wrap_pair(c.i0, nx, i0, i1);
        dev::wrap_pair(c.j0, ny, j0, j1);
        v[0] = (float)g[j0 * nx + i0];
        v[1] = (float)g[j0 * nx + i1];
        v[2] = (float)g[j1 * nx + i0];
        v[3] = (float)g[j1 * nx + i1];
    }
}

// d/dx and d/dy of sum_k W_k v_k (cell units)
__device__ __forceinline__ void grad4(const Cic& c, const float v[4], float& dx, float& dy) {
    const float gx = 1.f - c.fx, gy = 1.f - c.fy;
    dx = gy * (v[1] - v[0]) + c.fy * (v[3] - v[2]);
    dy = gx * (v[2] - v[0]) + c.fx * (v[3] - v[1]);
}

// One reverse step n (with the D_x^T part of step n+1 when xb is set), over
// a tile's particles: LDS holds E^n, lrho^{n+1} and the lE^n accumulators.
__device__ __forceinline__ float max4abs(float a, float b, float c, float d) {
    return fmaxf(fmaxf(fabsf(a), fabsf(b)), fmaxf(fabsf(c), fabsf(d)));
}

// block-wide max of non-negative v, one atomicMax per wavefront

// Exponent e of the step's fixed-point scale: every contribution le * W is
// bounded by B = |a| (U + c (U + |coef| 2 R)), U the entry-state bound and R
// the lrho bound (2 R bounds a CIC gradient of lrho), and e makes
// 2^log2n * B * 2^e < 2^62 and B * 2^e < 2^50.  The FP32 rounding of the
// products stays inside the 1e-3 margin.
__device__ __forceinline__ int adj_exponent(const AdjArgs& a) {
    const double U = (double)slots_max(a.bd->smax[a.slot_in]);
    const double R = a.xb ? (double)slots_max(a.bd->rmax) * (double)a.rfac : 0.0;
    const double c = fmax(fabs((double)a.cx), fabs((double)a.cy));
    // |le| <= (|h| + |h0|) (1 + sqrt 2) sqrt 2 |lu'| (rotations keep the 2-norm)
    const double B = 4.0 * (fabs((double)a.h) + fabs((double)a.h0)) *
                     (U + c * (U + fabs((double)a.coef) * 2.0 * R)) * (1.0 + 1e-3);
    if (!(B > 0.0) || !(B < 1e300)) return 0;
    int k;
    (void)frexp(B, &k);  // B < 2^k
    // each contribution below 2^50 (the quantiser's range), their sum below 2^62
    const int e = (62 - a.log2n < 50 ? 62 - a.log2n : 50) - k;
    return e > 1000 ? 1000 : (e < -1000 ? -1000 : e);
}

// One reverse step n (with the D_x^T part of step n+1 when xb is set), over
// a tile's particles: LDS holds E^n, lrho^{n+1} and the int64 lE^n tiles.
__global__ void __launch_bounds__(kThreads) adj_step_kernel(AdjArgs a) {
    extern __shared__ unsigned long long smem[];
    const int W = a.W, H = a.H, R = W * H;
    unsigned long long* s_lex = smem;
    unsigned long long* s_ley = smem + R;
    float* s_ex = reinterpret_cast<float*>(smem + 2 * R);
    float* s_ey = s_ex + R;
    float* s_lr = s_ey + R;
    const int tile = blockIdx.x / a.bpt, part = blockIdx.x % a.bpt;
    const int orgx = (tile % a.ntx) * a.tx - 2, orgy = (tile / a.ntx) * a.ty - 2;
    const int64_t nc = (int64_t)a.nx * a.ny;
    const int e = adj_exponent(a);
    const double scale = ldexp(1.0, e);
    if (blockIdx.x == 0 && threadIdx.x == 0) a.bd->e = e;
    for (int c = threadIdx.x; c < R; c += kThreads) {
        const int gc = dev::wrapfar(orgy + c / W, a.ny) * a.nx + dev::wrapfar(orgx + c % W, a.nx);
        s_lex[c] = 0ull;
        s_ley[c] = 0ull;
        s_ex[c] = a.Ex[gc];
        s_ey[c] = a.Ey[gc];
        s_lr[c] = a.xb ? a.lrho[gc] : 0.f;
    }
    __syncthreads();
    const int beg = a.off[tile], end = a.off[tile + 1];
    const int chunk = (end - beg + a.bpt - 1) / a.bpt;
    const int r0 = beg + part * chunk, r1 = min(end, r0 + chunk);
    float smax = 0.f;
    for (int p = r0 + (int)threadIdx.x; p < r1; p += kThreads) {
        float lx = a.lx[p], ly = a.ly[p], lux = a.lux[p], luy = a.luy[p];
        if (a.xb) {  // D_x^T lrho^{n+1} at x^{n+1}
            const Cic cb = cic_of(a.xb[p], a.yb[p]);
            float v[4], gx, gy;
            corners(a.lrho, s_lr, W, H, orgx, orgy, a.nx, a.ny, cb, v);
            grad4(cb, v, gx, gy);
            lx = fmaf(a.coef, gx, lx);
            ly = fmaf(a.coef, gy, ly);
        }
        lux = fmaf(a.cx, lx, lux);
        luy = fmaf(a.cy, ly, luy);
        // transpose of the Boris velocity update (and, after step 0, of the
        // half step back): le = h (I + R^T) lu, lu = R^T lu
        float rx = a.rc * lux - a.rs * luy, ry = a.rs * lux + a.rc * luy;
        float le0 = a.h * (lux + rx), le1 = a.h * (luy + ry);
        lux = rx;
        luy = ry;
        if (a.h0 != 0.f) {
            rx = a.rc0 * lux - a.rs0 * luy;
            ry = a.rs0 * lux + a.rc0 * luy;
            le0 = le0 + a.h0 * (lux + rx);  // unfused: without B0 this is exactly (qm dt/2) lu
            le1 = le1 + a.h0 * (luy + ry);
            lux = rx;
            luy = ry;
        }
        const Cic c = cic_of(a.xn[p], a.yn[p]);
        float ve[4], vf[4], dxx, dxy, dyx, dyy;
        corners(a.Ex, s_ex, W, H, orgx, orgy, a.nx, a.ny, c, ve);
        corners(a.Ey, s_ey, W, H, orgx, orgy, a.nx, a.ny, c, vf);
        grad4(c, ve, dxx, dxy);
        grad4(c, vf, dyx, dyy);
        lx += le0 * dxx + le1 * dyx;
        ly += le0 * dxy + le1 * dyy;
        a.lx[p] = lx;
        a.ly[p] = ly;
        a.lux[p] = lux;
        a.luy[p] = luy;
        smax = fmaxf(smax, max4abs(lx, ly, lux, luy));
        // gather^T: lE += le W, quantised: rint(le * W * 2^e) (exact FP64
        // product of the FP32 factors, one rounding)
        const float gx = 1.f - c.fx, gy = 1.f - c.fy;
        const float w4[4] = {gx * gy, c.fx * gy, gx * c.fy, c.fx * c.fy};
        const double q0 = dev::clamp_fix_f64((double)le0 * scale);
        const double q1 = dev::clamp_fix_f64((double)le1 * scale);
        const int li = local(c.i0, orgx, a.nx), lj = local(c.j0, orgy, a.ny);
        if (dev::in_tile(li, W - 2) && dev::in_tile(lj, H - 2)) {
            const int b = lj * W + li;
            const int idx[4] = {b, b + 1, b + W, b + W + 1};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                atomicAdd(&s_lex[idx[k]], dev::fix_mul(q0, (double)w4[k]));
                atomicAdd(&s_ley[idx[k]], dev::fix_mul(q1, (double)w4[k]));
            }
        } else {
            int i0, i1, j0, j1;
            dev::wrap_pair(c.i0, a.nx, i0, i1);
            dev::wrap_pair(c.j0, a.ny, j0, j1);
            const int idx[4] = {j0 * a.nx + i0, j0 * a.nx + i1, j1 * a.nx + i0, j1 * a.nx + i1};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                atomicAdd(&a.lE[idx[k]], dev::fix_mul(q0, (double)w4[k]));
                atomicAdd(&a.lE[nc + idx[k]], dev::fix_mul(q1, (double)w4[k]));
            }
        }
    }
    wave_max_to(smax, a.bd->smax[a.slot_in ^ 1]);
    __syncthreads();
    for (int c = threadIdx.x; c < R; c += kThreads) {
        const int gc = dev::wrapfar(orgy + c / W, a.ny) * a.nx + dev::wrapfar(orgx + c % W, a.nx);
        if (s_lex[c]) atomicAdd(&a.lE[gc], s_lex[c]);
        if (s_ley[c]) atomicAdd(&a.lE[nc + gc], s_ley[c]);
    }
}

// lx += coef D_x^T lrho at x (no tile: used at segment boundaries)
// and the bound of its outputs into *smax
__global__ void __launch_bounds__(kThreads)
adj_drho_kernel(const float* __restrict__ x, const float* __restrict__ y,
                const float* __restrict__ lrho, float* __restrict__ lx, float* __restrict__ ly,
                int n, int nx, int ny, float coef, unsigned int* __restrict__ smax) {
    const int p0 = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = p0 < n;
    const int p = live ? p0 : 0;  // lanes past the end read particle 0 and store nothing
    const Cic c = cic_of(x[p], y[p]);
    int i0, i1, j0, j1;
    dev::wrap_pair(c.i0, nx, i0, i1);
    dev::wrap_pair(c.j0, ny, j0, j1);
    const float v[4] = {lrho[j0 * nx + i0], lrho[j0 * nx + i1], lrho[j1 * nx + i0],
                        lrho[j1 * nx + i1]};
    float gx, gy;
    grad4(c, v, gx, gy);
    const float ox = fmaf(coef, gx, lx[p]), oy = fmaf(coef, gy, ly[p]);
    if (live) {
        lx[p] = ox;
        ly[p] = oy;
    }
    wave_max_to(live ? fmaxf(fabsf(ox), fabsf(oy)) : 0.f, smax);
}

// out[perm[i]] = in[i] for four arrays (back to the previous order)
__global__ void scatter4_kernel(const int32_t* __restrict__ perm, int n, const float* a0,
                                const float* a1, const float* a2, const float* a3, float* b0,
                                float* b1, float* b2, float* b3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = perm[i];
    b0[j] = a0[i];
    b1[j] = a1[i];
    b2[j] = a2[i];
    b3[j] = a3[i];
}

// p[0..nw) = 0 (32-bit words)
__global__ void zero_words_kernel(uint32_t* __restrict__ p, int64_t nw) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nw) p[i] = 0u;
}

__global__ void copy_words_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                  int64_t nw) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nw) dst[i] = src[i];
}

// dst[q][i] = src[q][i] for nq <= 4 arrays of n floats, one launch
struct Copy4 {
    const float* src[4];
    float* dst[4];
};

__global__ void copy4_kernel(Copy4 c, int nq, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int q = 0; q < nq; ++q) c.dst[q][i] = c.src[q][i];
}

// per-block partial sums of the parameter gradients: modes m (lx * sin((m+1)
// k1 xbase)/dx) and per species (lux, luy); the adjoint state in load order
struct GradArgs {
    const float *lx, *lux, *luy, *xbase;  // load order
    const int64_t* sp_end;  // cumulative particle count per species (load order)
    int n, nmodes, nspecies;
    double k1, inv_dx;
};

__global__ void __launch_bounds__(kThreads) grad_kernel(GradArgs a, double* __restrict__ part) {
    // part layout per block: [nmodes + 2 * nspecies]
    const int q = a.nmodes + 2 * a.nspecies;
    extern __shared__ double sh[];  // q * kThreads
    for (int k = 0; k < q; ++k) sh[k * kThreads + threadIdx.x] = 0.0;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < a.n; i += gridDim.x * kThreads) {
        const int slot = i;
        const double xb = a.xbase[slot];
        const double lx = a.lx[i];
        for (int m = 0; m < a.nmodes; ++m)
            sh[m * kThreads + threadIdx.x] += lx * sin((m + 1) * a.k1 * xb) * a.inv_dx;
        int s = 0;
        while (s + 1 < a.nspecies && slot >= a.sp_end[s]) ++s;
        sh[(a.nmodes + 2 * s) * kThreads + threadIdx.x] += a.lux[i];
        sh[(a.nmodes + 2 * s + 1) * kThreads + threadIdx.x] += a.luy[i];
    }
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < q; ++k) sh[k * kThreads + threadIdx.x] += sh[k * kThreads + threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < q; ++k) part[blockIdx.x * q + k] = sh[k * kThreads];
}


// ---- standalone adjoint primitives (aw_adj_deposit_t / gather_t / push_reverse) ----

// lx += c_p d/dx sum_k W_k lrho_k (the transpose of the CIC charge deposit
// with respect to the particle position; c_p = qm w_p / (dx dy))
__global__ void __launch_bounds__(kThreads)
deposit_t_kernel(const float* __restrict__ x, const float* __restrict__ y,
                 const float* __restrict__ w, float coef_u, float coef_w,
                 const float* __restrict__ lrho, float* __restrict__ lx, float* __restrict__ ly,
                 int n, int nx, int ny) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const Cic c = cic_of(x[p], y[p]);
    int i0, i1, j0, j1;
    dev::wrap_pair(c.i0, nx, i0, i1);
    dev::wrap_pair(c.j0, ny, j0, j1);
    const float v[4] = {lrho[j0 * nx + i0], lrho[j0 * nx + i1], lrho[j1 * nx + i0],
                        lrho[j1 * nx + i1]};
    float gx, gy;
    grad4(c, v, gx, gy);
    const float coef = w ? coef_w * w[p] : coef_u;
    lx[p] = fmaf(coef, gx, lx[p]);
    ly[p] = fmaf(coef, gy, ly[p]);
}

// max |le| over the particles (FP32 bit pattern, non-negative)
__global__ void __launch_bounds__(kThreads)
le_bound_kernel(const float* __restrict__ a, const float* __restrict__ b, int n,
                unsigned int* __restrict__ out) {
    float m = 0.f;
    for (int p = blockIdx.x * kThreads + threadIdx.x; p < n; p += gridDim.x * kThreads)
        m = fmaxf(m, fmaxf(fabsf(a[p]), fabsf(b[p])));
    wave_max_to(m, out);
}

// lE += sum_p W(x_p) le_p in int64 fixed point: scale 2^e with
// n * max|le| * 2^e < 2^62 (order-free, deterministic)
__device__ __forceinline__ int gather_t_exponent(const unsigned int* bound, int log2n) {
    const double B = (double)slots_max(bound) * (1.0 + 1e-6);
    if (!(B > 0.0) || !(B < 1e300)) return 0;
    int k;
    (void)frexp(B, &k);
    // each contribution below 2^50 (the quantiser's range), their sum below 2^62
    const int e = (62 - log2n < 50 ? 62 - log2n : 50) - k;
    return e > 1000 ? 1000 : (e < -1000 ? -1000 : e);
}

__global__ void __launch_bounds__(kThreads)
gather_t_kernel(const float* __restrict__ x, const float* __restrict__ y,
                const float* __restrict__ lex, const float* __restrict__ ley, int n, int nx,
                int ny, const unsigned int* __restrict__ bound, int log2n,
                unsigned long long* __restrict__ acc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double scale = ldexp(1.0, gather_t_exponent(bound, log2n));
    const Cic c = cic_of(x[p], y[p]);
    const float gx = 1.f - c.fx, gy = 1.f - c.fy;
    const float w4[4] = {gx * gy, c.fx * gy, gx * c.fy, c.fx * c.fy};
    int i0, i1, j0, j1;
    dev::wrap_pair(c.i0, nx, i0, i1);
    dev::wrap_pair(c.j0, ny, j0, j1);
    const int idx[4] = {j0 * nx + i0, j0 * nx + i1, j1 * nx + i0, j1 * nx + i1};
    const int64_t nc = (int64_t)nx * ny;
    const double q0 = dev::clamp_fix_f64((double)lex[p] * scale);
    const double q1 = dev::clamp_fix_f64((double)ley[p] * scale);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        atomicAdd(&acc[idx[k]], dev::fix_mul(q0, (double)w4[k]));
        atomicAdd(&acc[nc + idx[k]], dev::fix_mul(q1, (double)w4[k]));
    }
}

__global__ void gather_t_convert_kernel(unsigned long long* __restrict__ acc, int64_t nc,
                                        const unsigned int* __restrict__ bound, int log2n,
                                        float* __restrict__ lEx, float* __restrict__ lEy) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double inv = ldexp(1.0, -gather_t_exponent(bound, log2n));
    lEx[c] += (float)((double)(long long)acc[c] * inv);
    lEy[c] += (float)((double)(long long)acc[nc + c] * inv);
    acc[c] = 0ull;
    acc[nc + c] = 0ull;
}

struct PushT {
    const float *x, *y, *Ex, *Ey;
    float *lx, *ly, *lux, *luy, *lex, *ley;
    int n, nx, ny;
    float cx, cy, h, rc, rs;
};

// The transpose of one push at positions x^n: lu += c lx (the move);
// le = h (I + R^T) lu, lu = R^T lu (the Boris velocity update);
// lx += (de/dx)^T le (the gather's dependence on the position).
__global__ void __launch_bounds__(kThreads) push_t_kernel(PushT a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    float lux = fmaf(a.cx, a.lx[p], a.lux[p]), luy = fmaf(a.cy, a.ly[p], a.luy[p]);
    const float rx = a.rc * lux - a.rs * luy, ry = a.rs * lux + a.rc * luy;
    const float le0 = a.h * (lux + rx), le1 = a.h * (luy + ry);
    const Cic c = cic_of(a.x[p], a.y[p]);
    int i0, i1, j0, j1;
    dev::wrap_pair(c.i0, a.nx, i0, i1);
    dev::wrap_pair(c.j0, a.ny, j0, j1);
    const int idx[4] = {j0 * a.nx + i0, j0 * a.nx + i1, j1 * a.nx + i0, j1 * a.nx + i1};
    float ve[4], vf[4];
    for (int k = 0; k < 4; ++k) {
        ve[k] = a.Ex[idx[k]];
        vf[k] = a.Ey[idx[k]];
    }
    float dxx, dxy, dyx, dyy;
    grad4(c, ve, dxx, dxy);
    grad4(c, vf, dyx, dyy);
    a.lx[p] += le0 * dxx + le1 * dyx;
    a.ly[p] += le0 * dxy + le1 * dyy;
    a.lux[p] = rx;
    a.luy[p] = ry;
    a.lex[p] = le0;
    a.ley[p] = le1;
}
}  // namespace

}  // namespace aw

using namespace aw;

struct aw_adjoint {
    aw_grid g{};
    aw_run_params rp{};
    std::vector<aw_species> species;
    int n_modes = 0, K = 10;
    int64_t n = 0, nsteps = 0;
    double w_uniform = 0.0, loss = 0.0;
    std::unique_ptr<aw_ctx> ctx;
    hipStream_t stream = nullptr;
    int bpt = 1;
    // particle state (current order), its copy buffer, load-order bases
    DevBuf<float> P[2][5];
    int cur = 0;
    DevBuf<float> xbase;           // undisplaced x, load order
    DevBuf<float> init4;           // x, y, ux, uy at t = 0, load order
    DevBuf<int32_t> off;
    DevBuf<unsigned long long> strays;
    DevBuf<float> rho, E;          // rho^n, [Ex|Ey]^n
    // checkpoints: (x, y, ux, uy), tile offsets and the sort permutation
    int nck = 0;
    DevBuf<float> ck;              // nck * 4 * n
    DevBuf<int32_t> ck_off, ck_perm;
    // segment recompute: x^n, y^n and E^n for K+1 steps
    DevBuf<float> segx, segE;
    // adjoint state and grid work
    DevBuf<float> lx, ly, lux, luy, tmp4;
    DevBuf<unsigned long long> lE;  // int64 adjoint field accumulators
    DevBuf<float> lE0;              // the loss seed eps0 dx dy E^N
    DevBuf<double> part;
    DevBuf<AdjBound> bd;            // the sweep's order-free bounds
    DevBuf<float> lEf, lrho;
    DevBuf<int64_t> sp_end;
    bool forward_done = false;
    bool has_b0 = false;  // B0 along z (the run is unmagnetised otherwise)

    ~aw_adjoint() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void zero_words(void* p, size_t bytes) {
        const int64_t nw = (int64_t)(bytes / 4);
        if (!nw) return;
        zero_words_kernel<<<blocks_for(nw), kThreads, 0, stream>>>(static_cast<uint32_t*>(p), nw);
        AW_LAUNCH_CHECK();
    }

    void copy_words(const void* src, void* dst, size_t bytes) {
        const int64_t nw = (int64_t)(bytes / 4);
        if (!nw) return;
        copy_words_kernel<<<blocks_for(nw), kThreads, 0, stream>>>(
            static_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst), nw);
        AW_LAUNCH_CHECK();
    }

    aw_particles parts(int which) const {
        aw_particles p{};
        p.x = P[which][0].get();
        p.y = P[which][1].get();
        p.ux = P[which][2].get();
        p.uy = P[which][3].get();
        p.uz = P[which][4].get();
        p.w = nullptr;
        p.n = n;
        p.w_uniform = w_uniform;
        return p;
    }

    // stable tile sort of the current state into the other buffer; perm_out
    // (n) receives the permutation new[i] = old[perm[i]]
    void sort(int32_t* perm_out, int32_t* off_out) {
        const hipStream_t s = stream;
        const aw_particles src = parts(cur), dst = parts(cur ^ 1);
        if (!tile_sort_onepass(*ctx, src, dst, off_out, perm_out, s)) {
            ctx->sort.keyA.reserve((size_t)n);
            compute_tile_keys(*ctx, src, ctx->sort.keyA.get(), false, s);
            stable_sort_keys(*ctx, n, (uint32_t)ctx->ntiles, perm_out, off_out, s);
            permute_particles(src, dst, perm_out, s);
        }
        cur ^= 1;
    }

    void deposit_rho() {  // rho of the current state into the (zeroed) accumulator
        const hipStream_t s = stream;
        zero_words(ctx->acc.get(), ctx->acc.size() * sizeof(unsigned long long));
        deposit_rho_fixed(*ctx, parts(cur), rp.qm, AW_DEPOSIT_BINNED, off.get(), s);
    }

    // rho^n (in ctx->acc) -> rho, E2 = [Ex | Ey] (default: E); acc cleared
    void solve_current(float* E2 = nullptr) {
        ctx->rin.reserve((size_t)ncell(*ctx));
        fix_to_float_add(ctx->acc.get(), rho.get(), ncell(*ctx), ctx->fix_log2, true, stream, true,
                         strays.get(), ctx->rin.get(), 1, nullptr, ctx->rho_base);
        solve_es_contig(*ctx, ctx->rin.get(), E2 ? E2 : E.get(), rp.eps0, stream);
    }

    // up to four arrays of n floats, one launch
    void copy_arrays(std::initializer_list<std::pair<const float*, float*>> pairs) {
        AW_REQUIRE(pairs.size() <= 4, "copy_arrays takes at most four arrays");
        Copy4 c{};
        int q = 0;
        for (const auto& [src, dst] : pairs) {
            c.src[q] = src;
            c.dst[q] = dst;
            ++q;
        }
        copy4_kernel<<<blocks_for(n), kThreads, 0, stream>>>(c, q, (int)n);
        AW_LAUNCH_CHECK();
    }

    void save_checkpoint(int k) {
        const hipStream_t s = stream;
        const size_t b = (size_t)n * sizeof(float);
        float* c = ck.get() + (size_t)k * 4 * n;
        const aw_particles p = parts(cur);
        AW_HIP(hipMemcpyAsync(c, p.x, b, hipMemcpyDeviceToDevice, s));
        AW_HIP(hipMemcpyAsync(c + n, p.y, b, hipMemcpyDeviceToDevice, s));
        AW_HIP(hipMemcpyAsync(c + 2 * n, p.ux, b, hipMemcpyDeviceToDevice, s));
        AW_HIP(hipMemcpyAsync(c + 3 * n, p.uy, b, hipMemcpyDeviceToDevice, s));
        AW_HIP(hipMemcpyAsync(ck_off.get() + (size_t)k * (ctx->ntiles + 1), off.get(),
                              (size_t)(ctx->ntiles + 1) * 4, hipMemcpyDeviceToDevice, s));
    }

    void load_checkpoint(int k) {  // kernels only: part of the captured sweep
        const float* c = ck.get() + (size_t)k * 4 * n;
        const aw_particles p = parts(cur);
        copy_arrays({{c, p.x}, {c + n, p.y}, {c + 2 * n, p.ux}, {c + 3 * n, p.uy}});
        copy_words(ck_off.get() + (size_t)k * (ctx->ntiles + 1), off.get(),
                   (size_t)(ctx->ntiles + 1) * 4);
    }

    // one forward step n: E^n from rho^n, fused push + deposit of rho^{n+1}
    void fwd_step(int64_t step) {
        solve_current();
        aw_run_params r = rp;
        push_deposit_es(*ctx, parts(cur), off.get(), E.get(), E.get() + ncell(*ctx), r, has_b0,
                        strays.get(), nullptr, bpt, stream);
        (void)step;
    }

    void forward(const double* modes, const double* ufl, int64_t steps) {
        const hipStream_t s = stream;
        AW_REQUIRE(steps > 0, "nsteps must be positive");
        nsteps = steps;
        cur = 0;  // the load below rewrites both buffers: every run starts alike
        nck = (int)((steps + K - 1) / K);
        // load: the displaced state, then the undisplaced x (same draws)
        std::vector<aw_species> sp = species;
        if (ufl)
            for (size_t k = 0; k < sp.size(); ++k)
                for (int d = 0; d < 3; ++d) sp[k].ufl[d] = ufl[3 * k + d];
        std::vector<double> wts(sp.size());
        const aw_particles base = parts(cur ^ 1);
        load_species(*ctx, sp.data(), (int)sp.size(), rp, nullptr, 0, base, wts.data(), s);
        AW_HIP(hipMemcpyAsync(xbase.get(), base.x, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        load_species(*ctx, sp.data(), (int)sp.size(), rp, modes, n_modes, parts(cur), wts.data(),
                     s);
        {
            const aw_particles p = parts(cur);
            const float* src[4] = {p.x, p.y, p.ux, p.uy};
            for (int q = 0; q < 4; ++q)
                AW_HIP(hipMemcpyAsync(init4.get() + (size_t)q * n, src[q], (size_t)n * 4,
                                      hipMemcpyDeviceToDevice, s));
        }
        ck.reserve((size_t)nck * 4 * n);
        ck_off.reserve((size_t)nck * (ctx->ntiles + 1));
        ck_perm.reserve((size_t)nck * n);
        // t = 0: sort, rho^0, E^0, half step back
        sort(ck_perm.get(), off.get());
        deposit_rho();
        solve_current();
        push_es(*ctx, parts(cur), E.get(), E.get() + ncell(*ctx), rp, has_b0, true, s);
        // rho^0 again for step 0 (solve_current consumed the accumulator)
        deposit_rho();
        for (int64_t st = 0; st < steps; ++st) {
            if (st % K == 0) {
                const int k = (int)(st / K);
                if (k > 0) {  // re-sort at every checkpoint; rho^k is order-free
                    sort(ck_perm.get() + (size_t)k * n, off.get());
                }
                save_checkpoint(k);
            }
            fwd_step(st);
        }
        // loss from E^N
        solve_current();
        const int64_t nc = ncell(*ctx);
        const int nb = std::min(blocks_for(2 * nc), 256);
        part.reserve((size_t)std::max(nb, 4096));
        lE0.reserve((size_t)2 * nc);
        loss_kernel<<<nb, kThreads, 0, s>>>(E.get(), nc, rp.eps0 * g.dx * g.dy, lE0.get(),
                                            part.get());
        AW_LAUNCH_CHECK();
        std::vector<double> h(nb);
        AW_HIP(hipMemcpyAsync(h.data(), part.get(), nb * sizeof(double), hipMemcpyDeviceToHost, s));
        AW_HIP(hipStreamSynchronize(s));
        double sum = 0.0;
        for (double v : h) sum += v;
        loss = 0.5 * rp.eps0 * g.dx * g.dy * sum;
        forward_done = true;
    }

    // lrho = S^T lE: the int64 accumulators of the step just swept (cleared,
    // with the state-bound slot that step read), or the loss seed (slot < 0)
    void adj_solve(int free_slot) {
        const hipStream_t s = stream;
        const int64_t nc = ncell(*ctx), nk = nkmodes(*ctx);
        if (free_slot >= 0) {
            lE_convert_kernel<<<blocks_for(2 * nc), kThreads, 0, s>>>(lE.get(), lEf.get(), 2 * nc,
                                                                     bd.get(), free_slot);
            AW_LAUNCH_CHECK();
        } else {
            copy_words(lE0.get(), lEf.get(), (size_t)2 * nc * sizeof(float));
        }
        ctx->ktmp.reserve((size_t)6 * nk);
        float2* ak = reinterpret_cast<float2*>(ctx->ktmp.get());
        float2* rk = ak + 2 * nk;
        fft_exec(*ctx, ctx->r2c2, lEf.get(), ak, s);  // [lEx | lEy] -> [ak_x | ak_y]
        KG k{g.nx, g.ny, g.nx / 2 + 1, 6.28318530717958647692 / g.Lx,
             6.28318530717958647692 / g.Ly, rp.eps0};
        es_adj_kspace_kernel<<<blocks_for(nk), kThreads, 0, s>>>(ak, rk, k, bd.get()->rmax);
        AW_LAUNCH_CHECK();
        fft_exec(*ctx, ctx->c2r, rk, lrho.get(), s);
    }

    void backward(double* grad_modes, double* grad_ufl) {
        AW_REQUIRE(forward_done, "aw_adj_backward before aw_adj_forward");
        sweep();
        gradients(grad_modes, grad_ufl);
    }

    // seed, then every segment recomputed and swept in reverse; the adjoint
    // state ends in load order
    void sweep() {
        const hipStream_t s = stream;
        const int64_t nc = ncell(*ctx);
        const float coef = (float)(rp.qm * w_uniform / (g.dx * g.dy));
        const float cx = (float)(rp.dt / g.dx), cy = (float)(rp.dt / g.dy);
        // the Boris steps of the forward run: the step and the half step back
        const BorisConsts bs = boris_make(rp.qm, rp.dt, rp.B0, has_b0);
        const BorisConsts bh = boris_make(rp.qm, -0.5 * rp.dt, rp.B0, has_b0);
        auto rot_c = [](const BorisConsts& b) {
            return (float)(1.0 - (double)b.t[2] * (double)b.s[2]);
        };
        // seed: lrho^N = S^T (eps0 dx dy E^N), lx = lu = 0, bounds zero (the seed
        // is kept, so the backward sweep can be repeated)
        zero_words(bd.get(), sizeof(AdjBound));
        adj_solve(-1);
        const float rfac = (float)(2.0 * (double)nkmodes(*ctx) / (double)nc);
        int log2n = 0;
        while (log2n < 62 && ((int64_t)1 << log2n) < n) ++log2n;
        for (auto* b : {&lx, &ly, &lux, &luy}) zero_words(b->get(), (size_t)n * 4);
        segx.reserve((size_t)(K + 1) * 2 * n);
        segE.reserve((size_t)(K + 1) * 2 * nc);
        const size_t b = (size_t)n * sizeof(float);
        bool pending_drho = true;  // lrho of step "next" still to be applied at x^next
        for (int k = nck - 1; k >= 0; --k) {
            const int64_t s0 = (int64_t)k * K, s1 = std::min<int64_t>(nsteps, s0 + K);
            // recompute the segment from its checkpoint: x^n, E^n for n = s0..s1
            load_checkpoint(k);
            deposit_rho();
            for (int64_t st = s0; st <= s1; ++st) {
                const int64_t j = st - s0;
                const aw_particles p = parts(cur);
                float* xj = segx.get() + (size_t)j * 2 * n;
                if (st == s1) {  // the segment's last positions (no push follows)
                    copy_arrays({{p.x, xj}, {p.y, xj + n}});
                    break;
                }
                // E^n straight into the segment's store, and pushed from there;
                // the push saves x^n, y^n into the store as it moves them
                float* Ej = segE.get() + (size_t)j * 2 * nc;
                solve_current(Ej);
                push_deposit_es_save(*ctx, p, off.get(), Ej, Ej + nc, rp, has_b0, strays.get(),
                                     nullptr, bpt, xj, s);
            }
            // the last position of this segment is the first of the next one, in
            // this segment's order: apply the pending D_x^T at x^{s1}
            if (pending_drho) {
                const float* xs = segx.get() + (size_t)(s1 - s0) * 2 * n;
                adj_drho_kernel<<<blocks_for(n), kThreads, 0, s>>>(
                    xs, xs + n, lrho.get(), lx.get(), ly.get(), (int)n, g.nx, g.ny, coef,
                    bd.get()->smax[s1 & 1]);
                AW_LAUNCH_CHECK();
                pending_drho = false;
            }
            for (int64_t st = s1 - 1; st >= s0; --st) {
                const int64_t j = st - s0;
                AdjArgs aa{};
                aa.xn = segx.get() + (size_t)j * 2 * n;
                aa.yn = aa.xn + n;
                if (pending_drho) {
                    aa.xb = segx.get() + (size_t)(j + 1) * 2 * n;
                    aa.yb = aa.xb + n;
                }
                aa.lx = lx.get();
                aa.ly = ly.get();
                aa.lux = lux.get();
                aa.luy = luy.get();
                aa.Ex = segE.get() + (size_t)j * 2 * nc;
                aa.Ey = aa.Ex + nc;
                aa.lrho = lrho.get();
                aa.lE = lE.get();
                aa.bd = bd.get();
                aa.slot_in = (int)((st + 1) & 1);
                aa.rfac = rfac;
                aa.log2n = log2n;
                aa.off = off.get();
                aa.ntx = ctx->ntx;
                aa.tx = ctx->tx;
                aa.ty = ctx->ty;
                aa.W = ctx->tx + 5;
                aa.H = ctx->ty + 5;
                aa.bpt = bpt;
                aa.nx = g.nx;
                aa.ny = g.ny;
                aa.n = (int)n;
                aa.cx = cx;
                aa.cy = cy;
                aa.h = bs.qmh;
                aa.rc = rot_c(bs);
                aa.rs = bs.s[2];
                aa.h0 = st == 0 ? bh.qmh : 0.f;
                aa.rc0 = rot_c(bh);
                aa.rs0 = bh.s[2];
                aa.coef = coef;
                const size_t lds = (size_t)aa.W * aa.H * (16 + 12);
                adj_step_kernel<<<ctx->ntiles * bpt, kThreads, lds, s>>>(aa);
                AW_LAUNCH_CHECK();
                adj_solve(aa.slot_in);  // lrho^st
                pending_drho = true;   // applied at x^st by the next kernel
            }
            // D_x^T at x^{s0} in this order, then back to the previous segment's order
            const float* x0 = segx.get();
            adj_drho_kernel<<<blocks_for(n), kThreads, 0, s>>>(x0, x0 + n, lrho.get(), lx.get(),
                                                              ly.get(), (int)n, g.nx, g.ny, coef,
                                                              bd.get()->smax[s0 & 1]);
            AW_LAUNCH_CHECK();
            pending_drho = false;
            // undo the sort that produced this segment's order
            tmp4.reserve((size_t)4 * n);
            float* t = tmp4.get();
            scatter4_kernel<<<blocks_for(n), kThreads, 0, s>>>(
                ck_perm.get() + (size_t)k * n, (int)n, lx.get(), ly.get(), lux.get(), luy.get(),
                t, t + n, t + 2 * n, t + 3 * n);
            AW_LAUNCH_CHECK();
            copy_arrays({{t, lx.get()}, {t + n, ly.get()}, {t + 2 * n, lux.get()},
                         {t + 3 * n, luy.get()}});
        }
    }

    // parameter gradients from the adjoint state (load order)
    void gradients(double* grad_modes, double* grad_ufl) {
        const hipStream_t s = stream;
        GradArgs ga{};
        ga.lx = lx.get();
        ga.lux = lux.get();
        ga.luy = luy.get();
        ga.xbase = xbase.get();
        ga.sp_end = sp_end.get();
        ga.n = (int)n;
        ga.nmodes = n_modes;
        ga.nspecies = (int)species.size();
        ga.k1 = 6.28318530717958647692 / g.nx;
        ga.inv_dx = 1.0 / g.dx;
        const int q = n_modes + 2 * (int)species.size();
        AW_REQUIRE((size_t)q * kThreads * sizeof(double) <= ctx->lds_per_block,
                   "too many modes and species for the gradient reduction");
        const int nb = std::min(blocks_for(n), 512);
        part.reserve((size_t)nb * q);
        grad_kernel<<<nb, kThreads, (size_t)q * kThreads * sizeof(double), s>>>(ga, part.get());
        AW_LAUNCH_CHECK();
        std::vector<double> h((size_t)nb * q);
        AW_HIP(hipMemcpyAsync(h.data(), part.get(), h.size() * sizeof(double),
                              hipMemcpyDeviceToHost, s));
        AW_HIP(hipStreamSynchronize(s));
        std::vector<double> tot(q, 0.0);
        for (int blk = 0; blk < nb; ++blk)
            for (int i = 0; i < q; ++i) tot[i] += h[(size_t)blk * q + i];
        if (grad_modes)
            for (int m = 0; m < n_modes; ++m) grad_modes[m] = tot[m];
        if (grad_ufl)
            for (size_t sidx = 0; sidx < species.size(); ++sidx) {
                grad_ufl[3 * sidx] = tot[n_modes + 2 * sidx];
                grad_ufl[3 * sidx + 1] = tot[n_modes + 2 * sidx + 1];
                grad_ufl[3 * sidx + 2] = 0.0;  // uz does not enter (E in the plane, B0 along z)
            }
    }
};

extern "C" {

aw_status aw_adj_deposit_t(aw_ctx* ctx, const aw_particles* p, const float* lrho, double qm,
                           float* lx, float* ly, aw_stream s) {
    return guard("aw_adj_deposit_t", [&] {
        AW_REQUIRE(ctx && lrho && lx && ly, "NULL argument");
        check_particles(p);
        if (p->n == 0) return;
        const double area = ctx->g.dx * ctx->g.dy;
        deposit_t_kernel<<<blocks_for(p->n), kThreads, 0, S(s)>>>(
            p->x, p->y, p->w, (float)(qm * p->w_uniform / area), (float)(qm / area), lrho, lx,
            ly, (int)p->n, ctx->g.nx, ctx->g.ny);
        AW_LAUNCH_CHECK();
    });
}

aw_status aw_adj_gather_t(aw_ctx* ctx, const aw_particles* p, const float* lex, const float* ley,
                          float* lEx, float* lEy, aw_stream s) {
    return guard("aw_adj_gather_t", [&] {
        AW_REQUIRE(ctx && lEx && lEy, "NULL argument");
        check_particles(p);
        if (p->n == 0) return;
        AW_REQUIRE(lex && ley, "lex/ley is NULL");
        const int64_t nc = ncell(*ctx);
        const hipStream_t st = S(s);
        ctx->jacc.reserve((size_t)(3 * nc));
        ctx->jbound.reserve(kBoundWords);
        // cleared by kernels, not hipMemsetAsync, so the call can be graph-captured
        const int64_t nw = 4 * nc;  // the 2 nc int64 accumulators as 32-bit words
        zero_words_kernel<<<(int)((nw + kThreads - 1) / kThreads), kThreads, 0, st>>>(
            reinterpret_cast<uint32_t*>(ctx->jacc.get()), nw);
        AW_LAUNCH_CHECK();
        zero_words_kernel<<<(kBoundWords + kThreads - 1) / kThreads, kThreads, 0, st>>>(
            ctx->jbound.get(), kBoundWords);
        AW_LAUNCH_CHECK();
        unsigned int* bound = ctx->jbound.get();
        int log2n = 0;
        while (log2n < 62 && ((int64_t)1 << log2n) < p->n) ++log2n;
        const int nb = std::min(blocks_for(p->n), 4 * std::max(1, ctx->num_cu));
        le_bound_kernel<<<nb, kThreads, 0, st>>>(lex, ley, (int)p->n, bound);
        AW_LAUNCH_CHECK();
        gather_t_kernel<<<blocks_for(p->n), kThreads, 0, st>>>(p->x, p->y, lex, ley, (int)p->n,
                                                             ctx->g.nx, ctx->g.ny, bound, log2n,
                                                             ctx->jacc.get());
        AW_LAUNCH_CHECK();
        gather_t_convert_kernel<<<blocks_for(nc), kThreads, 0, st>>>(ctx->jacc.get(), nc, bound,
                                                                   log2n, lEx, lEy);
        AW_LAUNCH_CHECK();
    });
}

aw_status aw_adj_push_reverse(aw_ctx* ctx, const aw_particles* p, const float* Ex,
                              const float* Ey, const aw_run_params* rp, int32_t has_b0,
                              int32_t half_back, float* lx, float* ly, float* lux, float* luy,
                              float* lex, float* ley, aw_stream s) {
    return guard("aw_adj_push_reverse", [&] {
        AW_REQUIRE(ctx && Ex && Ey && rp && lx && ly && lux && luy && lex && ley,
                   "NULL argument");
        check_particles(p);
        AW_REQUIRE(!has_b0 || (rp->B0[0] == 0.f && rp->B0[1] == 0.f),
                   "the adjoint takes B0 along z only (rp.B0[0] = rp.B0[1] = 0)");
        AW_REQUIRE(!rp->rel, "the adjoint push is non-relativistic (rp.rel = 0)");
        if (p->n == 0) return;
        const double dt_eff = half_back ? -0.5 * rp->dt : rp->dt;
        const BorisConsts b = boris_make(rp->qm, dt_eff, rp->B0, has_b0 != 0);
        PushT a{};
        a.x = p->x;
        a.y = p->y;
        a.Ex = Ex;
        a.Ey = Ey;
        a.lx = lx;
        a.ly = ly;
        a.lux = lux;
        a.luy = luy;
        a.lex = lex;
        a.ley = ley;
        a.n = (int)p->n;
        a.nx = ctx->g.nx;
        a.ny = ctx->g.ny;
        // the half step back does not move the particle
        a.cx = half_back ? 0.f : (float)(rp->dt / ctx->g.dx);
        a.cy = half_back ? 0.f : (float)(rp->dt / ctx->g.dy);
        a.h = b.qmh;
        a.rc = (float)(1.0 - (double)b.t[2] * (double)b.s[2]);
        a.rs = b.s[2];
        push_t_kernel<<<blocks_for(p->n), kThreads, 0, S(s)>>>(a);
        AW_LAUNCH_CHECK();
    });
}

aw_status aw_adj_create(const aw_grid* g, const aw_run_params* rp, const aw_species* sp,
                        int32_t nspecies, int32_t n_modes, int32_t checkpoint_every,
                        int32_t tile, aw_adjoint** out) {
    return guard("aw_adj_create", [&] {
        AW_REQUIRE(g && rp && sp && nspecies > 0 && out, "NULL argument");
        AW_REQUIRE(n_modes >= 0 && n_modes <= 32, "n_modes must be in [0, 32]");
        AW_REQUIRE(checkpoint_every > 0, "checkpoint_every must be positive");
        AW_REQUIRE(rp->dt > 0 && rp->eps0 > 0, "dt and eps0 must be positive");
        AW_REQUIRE(rp->B0[0] == 0.f && rp->B0[1] == 0.f,
                   "the adjoint takes B0 along z only (rp.B0[0] = rp.B0[1] = 0)");
        AW_REQUIRE(!rp->rel, "the adjoint push is non-relativistic (rp.rel = 0)");
        *out = nullptr;
        auto a = std::make_unique<aw_adjoint>();
        a->g = *g;
        a->rp = *rp;
        a->species.assign(sp, sp + nspecies);
        a->n_modes = n_modes;
        a->K = checkpoint_every;
        a->has_b0 = rp->B0[2] != 0.f;
        const int64_t nc = (int64_t)g->nx * g->ny;
        std::vector<int64_t> ends;
        double w0 = -1.0;
        for (const auto& s : a->species) {
            AW_REQUIRE(s.ppc > 0, "species ppc must be positive");
            a->n += (int64_t)s.ppc * nc;
            ends.push_back(a->n);
            const double w = s.density * g->dx * g->dy / s.ppc;
            AW_REQUIRE(w0 < 0 || w == w0, "the adjoint run carries one particle weight");
            w0 = w;
        }
        AW_REQUIRE(a->n < (int64_t)INT32_MAX, "particle count must fit int32");
        a->w_uniform = w0;
        const int t = tile > 0 ? tile : (a->n >= 64 * nc ? 8 : 16);
        a->ctx.reset(make_ctx(g, t, t));
        a->ctx->fix_log2 = fixed_scale_for(a->species, *rp, false);
        {
            double dens = 0.0;
            for (const auto& s : a->species) dens += s.density;
            a->ctx->rho_base = rho_baseline(rp->qm, dens, a->ctx->fix_log2);
        }
        a->bpt = tile_bpt(*a->ctx, a->n);
        AW_HIP(hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking));
        const size_t n = (size_t)a->n;
        for (auto& buf : a->P)
            for (auto& v : buf) v.resize(n);
        a->xbase.resize(n);
        a->init4.resize(4 * n);
        a->off.resize((size_t)a->ctx->ntiles + 1);
        a->strays.resize(1);
        a->rho.resize((size_t)nc);
        a->E.resize((size_t)2 * nc);
        for (auto* v : {&a->lx, &a->ly, &a->lux, &a->luy}) v->resize(n);
        a->lE.resize((size_t)2 * nc);
        AW_HIP(hipMemset(a->lE.get(), 0, (size_t)2 * nc * sizeof(unsigned long long)));
        a->bd.resize(1);
        a->lEf.resize((size_t)2 * nc);
        a->lrho.resize((size_t)nc);
        a->sp_end.resize(ends.size());
        AW_HIP(hipMemcpy(a->sp_end.get(), ends.data(), ends.size() * sizeof(int64_t),
                         hipMemcpyHostToDevice));
        *out = a.release();
    });
}

aw_status aw_adj_destroy(aw_adjoint* a) {
    return guard("aw_adj_destroy", [&] { delete a; });
}

aw_status aw_adj_forward(aw_adjoint* a, const double* mode_amp, const double* ufl,
                         int64_t nsteps, double* loss) {
    return guard("aw_adj_forward", [&] {
        AW_REQUIRE(a, "adjoint is NULL");
        AW_REQUIRE(a->n_modes == 0 || mode_amp, "mode_amp is NULL");
        a->forward(mode_amp, ufl, nsteps);
        if (loss) *loss = a->loss;
    });
}

aw_status aw_adj_backward(aw_adjoint* a, double* grad_modes, double* grad_ufl) {
    return guard("aw_adj_backward", [&] {
        AW_REQUIRE(a, "adjoint is NULL");
        a->backward(grad_modes, grad_ufl);
    });
}

aw_status aw_adj_state_grad(aw_adjoint* a, float* gx, float* gy, float* gux, float* guy) {
    return guard("aw_adj_state_grad", [&] {
        AW_REQUIRE(a, "adjoint is NULL");
        const hipStream_t s = a->stream;
        const size_t b = (size_t)a->n * sizeof(float);
        const float* src[4] = {a->lx.get(), a->ly.get(), a->lux.get(), a->luy.get()};
        float* dst[4] = {gx, gy, gux, guy};
        for (int q = 0; q < 4; ++q)
            if (dst[q]) AW_HIP(hipMemcpyAsync(dst[q], src[q], b, hipMemcpyDeviceToDevice, s));
        AW_HIP(hipStreamSynchronize(s));
    });
}

aw_status aw_adj_initial_state(aw_adjoint* a, aw_particles* out) {
    return guard("aw_adj_initial_state", [&] {
        AW_REQUIRE(a && out, "NULL argument");
        AW_REQUIRE(a->forward_done, "aw_adj_initial_state before aw_adj_forward");
        AW_REQUIRE(out->n == a->n, "out->n must equal the particle count");
        const hipStream_t s = a->stream;
        const size_t b = (size_t)a->n * sizeof(float);
        float* dst[4] = {out->x, out->y, out->ux, out->uy};
        for (int q = 0; q < 4; ++q)
            if (dst[q])
                AW_HIP(hipMemcpyAsync(dst[q], a->init4.get() + (size_t)q * a->n, b,
                                      hipMemcpyDeviceToDevice, s));
        AW_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
