// darwin.hip -- spectral Darwin particle kernels.
//
//   em_moments_kernel  rho, J, amu of the current state (init / API):
//                      deposit_rho_j_tiled_kernel + deposit_amu_kernel
//                      (include/pic/depositor.hpp:160-270)
//   dcu_kernel         qm^2 w/area (E + v x B) with E, B gathered from an LDS
//                      field tile: deposit_dcu_kernel (:275-324)
//   em_push_kernel     E, B gathered from an LDS tile, full Boris, move, wrap,
//                      then rho, J, amu of the NEW state deposited for the next
//                      step: em_push_tiled_kernel (pusher.hpp:211-258) +
//                      particle_migrate_kernel + the next step's ρ/J/amu
//                      deposits, for 40 B of particle traffic.
//
// All eight moments (rho, Jx, Jy, Jz, amu0..amu3) and the three dcu
// components accumulate fixed-point int64 quanta: each contribution is the
// FP32 product v*W rounded to an integer, v carrying the quantity's 2^S scale.
#include "aw_internal.hpp"

namespace aw {

namespace {

constexpr int kThreads = 256;

inline int blocks_for(int64_t n) { return (int)((n + kThreads - 1) / kThreads); }

__device__ __forceinline__ int local_index(int i, int org, int n) {
    int l = i - org;
    if (l < 0) l += n;
    else if (l >= n) l -= n;
    return l;
}

struct EmArgs {
    float *x, *y, *ux, *uy, *uz;
    const float* w;
    const int32_t* off;
    const int32_t* idx;          // bin permutation (unsorted particles) or nullptr
    int ntx, tx, ty, apron, W, H, bpt;
    int nx, ny, ncell;
    // fields (E = total, B), each [ny][nx]
    const float *Ex, *Ey, *Ez, *Bx, *By, *Bz;
    unsigned long long* acc;     // nq grids of ncell int64
    const float* cq;             // device: rho, J, amu coefficients 2^S * qm/area
                                 // (* w_uniform when uniform)
    const float* c_dcu;          // device: 2^S_dcu * qm^2/area (* w_uniform)
    float qmh, cx, cy;
    double dxdt, dydt;           // dx/dt, dy/dt (Esirkepov current coefficients)
    unsigned long long* strays;
    double* ke_part;
    // uniform batch (EmBatch): run blockIdx.y offsets its pointers by these
    long long bp, be, bb, ba, bk;
    int bo;
};

// Moves every per-run pointer of a to run blockIdx.y of a uniform batch.
__device__ __forceinline__ void select_run(EmArgs& a) {
    const long long r = blockIdx.y;
    if (r == 0) return;
    a.x += r * a.bp;
    a.y += r * a.bp;
    a.ux += r * a.bp;
    a.uy += r * a.bp;
    a.uz += r * a.bp;
    if (a.w) a.w += r * a.bp;
    a.off += r * a.bo;
    if (a.Ex) {
        a.Ex += r * a.be;
        a.Ey += r * a.be;
        a.Ez += r * a.be;
        a.Bx += r * a.bb;
        a.By += r * a.bb;
        a.Bz += r * a.bb;
    }
    if (a.acc) a.acc += r * a.ba;
    if (a.c_dcu) a.c_dcu += r;
    if (a.ke_part) a.ke_part += r * a.bk;
}

// Range guard of the fixed-point moments: the static scales put the largest
// contribution allowed by |u| <= 2 max(1, c) at 2^48.  A moment beyond 2^50
// (|u| four times that bound for J, twice for amu) marks the thread's stray
// count with kRangeBit; the count reaches the host as stray_count +
// kRangeFlag, and the simulation reports the overflow instead of letting
// saturated (clamped) contributions pile up into a wrapped int64 sum.
constexpr float kRangeLimit = 1125899906842624.0f;  // 2^50
constexpr unsigned int kRangeBit = 0x80000000u;

__device__ __forceinline__ unsigned long long stray_word(unsigned int nstray) {
    return (unsigned long long)(nstray & ~kRangeBit) + ((nstray & kRangeBit) ? kRangeFlag : 0ull);
}

// the 8 moment values of one particle (already scaled); cq = a.cq
__device__ __forceinline__ void moments8(const float cq[3], float mw, float ux, float uy, float uz,
                                         float v[8]) {
    const float br = cq[0] * mw, bj = cq[1] * mw, ba = cq[2] * mw;
    v[0] = br;
    v[1] = bj * ux;
    v[2] = bj * uy;
    v[3] = bj * uz;
    v[4] = ba * (0.5f * (ux * ux - uy * uy));
    v[5] = ba * (ux * uy);
    v[6] = ba * (uz * ux);
    v[7] = ba * (uz * uy);
}

// Deposits v[q] for the q < NQ selected by MASK (grid q of the tile / global
// accumulators): 4 CIC contributions rint(v[q] * W_k) per moment, with the
// FP64 weights of cic_d (exact products) and one fma per contribution.
template <int NQ, unsigned MASK = (1u << NQ) - 1u>
__device__ __forceinline__ void deposit_q(unsigned long long* s_acc, int R, int W, int H,
                                          unsigned long long* g_acc, int ncell, int nx, int ny,
                                          int orgx, int orgy, const dev::CicD& d,
                                          const float* v_in, unsigned int& nstray) {
    double v[NQ];
    bool over = false;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        over = over || fabsf(v_in[q]) > kRangeLimit;
        v[q] = dev::clamp_fix_f64((double)v_in[q]);
    }
    if (over) nstray |= kRangeBit;
    const int mi = local_index(d.i0, orgx, nx), mj = local_index(d.j0, orgy, ny);
    if (dev::in_tile(mi, W - 2) && dev::in_tile(mj, H - 2)) {
        const int b = mj * W + mi;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (!((MASK >> q) & 1u)) continue;
            unsigned long long* s = s_acc + q * R;
            atomicAdd(&s[b], dev::fix_mul(v[q], d.w0));
            atomicAdd(&s[b + 1], dev::fix_mul(v[q], d.w1));
            atomicAdd(&s[b + W], dev::fix_mul(v[q], d.w2));
            atomicAdd(&s[b + W + 1], dev::fix_mul(v[q], d.w3));
        }
    } else {
        int i0, i1, j0, j1;
        dev::wrap_pair(d.i0, nx, i0, i1);
        dev::wrap_pair(d.j0, ny, j0, j1);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (!((MASK >> q) & 1u)) continue;
            unsigned long long* g = g_acc + (size_t)q * ncell;
            atomicAdd(&g[j0 * nx + i0], dev::fix_mul(v[q], d.w0));
            atomicAdd(&g[j0 * nx + i1], dev::fix_mul(v[q], d.w1));
            atomicAdd(&g[j1 * nx + i0], dev::fix_mul(v[q], d.w2));
            atomicAdd(&g[j1 * nx + i1], dev::fix_mul(v[q], d.w3));
        }
        ++nstray;
    }
}

template <int NQ>
__device__ __forceinline__ void flush_q(const unsigned long long* s_acc, int R, int W,
                                        unsigned long long* g_acc, int ncell, int nx, int ny,
                                        int orgx, int orgy) {
    for (int c = threadIdx.x; c < R; c += kThreads) {
        const int gc = dev::wrapfar(orgy + c / W, ny) * nx + dev::wrapfar(orgx + c % W, nx);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const unsigned long long v = s_acc[q * R + c];
            if (v) atomicAdd(&g_acc[(size_t)q * ncell + gc], v);
        }
    }
}

struct Frame {
    int W, H, R, orgx, orgy, r0, r1;
};

__device__ __forceinline__ Frame frame(const EmArgs& a) {
    Frame f;
    f.W = a.W;
    f.H = a.H;
    f.R = a.W * a.H;
    const int tile = blockIdx.x / a.bpt, part = blockIdx.x % a.bpt;
    f.orgx = (tile % a.ntx) * a.tx - a.apron;
    f.orgy = (tile / a.ntx) * a.ty - a.apron;
    dev::split_range(a.off[tile], a.off[tile + 1], part, a.bpt, f.r0, f.r1);
    return f;
}

// field tile: node c -> (Ex, Ey, Ez, Bx) float4 + (By, Bz) float2
__device__ __forceinline__ void load_fields(const EmArgs& a, const Frame& f, float4* fa,
                                            float2* fb) {
    for (int c = threadIdx.x; c < f.R; c += kThreads) {
        const int gc =
            dev::wrapfar(f.orgy + c / f.W, a.ny) * a.nx + dev::wrapfar(f.orgx + c % f.W, a.nx);
        fa[c] = make_float4(a.Ex[gc], a.Ey[gc], a.Ez[gc], a.Bx[gc]);
        fb[c] = make_float2(a.By[gc], a.Bz[gc]);
    }
}

// gather (E, B) at a particle: LDS tile when inside, global otherwise
__device__ __forceinline__ bool gather6(const EmArgs& a, const Frame& f, const float4* fa,
                                        const float2* fb, const dev::CicF& s, float e[3],
                                        float b[3]) {
    const int li = local_index(s.i0, f.orgx, a.nx), lj = local_index(s.j0, f.orgy, a.ny);
    if (dev::in_tile(li, f.W - 2) && dev::in_tile(lj, f.H - 2)) {
        const int n0 = lj * f.W + li, n1 = n0 + 1, n2 = n0 + f.W, n3 = n2 + 1;
        const float4 a0 = fa[n0], a1 = fa[n1], a2 = fa[n2], a3 = fa[n3];
        const float2 b0 = fb[n0], b1 = fb[n1], b2 = fb[n2], b3 = fb[n3];
        e[0] = dev::gather4(s, a0.x, a1.x, a2.x, a3.x);
        e[1] = dev::gather4(s, a0.y, a1.y, a2.y, a3.y);
        e[2] = dev::gather4(s, a0.z, a1.z, a2.z, a3.z);
        b[0] = dev::gather4(s, a0.w, a1.w, a2.w, a3.w);
        b[1] = dev::gather4(s, b0.x, b1.x, b2.x, b3.x);
        b[2] = dev::gather4(s, b0.y, b1.y, b2.y, b3.y);
        return true;
    }
    int i0, i1, j0, j1;
    dev::wrap_pair(s.i0, a.nx, i0, i1);
    dev::wrap_pair(s.j0, a.ny, j0, j1);
    const int c0 = j0 * a.nx + i0, c1 = j0 * a.nx + i1, c2 = j1 * a.nx + i0, c3 = j1 * a.nx + i1;
    const float* F[6] = {a.Ex, a.Ey, a.Ez, a.Bx, a.By, a.Bz};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const float v = dev::gather4(s, F[q][c0], F[q][c1], F[q][c2], F[q][c3]);
        if (q < 3) e[q] = v;
        else b[q - 3] = v;
    }
    return false;
}

// Esirkepov current of the move (x0, y0) -> (x1, y1) into the J grids
// (q = 1..3) of the moment tile, or into the global accumulators when the 4x4
// union stencil leaves the tile frame.  cj = the J coefficient (fixed-point
// scaled, times the particle weight): Jx = -cj dx/dt sum DSx(..),
// Jy = -cj dy/dt sum DSy(..), Jz = cj uz W_z.  Every contribution is bounded
// by cj |u|, the bound of the CIC current, so the J scale carries over.
__device__ __forceinline__ void esk_deposit(const EmArgs& a, const Frame& f,
                                            unsigned long long* s_acc, float x0, float y0,
                                            float x1, float y1, float uz, double cj) {
    dev::EskShape m;
    dev::esk_shape(x0, y0, x1, y1, a.nx, a.ny, m);
    const double cJx = -cj * a.dxdt, cJy = -cj * a.dydt, cJz = cj * (double)uz;
    const int li = local_index(m.ib - 1, f.orgx, a.nx), lj = local_index(m.jb - 1, f.orgy, a.ny);
    if (dev::in_tile(li, f.W - 4) && dev::in_tile(lj, f.H - 4)) {
        unsigned long long* s = s_acc + f.R + lj * f.W + li;
        dev::esk_emit(m, cJx, cJy, cJz, [&](int q, int k, int b, unsigned long long v) {
            atomicAdd(&s[q * f.R + b * f.W + k], v);
        });
    } else {
        unsigned long long* g = a.acc + (size_t)a.ncell;
        dev::esk_emit(m, cJx, cJy, cJz, [&](int q, int k, int b, unsigned long long v) {
            atomicAdd(&g[(size_t)q * a.ncell + (size_t)dev::wrapfar(m.jb - 1 + b, a.ny) * a.nx +
                         dev::wrapfar(m.ib - 1 + k, a.nx)],
                      v);
        });
    }
}

template <bool PER_W>
__global__ void __launch_bounds__(kThreads) em_moments_kernel(EmArgs a) {
    select_run(a);
    extern __shared__ unsigned long long smem[];
    const Frame f = frame(a);
    for (int c = threadIdx.x; c < 8 * f.R; c += kThreads) smem[c] = 0ull;
    __syncthreads();
    unsigned int nstray = 0;
    const float cq[3] = {a.cq[0], a.cq[1], a.cq[2]};
    for (int r = f.r0 + (int)threadIdx.x; r < f.r1; r += kThreads) {
        const int k = a.idx ? a.idx[r] : r;
        const float mw = PER_W ? a.w[k] : 1.0f;
        float v[8];
        moments8(cq, mw, a.ux[k], a.uy[k], a.uz[k], v);
        deposit_q<8>(smem, f.R, f.W, f.H, a.acc, a.ncell, a.nx, a.ny, f.orgx, f.orgy,
                     dev::cic_d(a.x[k], a.y[k]), v, nstray);
    }
    __syncthreads();
    flush_q<8>(smem, f.R, f.W, a.acc, a.ncell, a.nx, a.ny, f.orgx, f.orgy);
    if (nstray && a.strays) atomicAdd(a.strays, stray_word(nstray));
}

template <bool PER_W>
__global__ void __launch_bounds__(kThreads) dcu_kernel(EmArgs a) {
    select_run(a);
    extern __shared__ unsigned long long smem[];
    const Frame f = frame(a);
    unsigned long long* s_acc = smem;
    float4* fa = reinterpret_cast<float4*>(smem + 3 * f.R);
    float2* fb = reinterpret_cast<float2*>(fa + f.R);
    for (int c = threadIdx.x; c < 3 * f.R; c += kThreads) s_acc[c] = 0ull;
    load_fields(a, f, fa, fb);
    __syncthreads();
    const float cd = *a.c_dcu;
    unsigned int nstray = 0;
    // two particles per thread per iteration, both loaded first
    auto one = [&](float x, float y, float vx, float vy, float vz, float mw) {
        const dev::CicF s = dev::cic_f(x, y);
        float e[3], b[3];
        if (!gather6(a, f, fa, fb, s, e, b)) ++nstray;
        const float base = PER_W ? cd * mw : cd;
        float v[3];
        v[0] = base * (e[0] + (vy * b[2] - vz * b[1]));
        v[1] = base * (e[1] + (vz * b[0] - vx * b[2]));
        v[2] = base * (e[2] + (vx * b[1] - vy * b[0]));
        deposit_q<3>(s_acc, f.R, f.W, f.H, a.acc, a.ncell, a.nx, a.ny, f.orgx, f.orgy,
                     dev::cic_d(x, y), v, nstray);
    };
    int r = f.r0 + (int)threadIdx.x;
    for (; r + kThreads < f.r1; r += 2 * kThreads) {
        const int k0 = a.idx ? a.idx[r] : r;
        const int k1 = a.idx ? a.idx[r + kThreads] : r + kThreads;
        const float x0 = a.x[k0], y0 = a.y[k0], ux0 = a.ux[k0], uy0 = a.uy[k0], uz0 = a.uz[k0];
        const float x1 = a.x[k1], y1 = a.y[k1], ux1 = a.ux[k1], uy1 = a.uy[k1], uz1 = a.uz[k1];
        const float w0 = PER_W ? a.w[k0] : 1.0f, w1 = PER_W ? a.w[k1] : 1.0f;
        one(x0, y0, ux0, uy0, uz0, w0);
        one(x1, y1, ux1, uy1, uz1, w1);
    }
    if (r < f.r1) {
        const int k = a.idx ? a.idx[r] : r;
        one(a.x[k], a.y[k], a.ux[k], a.uy[k], a.uz[k], PER_W ? a.w[k] : 1.0f);
    }
    __syncthreads();
    flush_q<3>(s_acc, f.R, f.W, a.acc, a.ncell, a.nx, a.ny, f.orgx, f.orgy);
    if (nstray && a.strays) atomicAdd(a.strays, stray_word(nstray));
}

// One particle of em_push_kernel, its SoA values already loaded.
template <bool PER_W, bool KE, bool DEP, bool ESK>
__device__ __forceinline__ void em_one(const EmArgs& a, const Frame& f, const float4* fa,
                                       const float2* fb, unsigned long long* s_acc,
                                       const float cq[3], int k, float x, float y, float ux,
                                       float uy, float uz, float mw, unsigned int& nstray,
                                       double& ke) {
    const dev::CicF s = dev::cic_f(x, y);
    float e[3], b[3];
    if (!gather6(a, f, fa, fb, s, e, b)) ++nstray;
    dev::boris_full(a.qmh, e[0], e[1], e[2], b[0], b[1], b[2], ux, uy, uz);
    const float xn = dev::wrap_coord(fmaf(ux, a.cx, x), (float)a.nx);
    const float yn = dev::wrap_coord(fmaf(uy, a.cy, y), (float)a.ny);
    a.x[k] = xn;
    a.y[k] = yn;
    a.ux[k] = ux;
    a.uy[k] = uy;
    a.uz[k] = uz;
    if constexpr (KE)
        ke += (PER_W ? (double)mw : 1.0) * ((double)ux * ux + (double)uy * uy + (double)uz * uz);
    if constexpr (DEP) {
        float v[8];
        moments8(cq, mw, ux, uy, uz, v);
        if constexpr (ESK) {
            deposit_q<8, 0xF1u>(s_acc, f.R, f.W, f.H, a.acc, a.ncell, a.nx, a.ny, f.orgx,
                                f.orgy, dev::cic_d(xn, yn), v, nstray);
            esk_deposit(a, f, s_acc, x, y, xn, yn, uz, (double)cq[1] * (double)mw);
        } else {
            deposit_q<8>(s_acc, f.R, f.W, f.H, a.acc, a.ncell, a.nx, a.ny, f.orgx, f.orgy,
                         dev::cic_d(xn, yn), v, nstray);
        }
    }
}

// DEP = false: push, move and wrap only (Pusher::boris_em + migrate); the
// LDS moment tile is then not allocated.  ESK: the current of the next state
// is the Esirkepov current of this step's move (staggered Jx, Jy), not the
// CIC current at x^{n+1}; rho and amu stay CIC at x^{n+1}.  Two particles
// per thread per iteration, both loaded before either is pushed.
template <bool PER_W, bool KE, bool DEP, bool ESK = false>
__global__ void __launch_bounds__(kThreads) em_push_kernel(EmArgs a) {
    select_run(a);
    extern __shared__ unsigned long long smem[];
    const Frame f = frame(a);
    constexpr int kNq = DEP ? 8 : 0;
    unsigned long long* s_acc = smem;
    float4* fa = reinterpret_cast<float4*>(smem + kNq * f.R);
    float2* fb = reinterpret_cast<float2*>(fa + f.R);
    for (int c = threadIdx.x; c < kNq * f.R; c += kThreads) s_acc[c] = 0ull;
    load_fields(a, f, fa, fb);
    __syncthreads();
    unsigned int nstray = 0;
    double ke = 0.0;
    float cq[3] = {0.f, 0.f, 0.f};
    if constexpr (DEP) {
        cq[0] = a.cq[0];
        cq[1] = a.cq[1];
        cq[2] = a.cq[2];
    }
    int r = f.r0 + (int)threadIdx.x;
    for (; r + kThreads < f.r1; r += 2 * kThreads) {
        const int k0 = a.idx ? a.idx[r] : r;
        const int k1 = a.idx ? a.idx[r + kThreads] : r + kThreads;
        const float x0 = a.x[k0], y0 = a.y[k0], ux0 = a.ux[k0], uy0 = a.uy[k0], uz0 = a.uz[k0];
        const float x1 = a.x[k1], y1 = a.y[k1], ux1 = a.ux[k1], uy1 = a.uy[k1], uz1 = a.uz[k1];
        const float w0 = PER_W ? a.w[k0] : 1.0f, w1 = PER_W ? a.w[k1] : 1.0f;
        em_one<PER_W, KE, DEP, ESK>(a, f, fa, fb, s_acc, cq, k0, x0, y0, ux0, uy0, uz0, w0,
                                    nstray, ke);
        em_one<PER_W, KE, DEP, ESK>(a, f, fa, fb, s_acc, cq, k1, x1, y1, ux1, uy1, uz1, w1,
                                    nstray, ke);
    }
    if (r < f.r1) {
        const int k = a.idx ? a.idx[r] : r;
        em_one<PER_W, KE, DEP, ESK>(a, f, fa, fb, s_acc, cq, k, a.x[k], a.y[k], a.ux[k],
                                    a.uy[k], a.uz[k], PER_W ? a.w[k] : 1.0f, nstray, ke);
    }
    __shared__ double s_ke[kThreads / 64];
    if constexpr (KE) {
        for (int o = 32; o > 0; o >>= 1) ke += __shfl_down(ke, o, 64);
        if ((threadIdx.x & 63) == 0) s_ke[threadIdx.x >> 6] = ke;
    }
    __syncthreads();
    if constexpr (DEP) flush_q<8>(s_acc, f.R, f.W, a.acc, a.ncell, a.nx, a.ny, f.orgx, f.orgy);
    if (nstray && a.strays) atomicAdd(a.strays, stray_word(nstray));
    if constexpr (KE) {
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < kThreads / 64; ++w) t += s_ke[w];
            a.ke_part[blockIdx.x] = t;
        }
    }
}

// nq int64 grids -> float grids (assign), acc cleared; inv scale per grid
// from device memory (the dcu scale changes every step).  out2 receives a
// second copy: rocFFT may overwrite the input of an out-of-place R2C, so the
// transform reads the copy and the source grids stay intact.
// Batched (gridDim.y runs): run r uses acc + r*sa, out + r*so, out2 + r*so2,
// inv + r*si.
// base0 (per run, may be NULL): the integer baseline taken off grid 0's copy
// in out2 -- the run's mean charge density, which the E_L solve drops (k = 0),
// so the FP32 transform rounds relative to the fluctuation.
__global__ void convert_multi_kernel(unsigned long long* __restrict__ acc, float* __restrict__ out,
                                     float* __restrict__ out2, int ncell, int nq,
                                     const double* __restrict__ inv, long long sa, long long so,
                                     long long so2, int si, const long long* __restrict__ base0) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const long long r = blockIdx.y;
    acc += r * sa;
    out += r * so;
    if (out2) out2 += r * so2;
    inv += r * si;
    for (int q = 0; q < nq; ++q) {
        const size_t i = (size_t)q * ncell + c;
        const long long a = (long long)acc[i];
        const float v = (float)((double)a * inv[q]);
        out[i] = v;
        if (out2) {
            const long long b = (q == 0 && base0) ? base0[r] : 0ll;
            out2[i] = b ? (float)((double)(a - b) * inv[q]) : v;
        }
        acc[i] = 0ull;
    }
}

__device__ __forceinline__ double pump_ramp(double t, double trmp, double toff) {
    if (t >= toff) return 0.0;
    if (t < trmp) return t / trmp;
    if (t > toff - trmp) return (toff - t) / trmp;
    return 1.0;
}

struct PumpArgs {
    int on;
    double ex, ey, ez, k0, w0, trmp, toff, t, dx;
};

// E = (E_L + s*E_T) + pump(t) (combine_etotal_kernel, solver_darwin.hpp:134-144;
// add_pump_kernel, pump.hpp:34-47); ET[0] may be null (s = 0).
__global__ void form_e_kernel(FormE f, int nx, int ncell, PumpArgs p) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    if (const long long r = blockIdx.y) {  // run r of a batch
        for (int q = 0; q < 3; ++q) {
            f.E[q] += r * f.sE;
            if (f.ET[q]) f.ET[q] += r * f.sT;
        }
        f.EL[0] += r * f.sL;
        f.EL[1] += r * f.sL;
    }
    float ex = f.EL[0][c], ey = f.EL[1][c], ez = 0.0f;
    if (f.ET[0]) {
        ex = ex + 1.0f * f.ET[0][c];
        ey = ey + 1.0f * f.ET[1][c];
        ez = 1.0f * f.ET[2][c];
    }
    if (p.on && p.t < p.toff) {
        const double x = (c % nx) * p.dx;
        const double th = p.k0 * x - p.w0 * p.t;
        const double r = pump_ramp(p.t, p.trmp, p.toff);
        const double ct = cos(th), st = sin(th);
        ex += (float)(p.ex * ct * r);
        ey += (float)(-p.ey * st * r);
        ez += (float)(p.ez * ct * r);
    }
    f.E[0][c] = ex;
    f.E[1][c] = ey;
    f.E[2][c] = ez;
}

struct Field6 {
    const float* f[6];  // Ex, Ey, Ez, Bx, By, Bz
};

// block maxima of |E|, |B| components -> part[blockIdx]
__global__ void __launch_bounds__(kThreads)
field_max_kernel(Field6 F, int ncell, float* __restrict__ part, long long sE, long long sB) {
    __shared__ float sh[2][kThreads];
    if (const long long r = blockIdx.y) {  // run r of a batch
        for (int q = 0; q < 3; ++q) {
            F.f[q] += r * sE;
            F.f[3 + q] += r * sB;
        }
        part += r * 2 * gridDim.x;
    }
    float me = 0.f, mb = 0.f;
    for (int c = blockIdx.x * kThreads + threadIdx.x; c < ncell; c += gridDim.x * kThreads) {
        for (int q = 0; q < 3; ++q) {
            me = fmaxf(me, fabsf(F.f[q][c]));
            mb = fmaxf(mb, fabsf(F.f[3 + q][c]));
        }
    }
    sh[0][threadIdx.x] = me;
    sh[1][threadIdx.x] = mb;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[0][threadIdx.x] = fmaxf(sh[0][threadIdx.x], sh[0][threadIdx.x + o]);
            sh[1][threadIdx.x] = fmaxf(sh[1][threadIdx.x], sh[1][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sh[0][0];
        part[2 * blockIdx.x + 1] = sh[1][0];
    }
}

// dcu scale: per-contribution bound = |qm^2 w/area| (max|E| + umax max|B|);
// S = 48 - ceil(log2(bound)) maps the bound to 2^48.  Writes the float
// coefficient used by dcu_kernel and the inverse scale used by conversion.
// dev (optional, device): {coef_abs, umax} measured from the particles.
// Batched: one thread per run (threadIdx.x), run r reads part + 2*nb*r and
// writes c_dcu[r], inv_dcu + si*r.
__global__ void dcu_scale_kernel(const float* __restrict__ part, int nb, double coef_abs,
                                 double coef_signed, double umax, const double* __restrict__ dev,
                                 float* __restrict__ c_dcu, double* __restrict__ inv_dcu, int runs,
                                 int si) {
    const int r = threadIdx.x;
    if (r >= runs || blockIdx.x != 0) return;
    part += (size_t)2 * nb * r;
    c_dcu += r;
    inv_dcu += (size_t)si * r;
    if (dev) {
        coef_abs = dev[0];
        umax = dev[1];
    }
    float me = 0.f, mb = 0.f;
    for (int i = 0; i < nb; ++i) {
        me = fmaxf(me, part[2 * i]);
        mb = fmaxf(mb, part[2 * i + 1]);
    }
    double bound = coef_abs * ((double)me + umax * (double)mb);
    if (!(bound > 1e-300)) bound = 1e-300;
    int S = 48 - (int)ceil(log2(bound));
    if (S > 120) S = 120;
    if (S < -120) S = -120;
    *c_dcu = (float)ldexp(coef_signed, S);
    const double inv = ldexp(1.0, -S);
    inv_dcu[0] = inv;  // one entry per dcu component (convert_multi reads 3)
    inv_dcu[1] = inv;
    inv_dcu[2] = inv;
}

__device__ __forceinline__ int scale_exp(double bound) {
    if (!(bound > 1e-300)) bound = 1e-300;
    const int S = 48 - (int)ceil(log2(bound));
    return S > 120 ? 120 : (S < -120 ? -120 : S);
}

__global__ void zero_u32_kernel(unsigned int* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// max |w| (per-particle weights) and max |u| component, as positive-float bit
// patterns (unsigned order = value order)
__global__ void __launch_bounds__(kThreads)
moment_bound_kernel(const float* __restrict__ ux, const float* __restrict__ uy,
                    const float* __restrict__ uz, const float* __restrict__ w, int n,
                    unsigned int* bound) {
    float mw = 0.f, mu = 0.f;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        mu = fmaxf(mu, fmaxf(fabsf(ux[i]), fmaxf(fabsf(uy[i]), fabsf(uz[i]))));
        if (w) mw = fmaxf(mw, fabsf(w[i]));
    }
    for (int o = 32; o > 0; o >>= 1) {
        mw = fmaxf(mw, __shfl_xor(mw, o, 64));
        mu = fmaxf(mu, __shfl_xor(mu, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&bound[0], __float_as_uint(mw));
        atomicMax(&bound[1], __float_as_uint(mu));
    }
}

// Moment scales from the measured bounds (the same rule as the Simulation's
// static scales): cq = {rho, J, amu} coefficients, inv[0..7] the inverse
// scales of rho, Jx..Jz, amu0..3, dv = {dcu coef_abs, umax}.
__global__ void moment_scale_kernel(const unsigned int* __restrict__ bound, int per_w,
                                    double qm, double area, double w_uniform,
                                    float* __restrict__ cq, double* __restrict__ inv,
                                    double* __restrict__ dv) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double wmax = per_w ? (double)__uint_as_float(bound[0]) : fabs(w_uniform);
    const double umax = (double)__uint_as_float(bound[1]);
    const double cabs = fabs(qm) * wmax / area;
    const int S[3] = {scale_exp(cabs), scale_exp(cabs * umax), scale_exp(cabs * umax * umax)};
    const double mult = per_w ? 1.0 : w_uniform;
    for (int q = 0; q < 3; ++q) cq[q] = (float)ldexp(qm * mult / area, S[q]);
    inv[0] = ldexp(1.0, -S[0]);
    for (int q = 1; q < 4; ++q) inv[q] = ldexp(1.0, -S[1]);
    for (int q = 4; q < 8; ++q) inv[q] = ldexp(1.0, -S[2]);
    dv[0] = qm * qm * wmax / area;
    dv[1] = umax;
}

struct Out8 {
    float* o[8];
};

// out_q += acc_q * inv[q] for q < nq (separate output grids)
__global__ void convert_add_kernel(const unsigned long long* __restrict__ acc, Out8 out, int ncell,
                                   int nq, const double* __restrict__ inv) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    for (int q = 0; q < nq; ++q)  // a NULL output skips its quantity
        if (out.o[q]) out.o[q][c] += (float)((double)(long long)acc[(size_t)q * ncell + c] * inv[q]);
}

}  // namespace

// ---- host side ------------------------------------------------------------------------

int tile_bpt(const aw_ctx& c, int64_t n);

static EmArgs em_args(aw_ctx& c, const aw_particles& p, const int32_t* off, const int32_t* idx,
                      const float* const* F, unsigned long long* acc, const float* cq,
                      const float* c_dcu, double qm, double dt, int bpt,
                      unsigned long long* strays, double* ke_part) {
    EmArgs a{};
    a.x = p.x;
    a.y = p.y;
    a.ux = p.ux;
    a.uy = p.uy;
    a.uz = p.uz;
    a.w = p.w;
    a.off = off;
    a.idx = idx;
    a.ntx = c.ntx;
    a.tx = c.tx;
    a.ty = c.ty;
    a.apron = c.apron;
    a.W = c.tx + 2 * c.apron + 1;
    a.H = c.ty + 2 * c.apron + 1;
    a.bpt = bpt;
    a.nx = c.g.nx;
    a.ny = c.g.ny;
    a.ncell = (int)ncell(c);
    if (F) {
        a.Ex = F[0];
        a.Ey = F[1];
        a.Ez = F[2];
        a.Bx = F[3];
        a.By = F[4];
        a.Bz = F[5];
    }
    a.acc = acc;
    a.cq = cq;
    a.c_dcu = c_dcu;
    a.qmh = (float)(qm * 0.5 * dt);
    a.cx = (float)(dt / c.g.dx);
    a.cy = (float)(dt / c.g.dy);
    a.dxdt = dt != 0.0 ? c.g.dx / dt : 0.0;
    a.dydt = dt != 0.0 ? c.g.dy / dt : 0.0;
    a.strays = strays;
    a.ke_part = ke_part;
    return a;
}

static void set_batch(EmArgs& a, const EmBatch& b) {
    a.bp = b.pstride;
    a.be = b.estride;
    a.bb = b.bstride;
    a.ba = b.astride;
    a.bk = b.kstride;
    a.bo = b.ostride;
}

// Batched forms (b.runs > 1): p describes run 0 (p.n particles per run);
// runs 1.. lie at the strides of b.
void em_moments(aw_ctx& c, const aw_particles& p, const int32_t* off, const int32_t* idx,
                unsigned long long* acc8, const float* cq, int bpt, unsigned long long* strays,
                hipStream_t s, const EmBatch& b) {
    if (p.n == 0) return;
    EmArgs a = em_args(c, p, off, idx, nullptr, acc8, cq, nullptr, 0.0, 0.0, bpt, strays,
                       nullptr);
    set_batch(a, b);
    const size_t sh = (size_t)a.W * a.H * 8 * 8;
    AW_REQUIRE(sh <= c.lds_per_block, "tile too large for the LDS of the moment deposit");
    const dim3 grid(c.ntiles * bpt, b.runs);
    if (p.w) em_moments_kernel<true><<<grid, kThreads, sh, s>>>(a);
    else em_moments_kernel<false><<<grid, kThreads, sh, s>>>(a);
    AW_LAUNCH_CHECK();
}

void em_dcu(aw_ctx& c, const aw_particles& p, const int32_t* off, const int32_t* idx,
            const float* const* F, unsigned long long* acc3, const float* c_dcu, int bpt,
            unsigned long long* strays, hipStream_t s, const EmBatch& b) {
    if (p.n == 0) return;
    EmArgs a = em_args(c, p, off, idx, F, acc3, nullptr, c_dcu, 0.0, 0.0, bpt, strays, nullptr);
    set_batch(a, b);
    const size_t sh = (size_t)a.W * a.H * (3 * 8 + 16 + 8);
    AW_REQUIRE(sh <= c.lds_per_block, "tile too large for the LDS of the dcu deposit");
    const dim3 grid(c.ntiles * bpt, b.runs);
    if (p.w) dcu_kernel<true><<<grid, kThreads, sh, s>>>(a);
    else dcu_kernel<false><<<grid, kThreads, sh, s>>>(a);
    AW_LAUNCH_CHECK();
}

// acc8 == nullptr: push, move and wrap only.  esk: the J moments of acc8 are
// the Esirkepov current of the move (staggered Jx, Jy).
int em_push(aw_ctx& c, const aw_particles& p, const int32_t* off, const int32_t* idx,
            const float* const* F, unsigned long long* acc8, const float* cq, double qm,
            double dt, int bpt, unsigned long long* strays, double* ke_part, hipStream_t s,
            bool esk, const EmBatch& b) {
    if (p.n == 0) return 0;
    EmArgs a = em_args(c, p, off, idx, F, acc8, cq, nullptr, qm, dt, bpt, strays, ke_part);
    set_batch(a, b);
    const bool dep = acc8 != nullptr;
    const size_t sh = (size_t)a.W * a.H * ((dep ? 8 * 8 : 0) + 16 + 8);
    AW_REQUIRE(sh <= c.lds_per_block, "tile too large for the LDS of the EM push");
    const int nblk = c.ntiles * bpt;
    const bool pw = p.w != nullptr, ke = ke_part != nullptr;
    auto launch = [&](auto kernel) {
        kernel<<<dim3(nblk, b.runs), kThreads, sh, s>>>(a);
        AW_LAUNCH_CHECK();
    };
    if (dep && esk) {
        if (pw && ke) launch(em_push_kernel<true, true, true, true>);
        else if (pw) launch(em_push_kernel<true, false, true, true>);
        else if (ke) launch(em_push_kernel<false, true, true, true>);
        else launch(em_push_kernel<false, false, true, true>);
    } else if (dep) {
        if (pw && ke) launch(em_push_kernel<true, true, true>);
        else if (pw) launch(em_push_kernel<true, false, true>);
        else if (ke) launch(em_push_kernel<false, true, true>);
        else launch(em_push_kernel<false, false, true>);
    } else {
        if (pw && ke) launch(em_push_kernel<true, true, false>);
        else if (pw) launch(em_push_kernel<true, false, false>);
        else if (ke) launch(em_push_kernel<false, true, false>);
        else launch(em_push_kernel<false, false, false>);
    }
    return nblk;
}

// Batched (runs > 1): run r converts acc + r*sa into out + r*so (and
// out2 + r*so2) with the scales at inv + r*si.
void convert_multi(unsigned long long* acc, float* out, float* out2, int64_t ncell, int nq,
                   const double* inv, hipStream_t s, int runs, long long sa, long long so,
                   long long so2, int si, const long long* base0) {
    convert_multi_kernel<<<dim3(blocks_for(ncell), runs), kThreads, 0, s>>>(
        acc, out, out2, (int)ncell, nq, inv, sa, so, so2, si, base0);
    AW_LAUNCH_CHECK();
}

// Batched (runs > 1): f carries run 0's pointers and the per-run strides.
void form_e(const FormE& f, const aw_grid& g, const aw_run_params& rp, double t, hipStream_t s,
            int runs) {
    PumpArgs p{};
    p.on = rp.pump ? 1 : 0;
    p.ex = rp.pump_ex;
    p.ey = rp.pump_ey;
    p.ez = rp.pump_ez;
    p.k0 = rp.pump_k0;
    p.w0 = rp.pump_w0;
    p.trmp = rp.pump_trmp;
    p.toff = rp.pump_toff;
    p.t = t;
    p.dx = g.dx;
    const int64_t nc = (int64_t)g.nx * g.ny;
    form_e_kernel<<<dim3(blocks_for(nc), runs), kThreads, 0, s>>>(f, g.nx, (int)nc, p);
    AW_LAUNCH_CHECK();
}

// Batched (runs > 1): F holds run 0's fields, E components at stride sE and
// B at stride sB per run; part holds dcu_part_capacity(ncell) floats per run,
// c_dcu one float and inv_dcu si doubles per run.
void dcu_scale(const float* const* F, int64_t ncell, float* part, double coef_abs,
               double coef_signed, double umax, const double* dev, float* c_dcu, double* inv_dcu,
               hipStream_t s, int runs, long long sE, long long sB, int si) {
    AW_REQUIRE(runs >= 1 && runs <= 64, "dcu scale: 1..64 runs");
    const int nb = (int)std::min<int64_t>(256, std::max<int64_t>(1, (ncell + 4095) / 4096));
    Field6 f6{{F[0], F[1], F[2], F[3], F[4], F[5]}};
    field_max_kernel<<<dim3(nb, runs), kThreads, 0, s>>>(f6, (int)ncell, part, sE, sB);
    AW_LAUNCH_CHECK();
    dcu_scale_kernel<<<1, 64, 0, s>>>(part, nb, coef_abs, coef_signed, umax, dev, c_dcu, inv_dcu,
                                      runs, si);
    AW_LAUNCH_CHECK();
}

int dcu_part_capacity() { return 2 * 256; }

// ---- standalone operators (any particle order) ------------------------------------------
//
// aw_deposit_rho_j_amu, aw_deposit_dcu and aw_push_boris_em take particles in
// any order: they bin them by tile (stable counting sort into the context's
// bin permutation) and read them through it; the fixed-point scales come from
// the measured max |w| and max |u|.

static void bin_particles(aw_ctx& c, const aw_particles& p, hipStream_t s) {
    c.sort.bin_off.reserve((size_t)c.ntiles + 1);
    c.sort.bin_perm.reserve((size_t)std::max<int64_t>(1, p.n));
    bin_tiles(c, p, c.sort.bin_perm.get(), c.sort.bin_off.get(), s);
}

static void measure_scales(aw_ctx& c, const aw_particles& p, double qm, hipStream_t s) {
    EmScratch& e = c.em;
    e.bound.reserve(2);
    e.cq.reserve(3);
    e.inv.reserve(11);
    e.dv.reserve(2);
    zero_u32_kernel<<<1, 64, 0, s>>>(e.bound.get(), 2);
    AW_LAUNCH_CHECK();
    const int nb = (int)std::min<int64_t>(4ll * std::max(1, c.num_cu), (p.n + kThreads - 1) / kThreads);
    moment_bound_kernel<<<std::max(1, nb), kThreads, 0, s>>>(p.ux, p.uy, p.uz, p.w, (int)p.n,
                                                            e.bound.get());
    AW_LAUNCH_CHECK();
    moment_scale_kernel<<<1, 64, 0, s>>>(e.bound.get(), p.w ? 1 : 0, qm, c.g.dx * c.g.dy,
                                         p.w_uniform, e.cq.get(), e.inv.get(), e.dv.get());
    AW_LAUNCH_CHECK();
}

void deposit_rho_j_amu(aw_ctx& c, const aw_particles& p, double qm, float* const out[8],
                       hipStream_t s) {
    if (p.n == 0) return;
    AW_REQUIRE(p.n < (int64_t)INT32_MAX, "particle count must fit int32");
    const int64_t nc = ncell(c);
    EmScratch& e = c.em;
    e.acc.reserve((size_t)(8 * nc));
    AW_HIP(hipMemsetAsync(e.acc.get(), 0, (size_t)(8 * nc) * sizeof(unsigned long long), s));
    bin_particles(c, p, s);
    measure_scales(c, p, qm, s);
    em_moments(c, p, c.sort.bin_off.get(), c.sort.bin_perm.get(), e.acc.get(), e.cq.get(),
               tile_bpt(c, p.n), nullptr, s, EmBatch{});
    Out8 o{};
    for (int q = 0; q < 8; ++q) o.o[q] = out[q];
    convert_add_kernel<<<blocks_for(nc), kThreads, 0, s>>>(e.acc.get(), o, (int)nc, 8, e.inv.get());
    AW_LAUNCH_CHECK();
}

void deposit_dcu(aw_ctx& c, const aw_particles& p, double qm, const float* const* F,
                 float* const out[3], hipStream_t s) {
    if (p.n == 0) return;
    AW_REQUIRE(p.n < (int64_t)INT32_MAX, "particle count must fit int32");
    const int64_t nc = ncell(c);
    EmScratch& e = c.em;
    e.acc.reserve((size_t)(8 * nc));
    e.cdcu.reserve(1);
    e.fpart.reserve(512);
    AW_HIP(hipMemsetAsync(e.acc.get(), 0, (size_t)(3 * nc) * sizeof(unsigned long long), s));
    bin_particles(c, p, s);
    measure_scales(c, p, qm, s);
    const double mult = p.w ? 1.0 : p.w_uniform;
    dcu_scale(F, nc, e.fpart.get(), 0.0, qm * qm * mult / (c.g.dx * c.g.dy), 0.0, e.dv.get(),
              e.cdcu.get(), e.inv.get() + 8, s, 1, 0, 0, 0);
    em_dcu(c, p, c.sort.bin_off.get(), c.sort.bin_perm.get(), F, e.acc.get(), e.cdcu.get(),
           tile_bpt(c, p.n), nullptr, s, EmBatch{});
    Out8 o{};
    for (int q = 0; q < 3; ++q) o.o[q] = out[q];
    convert_add_kernel<<<blocks_for(nc), kThreads, 0, s>>>(e.acc.get(), o, (int)nc, 3,
                                                            e.inv.get() + 8);
    AW_LAUNCH_CHECK();
}

void push_boris_em(aw_ctx& c, const aw_particles& p, double qm, double dt, const float* const* F,
                   hipStream_t s) {
    if (p.n == 0) return;
    AW_REQUIRE(p.n < (int64_t)INT32_MAX, "particle count must fit int32");
    bin_particles(c, p, s);
    em_push(c, p, c.sort.bin_off.get(), c.sort.bin_perm.get(), F, nullptr, nullptr, qm, dt,
            tile_bpt(c, p.n), nullptr, nullptr, s, false, EmBatch{});
}

}  // namespace aw
