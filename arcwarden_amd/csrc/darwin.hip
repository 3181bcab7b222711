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
    int ntx, tx, ty, apron, W, H, bpt;
    int nx, ny, ncell;
    // fields (E = total, B), each [ny][nx]
    const float *Ex, *Ey, *Ez, *Bx, *By, *Bz;
    unsigned long long* acc;     // nq grids of ncell int64
    float c_rho, c_j, c_amu;     // 2^S * qm/area (* w_uniform when uniform)
    const float* c_dcu;          // device: 2^S_dcu * qm^2/area (* w_uniform)
    float qmh, cx, cy;
    unsigned long long* strays;
    double* ke_part;
};

// the 8 moment values of one particle (already scaled)
__device__ __forceinline__ void moments8(const EmArgs& a, float mw, float ux, float uy, float uz,
                                         float v[8]) {
    const float br = a.c_rho * mw, bj = a.c_j * mw, ba = a.c_amu * mw;
    v[0] = br;
    v[1] = bj * ux;
    v[2] = bj * uy;
    v[3] = bj * uz;
    v[4] = ba * (0.5f * (ux * ux - uy * uy));
    v[5] = ba * (ux * uy);
    v[6] = ba * (uz * ux);
    v[7] = ba * (uz * uy);
}

template <int NQ>
__device__ __forceinline__ void deposit_q(unsigned long long* s_acc, int R, int W, int H,
                                          unsigned long long* g_acc, int ncell, int nx, int ny,
                                          int orgx, int orgy, const dev::CicF& d,
                                          const float* v, unsigned int& nstray) {
    const int mi = local_index(d.i0, orgx, nx), mj = local_index(d.j0, orgy, ny);
    if (dev::in_tile(mi, W - 2) && dev::in_tile(mj, H - 2)) {
        const int b = mj * W + mi;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            unsigned long long* s = s_acc + q * R;
            atomicAdd(&s[b], dev::fixq32(v[q], d.w0));
            atomicAdd(&s[b + 1], dev::fixq32(v[q], d.w1));
            atomicAdd(&s[b + W], dev::fixq32(v[q], d.w2));
            atomicAdd(&s[b + W + 1], dev::fixq32(v[q], d.w3));
        }
    } else {
        int i0, i1, j0, j1;
        dev::wrap_pair(d.i0, nx, i0, i1);
        dev::wrap_pair(d.j0, ny, j0, j1);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            unsigned long long* g = g_acc + (size_t)q * ncell;
            atomicAdd(&g[j0 * nx + i0], dev::fixq32(v[q], d.w0));
            atomicAdd(&g[j0 * nx + i1], dev::fixq32(v[q], d.w1));
            atomicAdd(&g[j1 * nx + i0], dev::fixq32(v[q], d.w2));
            atomicAdd(&g[j1 * nx + i1], dev::fixq32(v[q], d.w3));
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
    const int beg = a.off[tile], end = a.off[tile + 1];
    const int chunk = (end - beg + a.bpt - 1) / a.bpt;
    f.r0 = beg + part * chunk;
    f.r1 = min(end, f.r0 + chunk);
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

template <bool PER_W>
__global__ void __launch_bounds__(kThreads) em_moments_kernel(EmArgs a) {
    extern __shared__ unsigned long long smem[];
    const Frame f = frame(a);
    for (int c = threadIdx.x; c < 8 * f.R; c += kThreads) smem[c] = 0ull;
    __syncthreads();
    unsigned int nstray = 0;
    for (int k = f.r0 + (int)threadIdx.x; k < f.r1; k += kThreads) {
        const float mw = PER_W ? a.w[k] : 1.0f;
        float v[8];
        moments8(a, mw, a.ux[k], a.uy[k], a.uz[k], v);
        deposit_q<8>(smem, f.R, f.W, f.H, a.acc, a.ncell, a.nx, a.ny, f.orgx, f.orgy,
                     dev::cic_f(a.x[k], a.y[k]), v, nstray);
    }
    __syncthreads();
    flush_q<8>(smem, f.R, f.W, a.acc, a.ncell, a.nx, a.ny, f.orgx, f.orgy);
    if (nstray && a.strays) atomicAdd(a.strays, (unsigned long long)nstray);
}

template <bool PER_W>
__global__ void __launch_bounds__(kThreads) dcu_kernel(EmArgs a) {
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
    for (int k = f.r0 + (int)threadIdx.x; k < f.r1; k += kThreads) {
        const dev::CicF s = dev::cic_f(a.x[k], a.y[k]);
        float e[3], b[3];
        if (!gather6(a, f, fa, fb, s, e, b)) ++nstray;
        const float vx = a.ux[k], vy = a.uy[k], vz = a.uz[k];
        const float base = PER_W ? cd * a.w[k] : cd;
        float v[3];
        v[0] = base * (e[0] + (vy * b[2] - vz * b[1]));
        v[1] = base * (e[1] + (vz * b[0] - vx * b[2]));
        v[2] = base * (e[2] + (vx * b[1] - vy * b[0]));
        deposit_q<3>(s_acc, f.R, f.W, f.H, a.acc, a.ncell, a.nx, a.ny, f.orgx, f.orgy, s, v,
                     nstray);
    }
    __syncthreads();
    flush_q<3>(s_acc, f.R, f.W, a.acc, a.ncell, a.nx, a.ny, f.orgx, f.orgy);
    if (nstray && a.strays) atomicAdd(a.strays, (unsigned long long)nstray);
}

template <bool PER_W, bool KE>
__global__ void __launch_bounds__(kThreads) em_push_kernel(EmArgs a) {
    extern __shared__ unsigned long long smem[];
    const Frame f = frame(a);
    unsigned long long* s_acc = smem;
    float4* fa = reinterpret_cast<float4*>(smem + 8 * f.R);
    float2* fb = reinterpret_cast<float2*>(fa + f.R);
    for (int c = threadIdx.x; c < 8 * f.R; c += kThreads) s_acc[c] = 0ull;
    load_fields(a, f, fa, fb);
    __syncthreads();
    unsigned int nstray = 0;
    double ke = 0.0;
    for (int k = f.r0 + (int)threadIdx.x; k < f.r1; k += kThreads) {
        const float x = a.x[k], y = a.y[k];
        float ux = a.ux[k], uy = a.uy[k], uz = a.uz[k];
        const float mw = PER_W ? a.w[k] : 1.0f;
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
            ke += (PER_W ? (double)mw : 1.0) *
                  ((double)ux * ux + (double)uy * uy + (double)uz * uz);
        float v[8];
        moments8(a, mw, ux, uy, uz, v);
        deposit_q<8>(s_acc, f.R, f.W, f.H, a.acc, a.ncell, a.nx, a.ny, f.orgx, f.orgy,
                     dev::cic_f(xn, yn), v, nstray);
    }
    __shared__ double s_ke[kThreads / 64];
    if constexpr (KE) {
        for (int o = 32; o > 0; o >>= 1) ke += __shfl_down(ke, o, 64);
        if ((threadIdx.x & 63) == 0) s_ke[threadIdx.x >> 6] = ke;
    }
    __syncthreads();
    flush_q<8>(s_acc, f.R, f.W, a.acc, a.ncell, a.nx, a.ny, f.orgx, f.orgy);
    if (nstray && a.strays) atomicAdd(a.strays, (unsigned long long)nstray);
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
__global__ void convert_multi_kernel(unsigned long long* __restrict__ acc, float* __restrict__ out,
                                     float* __restrict__ out2, int ncell, int nq,
                                     const double* __restrict__ inv) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    for (int q = 0; q < nq; ++q) {
        const size_t i = (size_t)q * ncell + c;
        const float v = (float)((double)(long long)acc[i] * inv[q]);
        out[i] = v;
        if (out2) out2[i] = v;
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
// add_pump_kernel, pump.hpp:34-47); ET may be null (s = 0).
__global__ void form_e_kernel(float* __restrict__ E3, const float* __restrict__ EL2,
                              const float* __restrict__ ET3, int nx, int ncell, PumpArgs p) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    float ex = EL2[c], ey = EL2[ncell + c], ez = 0.0f;
    if (ET3) {
        ex = ex + 1.0f * ET3[c];
        ey = ey + 1.0f * ET3[ncell + c];
        ez = 1.0f * ET3[2 * ncell + c];
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
    E3[c] = ex;
    E3[ncell + c] = ey;
    E3[2 * ncell + c] = ez;
}

// block maxima of |E|, |B| components -> part[blockIdx]
__global__ void __launch_bounds__(kThreads)
field_max_kernel(const float* __restrict__ E3, const float* __restrict__ B3, int ncell,
                 float* __restrict__ part) {
    __shared__ float sh[2][kThreads];
    float me = 0.f, mb = 0.f;
    for (int c = blockIdx.x * kThreads + threadIdx.x; c < ncell; c += gridDim.x * kThreads) {
        for (int q = 0; q < 3; ++q) {
            me = fmaxf(me, fabsf(E3[(size_t)q * ncell + c]));
            mb = fmaxf(mb, fabsf(B3[(size_t)q * ncell + c]));
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
__global__ void dcu_scale_kernel(const float* __restrict__ part, int nb, double coef_abs,
                                 double coef_signed, double umax, float* __restrict__ c_dcu,
                                 double* __restrict__ inv_dcu) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
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

}  // namespace

// ---- host side ------------------------------------------------------------------------

int tile_bpt(const aw_ctx& c, int64_t n);

struct EmLaunch {
    EmArgs a;
    int nblk;
    size_t lds_fields;
};

static EmArgs em_args(aw_ctx& c, const aw_particles& p, const int32_t* off, const float* E3,
                      const float* B3, unsigned long long* acc, const float cq[3],
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
    if (E3) {
        a.Ex = E3;
        a.Ey = E3 + a.ncell;
        a.Ez = E3 + 2 * a.ncell;
    }
    if (B3) {
        a.Bx = B3;
        a.By = B3 + a.ncell;
        a.Bz = B3 + 2 * a.ncell;
    }
    a.acc = acc;
    a.c_rho = cq[0];
    a.c_j = cq[1];
    a.c_amu = cq[2];
    a.c_dcu = c_dcu;
    a.qmh = (float)(qm * 0.5 * dt);
    a.cx = (float)(dt / c.g.dx);
    a.cy = (float)(dt / c.g.dy);
    a.strays = strays;
    a.ke_part = ke_part;
    return a;
}

void em_moments(aw_ctx& c, const aw_particles& p, const int32_t* off, unsigned long long* acc8,
                const float cq[3], int bpt, unsigned long long* strays, hipStream_t s) {
    if (p.n == 0) return;
    EmArgs a = em_args(c, p, off, nullptr, nullptr, acc8, cq, nullptr, 0.0, 0.0, bpt, strays,
                       nullptr);
    const size_t sh = (size_t)a.W * a.H * 8 * 8;
    const int nblk = c.ntiles * bpt;
    if (p.w) em_moments_kernel<true><<<nblk, kThreads, sh, s>>>(a);
    else em_moments_kernel<false><<<nblk, kThreads, sh, s>>>(a);
    AW_LAUNCH_CHECK();
}

void em_dcu(aw_ctx& c, const aw_particles& p, const int32_t* off, const float* E3,
            const float* B3, unsigned long long* acc3, const float* c_dcu, int bpt,
            unsigned long long* strays, hipStream_t s) {
    if (p.n == 0) return;
    const float zero[3] = {0.f, 0.f, 0.f};
    EmArgs a = em_args(c, p, off, E3, B3, acc3, zero, c_dcu, 0.0, 0.0, bpt, strays, nullptr);
    const size_t sh = (size_t)a.W * a.H * (3 * 8 + 16 + 8);
    const int nblk = c.ntiles * bpt;
    if (p.w) dcu_kernel<true><<<nblk, kThreads, sh, s>>>(a);
    else dcu_kernel<false><<<nblk, kThreads, sh, s>>>(a);
    AW_LAUNCH_CHECK();
}

int em_push(aw_ctx& c, const aw_particles& p, const int32_t* off, const float* E3,
            const float* B3, unsigned long long* acc8, const float cq[3], double qm, double dt,
            int bpt, unsigned long long* strays, double* ke_part, hipStream_t s) {
    if (p.n == 0) return 0;
    EmArgs a = em_args(c, p, off, E3, B3, acc8, cq, nullptr, qm, dt, bpt, strays, ke_part);
    const size_t sh = (size_t)a.W * a.H * (8 * 8 + 16 + 8);
    const int nblk = c.ntiles * bpt;
    const bool pw = p.w != nullptr, ke = ke_part != nullptr;
    if (pw && ke) em_push_kernel<true, true><<<nblk, kThreads, sh, s>>>(a);
    else if (pw) em_push_kernel<true, false><<<nblk, kThreads, sh, s>>>(a);
    else if (ke) em_push_kernel<false, true><<<nblk, kThreads, sh, s>>>(a);
    else em_push_kernel<false, false><<<nblk, kThreads, sh, s>>>(a);
    AW_LAUNCH_CHECK();
    return nblk;
}

void convert_multi(unsigned long long* acc, float* out, float* out2, int64_t ncell, int nq,
                   const double* inv, hipStream_t s) {
    convert_multi_kernel<<<blocks_for(ncell), kThreads, 0, s>>>(acc, out, out2, (int)ncell, nq,
                                                                inv);
    AW_LAUNCH_CHECK();
}

void form_e(float* E3, const float* EL2, const float* ET3, const aw_grid& g,
            const aw_run_params& rp, double t, hipStream_t s) {
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
    form_e_kernel<<<blocks_for(nc), kThreads, 0, s>>>(E3, EL2, ET3, g.nx, (int)nc, p);
    AW_LAUNCH_CHECK();
}

void dcu_scale(const float* E3, const float* B3, int64_t ncell, float* part, double coef_abs,
               double coef_signed, double umax, float* c_dcu, double* inv_dcu, hipStream_t s) {
    const int nb = (int)std::min<int64_t>(256, std::max<int64_t>(1, (ncell + 4095) / 4096));
    field_max_kernel<<<nb, kThreads, 0, s>>>(E3, B3, (int)ncell, part);
    AW_LAUNCH_CHECK();
    dcu_scale_kernel<<<1, 64, 0, s>>>(part, nb, coef_abs, coef_signed, umax, c_dcu, inv_dcu);
    AW_LAUNCH_CHECK();
}

}  // namespace aw
