// yee.hip -- the full-Maxwell (Yee) field model: FDTD field kernels, the
// particle push on the staggered mesh, the charge-conserving current, the
// delta-f weight equation, the mirror-equilibrium loader and the run's
// conservation diagnostics.
//
// Reference: MaxwellSimulation (include/pic/simulation_maxwell.hpp:24-227)
// and include/pic/yee2d.hpp (staggering :1-31, k_faraday / k_ampere :70-97,
// staggered gather :102-112, yee_advance_particle :145-302, k_binomial3x3
// :494-508, k_damp_x :531-545, k_cold_fluid :554-578, k_antenna :582-607,
// k_gauss_residual / k_div_stats / k_wd_stats / k_energy :614-678), the
// mirror loader (include/pic/particles.hpp:381-449,573-612) and the
// background profile (include/pic/background_b0.hpp:28-55).
//
// Staggering (periodic, cell units):
//   Ex(i+1/2, j)  Ey(i, j+1/2)  Ez(i, j)  Bx(i, j+1/2)  By(i+1/2, j)  Bz(i+1/2, j+1/2)
// One step: B half step (Faraday), push with E^n, B^n and the charge-conserving
// current of the move, periodic wrap, optional binomial current smoothing,
// antenna and cold-fluid currents, B half step, E full step (Ampere), optional
// x damping layers.
//
// MI355X design: the field kernels are one thread per cell over the
// row-major grid (coalesced rows); the push is one thread per particle,
// gathering from the L2-resident field arrays; the current is the library's
// Esirkepov deposit (int64 fixed point, tile-binned in LDS when dense), so
// the result is independent of thread order and of the particle order, which
// this engine never changes.  Every FP32 field and push operation is written
// out (the build disables contraction; fused multiply-adds are explicit), and
// the oracle restates the same sequence.
#include "esk_device.hpp"

#include <cmath>
#include <memory>
#include <vector>

namespace aw {

int tile_bpt(const aw_ctx& c, int64_t n);

namespace {

constexpr int kThreads = 256;
constexpr int kRedBlocks = 256;
constexpr int kPushBlocksPerCu = 8;  // grid-stride push: 8 workgroups of 256 per CU
constexpr int kPushApron = 2;        // tile push: cells of drift since the sort in LDS
constexpr int kRebinEvery = 8;       // steps between tile binnings without tile_sort

inline int blocks_for(int64_t n) { return (int)((n + kThreads - 1) / kThreads); }

// The nine grids of the mesh, each nx*ny floats, row-major.
struct Mesh {
    float *ex, *ey, *ez, *bx, *by, *bz, *jx, *jy, *jz;
    int nx, ny;
    float dx, dy, c2, dt;
};

__device__ __forceinline__ int wrap1(int i, int n) {
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

__device__ __forceinline__ int at(const Mesh& m, int i, int j) {
    return wrap1(j, m.ny) * m.nx + wrap1(i, m.nx);
}

// B += h * (-curl E)  (Faraday, half step h = dt/2), one cell
__device__ __forceinline__ void faraday_cell(const Mesh& m, float h, int c) {
    const int i = c % m.nx, j = c / m.nx;
    const float ez = m.ez[c];
    const float dez_y = (m.ez[at(m, i, j + 1)] - ez) / m.dy;
    const float dez_x = (m.ez[at(m, i + 1, j)] - ez) / m.dx;
    const float dey_x = (m.ey[at(m, i + 1, j)] - m.ey[c]) / m.dx;
    const float dex_y = (m.ex[at(m, i, j + 1)] - m.ex[c]) / m.dy;
    m.bx[c] = m.bx[c] - h * dez_y;
    m.by[c] = m.by[c] + h * dez_x;
    m.bz[c] = m.bz[c] - h * (dey_x - dex_y);
}

// E += dt * (c^2 curl B - J)  (Ampere, eps0 = 1), one cell
__device__ __forceinline__ void ampere_cell(const Mesh& m, int c) {
    const int i = c % m.nx, j = c / m.nx;
    const float bz = m.bz[c];
    const float dbz_y = (bz - m.bz[at(m, i, j - 1)]) / m.dy;
    const float dbz_x = (bz - m.bz[at(m, i - 1, j)]) / m.dx;
    const float dby_x = (m.by[c] - m.by[at(m, i - 1, j)]) / m.dx;
    const float dbx_y = (m.bx[c] - m.bx[at(m, i, j - 1)]) / m.dy;
    m.ex[c] = m.ex[c] + m.dt * (m.c2 * dbz_y - m.jx[c]);
    m.ey[c] = m.ey[c] + m.dt * (-m.c2 * dbz_x - m.jy[c]);
    m.ez[c] = m.ez[c] + m.dt * (m.c2 * (dby_x - dbx_y) - m.jz[c]);
}

// Faraday half step (include/pic/yee2d.hpp:70-81); zero_j also clears the
// current for the step's deposit
__global__ void __launch_bounds__(kThreads) faraday_kernel(Mesh m, float h, int zero_j) {
    const int c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= m.nx * m.ny) return;
    faraday_cell(m, h, c);
    if (zero_j) m.jx[c] = m.jy[c] = m.jz[c] = 0.f;
}

// Ampere step (include/pic/yee2d.hpp:84-97)
__global__ void __launch_bounds__(kThreads) ampere_kernel(Mesh m) {
    const int c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= m.nx * m.ny) return;
    ampere_cell(m, c);
}

// u' = u + u x t, u += u' x s with t = h B, s = 2t/(1 + t^2)
// (boris_rotate, include/pic/pusher.hpp:85-101)
__device__ __forceinline__ void rotate(float& ux, float& uy, float& uz, float bx, float by,
                                       float bz, float h) {
    const float tx = h * bx, ty = h * by, tz = h * bz;
    const float t2 = fmaf(tz, tz, fmaf(ty, ty, tx * tx));
    const float sf = 2.0f / (1.0f + t2);
    const float sx = sf * tx, sy = sf * ty, sz = sf * tz;
    const float upx = fmaf(uy, tz, fmaf(-uz, ty, ux));
    const float upy = fmaf(uz, tx, fmaf(-ux, tz, uy));
    const float upz = fmaf(ux, ty, fmaf(-uy, tx, uz));
    ux = fmaf(upy, sz, fmaf(-upz, sy, ux));
    uy = fmaf(upz, sx, fmaf(-upx, sz, uy));
    uz = fmaf(upx, sy, fmaf(-upy, sx, uz));
}

// The parabolic (b0_prof = 1) or fitted-polynomial (b0_prof = 2) background
// field along x and its derivative (include/pic/background_b0.hpp:28-55).
__device__ __forceinline__ float b0_profile(const aw_run_params& rp, float xph) {
    const float d = xph - (float)rp.b0_xc;
    if (rp.b0_prof == 2) {
        const float sn = d / (float)rp.b0_sref;
        const float u = sn * sn;
        float p = (float)rp.b0_c[5];
        for (int k = 4; k >= 0; --k) p = p * u + (float)rp.b0_c[k];
        return rp.B0[0] * p;
    }
    return rp.B0[0] * (1.f + (float)rp.b0_a * d * d);
}

__device__ __forceinline__ float b0_profile_dx(const aw_run_params& rp, float xph) {
    const float d = xph - (float)rp.b0_xc;
    if (rp.b0_prof == 2) {
        const float sn = d / (float)rp.b0_sref;
        const float u = sn * sn;
        float dp = 5.f * (float)rp.b0_c[5];
        for (int k = 4; k >= 1; --k) dp = dp * u + (float)k * (float)rp.b0_c[k];
        return rp.B0[0] * dp * 2.f * d / (float)(rp.b0_sref * rp.b0_sref);
    }
    return rp.B0[0] * 2.f * (float)rp.b0_a * d;
}

__device__ __forceinline__ double ramp_of(double t, double trmp, double toff) {
    if (t >= toff) return 0.0;
    if (t < trmp) return t / trmp;
    if (t > toff - trmp) return (toff - t) / trmp;
    return 1.0;
}

struct PushArgs {
    float *x, *y, *ux, *uy, *uz;
    const float* w;      // per-particle weight (NULL: w_uniform)
    float w_uniform;
    float *wd, *wc;      // delta-f weight and its Kahan compensation
    double* wdd;         // FP64 delta-f accumulator
    float *x0, *y0;      // out: positions before the move
    float* vz;           // out: deposit velocity uz / gamma (rel), else NULL
    float* wdep;         // out: deposit weight w * wd (delta-f), else NULL
    // the current's fixed scale holds while |uz/gamma| <= ucap and |w| <= wmax;
    // a particle past either sets *guard
    float ucap, wmax;
    unsigned int* guard;
    int n;
    double tnow;
};

// what the current deposit needs of one pushed particle
struct PushOut {
    float x0, y0, x1, y1;  // the move (x1 before the periodic wrap)
    float uz;              // deposit velocity: uz / gamma (rel, walls) or uz
    float w;               // deposit weight (w wd for delta-f, else w)
};

// yee_advance_particle (include/pic/yee2d.hpp:145-302): staggered gather of
// E (+ pump) and the wave B (+ B0 or the mirror profile), kick-rotate-kick
// (the rotation divided by the mid-kick gamma when rp.rel), the delta-f
// weight equation at the time-centred u, the move, and the reflecting x
// walls of bnd_x.  Positions are stored unwrapped (the wrap runs after the
// current deposit).
// The six fields sit on four staggerings: (ex, by) at (i + 1/2, j), (ey, bx)
// at (i, j + 1/2), ez at (i, j), bz at (i + 1/2, j + 1/2).  A gather is a
// staggering's CIC stencil, set up once and applied to its fields.
struct Cic {
    float fx, fy, gx, gy;
    int i0, j0;  // lower corner (cell index before wrapping)
    __device__ __forceinline__ Cic(float x, float y, float ox, float oy) {
        const float xs = x - ox, ys = y - oy;
        const float fi = floorf(xs), fj = floorf(ys);
        i0 = (int)fi;
        j0 = (int)fj;
        fx = xs - fi;
        fy = ys - fj;
        gx = 1.f - fx;
        gy = 1.f - fy;
    }
    __device__ __forceinline__ float mix(float v00, float v10, float v01, float v11) const {
        return fmaf(fx * fy, v11, fmaf(gx * fy, v01, fmaf(fx * gy, v10, (gx * gy) * v00)));
    }
};

__device__ __forceinline__ float cic_mesh(const float* a, const Mesh& m, const Cic& c) {
    const int i0 = dev::wrapfar(c.i0, m.nx), j0 = dev::wrapfar(c.j0, m.ny);
    const int i1 = i0 + 1 == m.nx ? 0 : i0 + 1, j1 = j0 + 1 == m.ny ? 0 : j0 + 1;
    return c.mix(a[j0 * m.nx + i0], a[j0 * m.nx + i1], a[j1 * m.nx + i0], a[j1 * m.nx + i1]);
}

constexpr float kOx[4] = {0.5f, 0.f, 0.f, 0.5f}, kOy[4] = {0.f, 0.5f, 0.f, 0.5f};

// gathers straight from the mesh; staggering s: 0 (ex, by), 1 (ey, bx), 2 ez, 3 bz
struct MeshGather {
    const Mesh& m;
    __device__ __forceinline__ void two(int s, float x, float y, float& a, float& b) const {
        const Cic c(x, y, kOx[s], kOy[s]);
        a = cic_mesh(s == 0 ? m.ex : m.ey, m, c);
        b = cic_mesh(s == 0 ? m.by : m.bx, m, c);
    }
    __device__ __forceinline__ float one(int s, float x, float y) const {
        const Cic c(x, y, kOx[s], kOy[s]);
        return cic_mesh(s == 2 ? m.ez : m.bz, m, c);
    }
};

// gathers from a workgroup's LDS copy of the fields over the cells
// [orgx, orgx + W) x [orgy, orgy + H) (periodically wrapped): planes of
// float2 (ex, by), float2 (ey, bx), float ez, float bz.  A stencil that
// leaves the window reads the mesh; the window holds the mesh's own values,
// so both return the same bits.
struct WindowGather {
    const float* s;
    const Mesh& m;
    int orgx, orgy, W, H;
    __device__ __forceinline__ bool local(const Cic& c, int& l) const {
        int li = c.i0 - orgx, lj = c.j0 - orgy;
        if ((unsigned)li > (unsigned)(W - 2)) li += li < 0 ? m.nx : -m.nx;
        if ((unsigned)lj > (unsigned)(H - 2)) lj += lj < 0 ? m.ny : -m.ny;
        l = lj * W + li;
        return (unsigned)li <= (unsigned)(W - 2) && (unsigned)lj <= (unsigned)(H - 2);
    }
    __device__ __forceinline__ void two(int st, float x, float y, float& a, float& b) const {
        const Cic c(x, y, kOx[st], kOy[st]);
        int l;
        if (!local(c, l)) {
            a = cic_mesh(st == 0 ? m.ex : m.ey, m, c);
            b = cic_mesh(st == 0 ? m.by : m.bx, m, c);
            return;
        }
        const float2* p = reinterpret_cast<const float2*>(s) + st * W * H + l;
        const float2 v00 = p[0], v10 = p[1], v01 = p[W], v11 = p[W + 1];
        a = c.mix(v00.x, v10.x, v01.x, v11.x);
        b = c.mix(v00.y, v10.y, v01.y, v11.y);
    }
    __device__ __forceinline__ float one(int st, float x, float y) const {
        const Cic c(x, y, kOx[st], kOy[st]);
        int l;
        if (!local(c, l)) return cic_mesh(st == 2 ? m.ez : m.bz, m, c);
        const float* p = s + (st + 2) * W * H + l;
        return c.mix(p[0], p[1], p[W], p[W + 1]);
    }
};

// One particle: the push updates u (and the delta-f weight) in place and
// returns the move; the caller stores the positions.
template <bool PLAIN, class Gather, bool NOPR = false>
__device__ __forceinline__ void push_one(const PushArgs& a, const Mesh& m, const aw_run_params& rp,
                                         int t, PushOut& o, const Gather& gat) {
    // PLAIN: none of the pump, relativistic push, B0 profile, delta-f weights
    // or walls; NOPR: no pump or relativistic push (compile-time instances
    // without their registers)
    const bool pump = !PLAIN && !NOPR && rp.pump, rel = !PLAIN && !NOPR && rp.rel;
    const bool b0_prof = !PLAIN && rp.b0_prof, deltaf = !PLAIN && rp.deltaf;
    const int bnd_x = PLAIN ? 0 : rp.bnd_x;
    const float x0 = a.x[t], y0 = a.y[t];
    float Ex, Ey, dBx, dBy;
    gat.two(0, x0, y0, Ex, dBy);
    gat.two(1, x0, y0, Ey, dBx);
    float Ez = gat.one(2, x0, y0);
    const float dBz = gat.one(3, x0, y0);
    if (pump && a.tnow < rp.pump_toff) {
        const double th = rp.pump_k0 * (double)x0 * (double)m.dx - rp.pump_w0 * a.tnow;
        const double r = ramp_of(a.tnow, rp.pump_trmp, rp.pump_toff);
        Ex += (float)(rp.pump_ex * cos(th) * r);
        Ey += (float)(-rp.pump_ey * sin(th) * r);
        Ez += (float)(rp.pump_ez * cos(th) * r);
    }
    float ux = a.ux[t], uy = a.uy[t], uz = a.uz[t];
    const float uxo = ux, uyo = uy, uzo = uz;  // u^{n-1/2}
    const float qmh = (float)(rp.qm * 0.5 * rp.dt);
    ux = fmaf(qmh, Ex, ux);
    uy = fmaf(qmh, Ey, uy);
    uz = fmaf(qmh, Ez, uz);
    const float gmid =
        rel ? 1.f / sqrtf(1.f + fmaf(uz, uz, fmaf(uy, uy, ux * ux))) : 1.f;
    if (b0_prof) {
        // B0 along x varying with x; the mirror force enters as the effective
        // transverse field (include/pic/yee2d.hpp:177-186)
        const float xph = x0 * m.dx;
        const float b0 = b0_profile(rp, xph);
        const float mc = b0_profile_dx(rp, xph) / (2.f * b0 * (float)rp.qm);
        rotate(ux, uy, uz, dBx + b0, dBy + mc * uz, dBz - mc * uy, qmh * gmid);
    } else {
        rotate(ux, uy, uz, dBx + rp.B0[0], dBy + rp.B0[1], dBz + rp.B0[2], qmh * gmid);
    }
    ux = fmaf(qmh, Ex, ux);
    uy = fmaf(qmh, Ey, uy);
    uz = fmaf(qmh, Ez, uz);
    const float gnew = rel ? 1.f / sqrtf(1.f + fmaf(uz, uz, fmaf(uy, uy, ux * ux))) : 1.f;
    float wdv = 0.f;
    if (deltaf) {
        // dwd/dt = -(1 - wd) (q/m) F . dln f0/du at u^n = (u^{n-1/2} + u^{n+1/2})/2,
        // F = E + v x dB (wave fields and pump), f0 the bi-Maxwellian with B0
        // along x (include/pic/yee2d.hpp:191-271)
        const float uxc = 0.5f * (ux + uxo), uyc = 0.5f * (uy + uyo), uzc = 0.5f * (uz + uzo);
        const float gc =
            rel ? 1.f / sqrtf(1.f + fmaf(uzc, uzc, fmaf(uyc, uyc, uxc * uxc))) : 1.f;
        const float Fx = Ex + gc * (uyc * dBz - uzc * dBy);
        const float Fy = Ey + gc * (uzc * dBx - uxc * dBz);
        const float Fz = Ez + gc * (uxc * dBy - uyc * dBx);
        float tperp = (float)rp.df_tperp;
        float t2 = (float)(rp.df_kappa * rp.df_tperp);
        if (b0_prof) {
            const float b = b0_profile(rp, x0 * m.dx) / rp.B0[0];
            const float par = (1.f - 1.f / b) / (float)rp.df_tpar;
            tperp = 1.f / (par + (1.f / b) / (float)rp.df_tperp);
            t2 = 1.f / (par + (1.f / b) / (float)(rp.df_kappa * rp.df_tperp));
        }
        float perp = 1.f / tperp;
        if (rp.df_dist == 1) {  // loss-cone subtracted bi-Maxwellian, capped
            const float u2p = fmaf(uzc, uzc, uyc * uyc);
            const float g = (float)rp.df_rho * expf(-0.5f * u2p * (1.f / t2 - 1.f / tperp));
            const float omg = fmaxf(1.f - g, 1e-3f);
            perp = fmaxf((1.f / tperp - g / t2) / omg, -10.f / tperp);
        }
        const float S = Fx * uxc / (float)rp.df_tpar + (Fy * uyc + Fz * uzc) * perp;
        const float relax = rp.df_taud > 0.0 ? (float)(1.0 - rp.dt / rp.df_taud) : 1.f;
        const float hi = 1.f - 1e-4f, lo = -8.f;
        float wn;
        if (rp.df_wprec == 2) {
            const double wo = a.wdd[t];
            double wnd = (wo + rp.dt * rp.qm * (1.0 - wo) * (double)S) * (double)relax;
            wnd = fmin(fmax(wnd, (double)lo), (double)hi);
            a.wdd[t] = wnd;
            wn = (float)wnd;
        } else if (rp.df_wprec == 1) {  // Kahan-compensated accumulation
            const float wd = a.wd[t];
            const float inc = (float)(rp.dt * rp.qm) * (1.f - wd) * S;
            const float cmp = a.wc[t];
            const float yk = inc - cmp;
            const float sum = wd + yk;
            a.wc[t] = (sum - wd) - yk;
            wn = fminf(fmaxf(sum * relax, lo), hi);
        } else {
            const float wd = a.wd[t];
            wn = fminf(fmaxf((wd + (float)(rp.dt * rp.qm) * (1.f - wd) * S) * relax, lo), hi);
        }
        a.wd[t] = wn;
        wdv = (a.w ? a.w[t] : a.w_uniform) * wn;
        if (a.wdep) a.wdep[t] = wdv;
    }
    float x1 = x0 + (float)((double)(ux * gnew) * rp.dt / (double)m.dx);
    const float y1 = y0 + (float)((double)(uy * gnew) * rp.dt / (double)m.dy);
    float uxs = ux, uys = uy, uzs = uz;
    if (bnd_x) {  // specular walls at the domain ends (include/pic/yee2d.hpp:276-297)
        const float fnx = (float)m.nx;
        if (x1 < 0.f) {
            x1 = -x1;
            uxs = -ux;
        } else if (x1 >= fnx) {
            x1 = 2.f * fnx - x1;
            uxs = -ux;
        }
        if (x1 >= fnx) x1 = nextafterf(fnx, 0.f);
        if (bnd_x == 2) {  // transverse momentum damped inside the layers
            const float nd = (float)rp.bnd_nd;
            float d = 0.f;
            if (x1 < nd) d = (nd - x1) / nd;
            else if (x1 > fnx - nd) d = (x1 - (fnx - nd)) / nd;
            if (d > 0.f) {
                const float f = expf((float)(-rp.bnd_numax * rp.dt) * d * d);
                uys = uy * f;
                uzs = uz * f;
            }
        }
    }
    a.ux[t] = uxs;
    a.uy[t] = uys;
    a.uz[t] = uzs;
    // the current's uz/gamma is taken before the walls damp u
    const float uzd = (rel || bnd_x) ? uz * gnew : uzs;
    const float wp = a.w ? a.w[t] : a.w_uniform;
    if (!(fabsf(uzd) <= a.ucap) || (a.w && !(fabsf(wp) <= a.wmax))) atomicOr(a.guard, 1u);
    o.x0 = x0;
    o.y0 = y0;
    o.x1 = x1;
    o.y1 = y1;
    o.uz = uzd;
    o.w = deltaf ? wdv : wp;
}

// MaxwellSimulation's particle push (include/pic/yee2d.hpp:140-300), grid-
// stride; stores the move for the current deposit that follows.  4 waves per
// SIMD fit without spills; without the pump and the relativistic push
// (NOPR: delta-f, profiles and walls stay) 5 waves do, which took the
// chen2026 delta-f line from 0.173 to 0.170 ms a step.
template <bool PLAIN, bool NOPR = false>
__global__ void __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(NOPR ? 5 : 4, NOPR ? 5 : 4)))
yee_push_kernel(PushArgs a, Mesh m, aw_run_params rp) {
    const MeshGather gat{m};
    for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < a.n;
         t += (int64_t)kThreads * gridDim.x) {
        PushOut o;
        push_one<PLAIN, MeshGather, NOPR>(a, m, rp, (int)t, o, gat);
        a.x0[t] = o.x0;
        a.y0[t] = o.y0;
        a.x[t] = o.x1;
        a.y[t] = o.y1;
        if (a.vz) a.vz[t] = o.uz;
        if (a.wdep) a.wdep[t] = o.w;
    }
}

// Particles in the tile order of an earlier sort: a workgroup per (tile,
// part) pushes its part of the tile's slice and deposits the Esirkepov
// current of each move on the spot (k_push_esirkepov_tiled,
// include/pic/yee2d.hpp:401-469).  The six fields sit in LDS over the tile
// with a D-cell apron for the drift since the sort (the mesh for what leaves
// it); the current goes to an LDS copy of the tile with a 4-node apron
// (global atomics for what leaves it), flushed at the end.  The move never
// touches HBM: the positions are stored wrapped, as the migrate does.
struct PushTiles {
    const int32_t* off;  // ntiles + 1
    const int32_t* idx;  // bin permutation (NULL: the particles are in tile order)
    int ntx, tx, ty, D, W, H, bpt;  // field window: W = tx + 2 D + 2, H = ty + 2 D + 2
    int JW, JH;                     // current window: tx + 2 A, ty + 2 A
};

// Register budgets, measured on the 156 M-particle eaw_case7 step: the
// general instance at 4 waves per SIMD (128 VGPRs, 12 spilled) took 3.32 ms
// against 3.81 ms at the compiler's 152 VGPRs (3 waves) and 4.08 ms at 5
// waves (39 spilled); the plain instance at 6 waves (80 VGPRs, 4 spilled),
// which is also as many as its 25 KB of LDS lets a CU hold, took 2.70 ms
// against 3.04 ms at its own 91 VGPRs (5 waves).
template <bool ONE_ROW, bool PLAIN>
__global__ void __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(PLAIN ? 6 : 4, PLAIN ? 6 : 4)))
yee_push_esk_tiles_kernel(PushArgs a, Mesh m, aw_run_params rp, PushTiles t, esk::EskArgs e,
                          unsigned long long* __restrict__ acc) {
    extern __shared__ unsigned long long s_mem[];
    const int RJ = t.JW * t.JH, R = t.W * t.H;
    unsigned long long* s_j = s_mem;                       // [3][JH][JW]
    float* s_f = reinterpret_cast<float*>(s_mem + 3 * RJ);  // field planes
    const int lb = dev::xcd_block(blockIdx.x, gridDim.x);
    const int tile = lb / t.bpt, part = lb % t.bpt;
    const int tx0 = (tile % t.ntx) * t.tx, ty0 = (tile / t.ntx) * t.ty;
    const int orgx = tx0 - t.D - 1, orgy = ty0 - t.D - 1;
    const int jorgx = tx0 - esk::kSortedApron, jorgy = ty0 - esk::kSortedApron;
    float2* s2 = reinterpret_cast<float2*>(s_f);
    // (row, column) of a thread's cells advance by the stride without dividing
    const int dr = kThreads / t.W, dc = kThreads % t.W;
    int row = threadIdx.x / t.W, col = threadIdx.x % t.W;
    for (int l = threadIdx.x; l < R; l += kThreads) {
        const int g = dev::wrapfar(orgy + row, m.ny) * m.nx + dev::wrapfar(orgx + col, m.nx);
        row += dr;
        col += dc;
        if (col >= t.W) {
            col -= t.W;
            ++row;
        }
        s2[l] = make_float2(m.ex[g], m.by[g]);
        s2[R + l] = make_float2(m.ey[g], m.bx[g]);
        s_f[4 * R + l] = m.ez[g];
        s_f[5 * R + l] = m.bz[g];
    }
    for (int c = threadIdx.x; c < 3 * RJ; c += kThreads) s_j[c] = 0ull;
    __syncthreads();
    const WindowGather gat{s_f, m, orgx, orgy, t.W, t.H};
    double kc[3];
    esk::esk_coefs(e, kc);
    const int64_t nc = (int64_t)m.nx * m.ny;
    const int beg = t.off[tile], end = t.off[tile + 1];
    const int chunk = (end - beg + t.bpt - 1) / t.bpt;
    const int r0 = beg + part * chunk, r1 = min(end, r0 + chunk);
    for (int r = r0 + (int)threadIdx.x; r < r1; r += kThreads) {
        const int i = t.idx ? t.idx[r] : r;
        PushOut o;
        push_one<PLAIN>(a, m, rp, i, o, gat);
        a.x[i] = dev::wrap_coord(o.x1, (float)m.nx);
        a.y[i] = dev::wrap_coord(o.y1, (float)m.ny);
        esk::EskMove mv;
        esk::esk_prepare_v(e, o.x0, o.y0, o.x1, o.y1, o.uz, e.per_w ? (double)o.w : e.w_uniform,
                           kc, mv);
        const int li = esk::wrap_local(mv.sh.ib - 1 - jorgx, m.nx);
        const int lj = esk::wrap_local(mv.sh.jb - 1 - jorgy, m.ny);
        if (dev::in_tile(li, t.JW - 4) && dev::in_tile(lj, t.JH - 4)) {
            const int base = lj * t.JW + li;
            esk::esk_emit<ONE_ROW>(mv, [&](int q, int k, int b, unsigned long long v) {
                atomicAdd(&s_j[q * RJ + base + b * t.JW + k], v);
            });
        } else {
            esk::esk_emit<ONE_ROW>(mv, [&](int q, int k, int b, unsigned long long v) {
                atomicAdd(&acc[q * nc + (int64_t)dev::wrapfar(mv.sh.jb - 1 + b, m.ny) * m.nx +
                               dev::wrapfar(mv.sh.ib - 1 + k, m.nx)],
                          v);
            });
        }
    }
    __syncthreads();
    // rows of the three components follow each other: (row, column) over
    // 3 JH rows, advanced by the stride without dividing
    const int jdr = kThreads / t.JW, jdc = kThreads % t.JW;
    int jrow = threadIdx.x / t.JW, jcol = threadIdx.x % t.JW;
    for (int c = threadIdx.x; c < 3 * RJ; c += kThreads) {
        const unsigned long long v = s_j[c];
        if (v) {
            const int q = jrow >= 2 * t.JH ? 2 : (jrow >= t.JH ? 1 : 0);
            const int gx = dev::wrapfar(jorgx + jcol, m.nx);
            const int gy = dev::wrapfar(jorgy + jrow - q * t.JH, m.ny);
            atomicAdd(&acc[q * nc + (int64_t)gy * m.nx + gx], v);
        }
        jrow += jdr;
        jcol += jdc;
        if (jcol >= t.JW) {
            jcol -= t.JW;
            ++jrow;
        }
    }
}

// per-workgroup max |a[i]|
__global__ void __launch_bounds__(kThreads) max_abs_kernel(const float* __restrict__ a, int64_t n,
                                                           float* __restrict__ part) {
    float v = 0.f, u = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (int64_t)kThreads * gridDim.x)
        v = fmaxf(v, fabsf(a[i]));
    dev::block_max2(v, u);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void gather_kernel(const float* __restrict__ src, const int32_t* __restrict__ perm,
                              float* __restrict__ dst, int n) {
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k < n) dst[k] = src[perm[k]];
}

__global__ void zero_kernel(float* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// [1,2,1] x [1,2,1] / 16 at cell c of one periodic grid (include/pic/yee2d.hpp:494-508)
__device__ __forceinline__ float binomial_at(const float* src, const Mesh& m, int c) {
    const int i = c % m.nx, j = c / m.nx;
    const float w[3] = {0.25f, 0.5f, 0.25f};
    float acc = 0.f;
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
        for (int di = -1; di <= 1; ++di)
            acc = fmaf(w[di + 1] * w[dj + 1], src[at(m, i + di, j + dj)], acc);
    return acc;
}

// one smoothing pass over grid blockIdx.y of consecutive nx*ny grids
__global__ void __launch_bounds__(kThreads)
binomial_kernel(const float* __restrict__ src, float* __restrict__ dst, Mesh m) {
    const int64_t nc = (int64_t)m.nx * m.ny;
    const int c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= nc) return;
    dst[blockIdx.y * nc + c] = binomial_at(src + blockIdx.y * nc, m, c);
}

// the current's conversion (J += acc 2^-S) and every smoothing pass of
// component blockIdx.x in LDS, for grids of up to kSmoothLdsCells
__global__ void __launch_bounds__(1024)
convert_smooth_kernel(float* __restrict__ j, Mesh m, int passes,
                      const unsigned long long* __restrict__ acc, esk::EskArgs e) {
    extern __shared__ float s_c[];
    const int nc = m.nx * m.ny, q = blockIdx.x;
    int S[3];
    esk::esk_exponents(e, S);
    float* g = j + (int64_t)q * nc;
    float *cur = s_c, *nxt = s_c + nc;
    for (int c = threadIdx.x; c < nc; c += blockDim.x)
        cur[c] = g[c] + esk::esk_value((long long)acc[(int64_t)q * nc + c], S[q]);
    __syncthreads();
    for (int k = 0; k < passes; ++k) {
        for (int c = threadIdx.x; c < nc; c += blockDim.x) nxt[c] = binomial_at(cur, m, c);
        __syncthreads();
        float* t = cur;
        cur = nxt;
        nxt = t;
    }
    for (int c = threadIdx.x; c < nc; c += blockDim.x) g[c] = cur[c];
}

// every pass of grid blockIdx.x in LDS, for grids of up to kSmoothLdsCells
constexpr int kSmoothThreads = 1024;
constexpr int kSmoothLdsCells = 8192;
__global__ void __launch_bounds__(kSmoothThreads)
binomial_lds_kernel(float* __restrict__ a, Mesh m, int passes) {
    extern __shared__ float s_b[];
    const int nc = m.nx * m.ny;
    float* g = a + (int64_t)blockIdx.x * nc;
    float *cur = s_b, *nxt = s_b + nc;
    for (int c = threadIdx.x; c < nc; c += kSmoothThreads) cur[c] = g[c];
    __syncthreads();
    for (int k = 0; k < passes; ++k) {
        for (int c = threadIdx.x; c < nc; c += kSmoothThreads) nxt[c] = binomial_at(cur, m, c);
        __syncthreads();
        float* t = cur;
        cur = nxt;
        nxt = t;
    }
    for (int c = threadIdx.x; c < nc; c += kSmoothThreads) g[c] = cur[c];
}

// the rotating transverse antenna current (include/pic/yee2d.hpp:582-607), one cell
__device__ __forceinline__ void antenna_cell(const Mesh& m, const aw_run_params& rp, double t,
                                             int c) {
    const int i = c % m.nx;
    const double d = i - rp.ant_x0;
    if (fabs(d) > 4.0 * rp.ant_sigma) return;
    if (rp.ant_toff > 0 && t >= rp.ant_toff) return;
    double r = 1.0;
    if (rp.ant_trmp > 0) {
        if (t < rp.ant_trmp) r = t / rp.ant_trmp;
        else if (rp.ant_toff > 0 && t > rp.ant_toff - rp.ant_trmp)
            r = (rp.ant_toff - t) / rp.ant_trmp;
    }
    const double g = rp.ant_amp * r * exp(-d * d / (2.0 * rp.ant_sigma * rp.ant_sigma));
    const double ph = rp.ant_w0 * t;
    m.jy[c] += (float)(g * cos(ph));
    m.jz[c] += (float)(g * sin(ph));
}

// linearised cold electron fluid on the nodes, ny = 1 (include/pic/yee2d.hpp:554-578), one cell
__device__ __forceinline__ void cold_cell(const Mesh& m, const aw_run_params& rp,
                                          float* __restrict__ vcy, float* __restrict__ vcz,
                                          int c) {
    const int i = c % m.nx;
    const float Ey = m.ey[c], Ez = m.ez[c];
    const float h = (float)(0.5 * rp.dt * rp.qm);
    float uy = vcy[c] + h * Ey;
    float uz = vcz[c] + h * Ez;
    const float b0 = rp.b0_prof ? b0_profile(rp, i * m.dx) : rp.B0[0];
    float sn, cs;
    sincosf((float)(-rp.qm * rp.dt) * b0, &sn, &cs);
    const float uy2 = uy * cs - uz * sn;
    const float uz2 = uy * sn + uz * cs;
    uy = uy2 + h * Ey;
    uz = uz2 + h * Ez;
    vcy[c] = uy;
    vcz[c] = uz;
    const float jc = (float)(rp.qm * rp.cold_nc);
    m.jy[c] += jc * uy;
    m.jz[c] += jc * uz;
}

// x damping layers: node mask on Ey, Ez, Bx (and the cold fluid), half-node
// mask on Ex, By, Bz (include/pic/yee2d.hpp:531-545), one cell
__device__ __forceinline__ void damp_cell(const Mesh& m, const float* __restrict__ mn,
                                          const float* __restrict__ mh, float* vcy, float* vcz,
                                          int c) {
    const int i = c % m.nx;
    const float a = mn[i], b = mh[i];
    if (a == 1.f && b == 1.f) return;
    m.ey[c] *= a;
    m.ez[c] *= a;
    m.bx[c] *= a;
    m.ex[c] *= b;
    m.by[c] *= b;
    m.bz[c] *= b;
    if (vcy) {
        vcy[c] *= a;
        vcz[c] *= a;
    }
}

// the field side between the particles' current and Ampere, per cell:
// [J cleared], antenna, cold fluid, the second Faraday half step.  None of
// them reads what another cell's update writes (Faraday reads E only).
__global__ void __launch_bounds__(kThreads)
field_mid_kernel(Mesh m, aw_run_params rp, double t, float h, int clear, int ant,
                 float* __restrict__ vcy, float* __restrict__ vcz) {
    const int c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= m.nx * m.ny) return;
    if (clear) m.jx[c] = m.jy[c] = m.jz[c] = 0.f;
    if (ant) antenna_cell(m, rp, t, c);
    if (vcy) cold_cell(m, rp, vcy, vcz, c);
    faraday_cell(m, h, c);
}

__global__ void damp_kernel(Mesh m, const float* __restrict__ mn, const float* __restrict__ mh,
                            float* vcy, float* vcz) {
    const int c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= m.nx * m.ny) return;
    damp_cell(m, mn, mh, vcy, vcz, c);
}

// Gauss residual at the nodes: div E - rho (include/pic/yee2d.hpp:614-623)
__global__ void gauss_kernel(Mesh m, const float* __restrict__ rho, float* __restrict__ r) {
    const int c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= m.nx * m.ny) return;
    const int i = c % m.nx, j = c / m.nx;
    const float dive = (m.ex[c] - m.ex[at(m, i - 1, j)]) / m.dx +
                       (m.ey[c] - m.ey[at(m, i, j - 1)]) / m.dy;
    r[c] = dive - rho[c];
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kThreads / 64; ++w) t += sh[w];
    __syncthreads();
    return t;
}

// Per-block FP64 partials over fixed slices of the cells (deterministic):
// what = 0: {0.5 |E|^2 dA, 0.5 c^2 |B|^2 dA}; 1: {divB^2, (r - r0)^2}.
__global__ void __launch_bounds__(kThreads)
mesh_stats_kernel(Mesh m, int what, const float* __restrict__ r, const float* __restrict__ r0,
                  double* __restrict__ part) {
    __shared__ double sh[kThreads / 64];
    const int64_t nc = (int64_t)m.nx * m.ny;
    const int64_t per = (nc + gridDim.x - 1) / gridDim.x;
    const int64_t b = (int64_t)blockIdx.x * per, e = min(nc, b + per);
    double s0 = 0.0, s1 = 0.0;
    const double dA = (double)m.dx * (double)m.dy;
    for (int64_t c = b + threadIdx.x; c < e; c += kThreads) {
        if (what == 0) {
            s0 += 0.5 * ((double)m.ex[c] * m.ex[c] + (double)m.ey[c] * m.ey[c] +
                         (double)m.ez[c] * m.ez[c]) * dA;
            s1 += 0.5 * (double)m.c2 *
                  ((double)m.bx[c] * m.bx[c] + (double)m.by[c] * m.by[c] +
                   (double)m.bz[c] * m.bz[c]) * dA;
        } else {
            const int i = (int)(c % m.nx), j = (int)(c / m.nx);
            const float divb = (m.bx[at(m, i + 1, j)] - m.bx[c]) / m.dx +
                               (m.by[at(m, i, j + 1)] - m.by[c]) / m.dy;
            s0 += (double)divb * divb;
            const double d = (double)r[c] - r0[c];
            s1 += d * d;
        }
    }
    const double t0 = block_sum(s0, sh), t1 = block_sum(s1, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = t0;
        part[2 * blockIdx.x + 1] = t1;
    }
}

// per-block {sum wd, sum wd^2, max |wd|} over fixed particle slices
__global__ void __launch_bounds__(kThreads)
wd_stats_kernel(const float* __restrict__ wd, int64_t n, double* __restrict__ part) {
    __shared__ double sh[kThreads / 64];
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b = (int64_t)blockIdx.x * per, e = min(n, b + per);
    double s1 = 0.0, s2 = 0.0, mx = 0.0;
    for (int64_t i = b + threadIdx.x; i < e; i += kThreads) {
        const double v = wd[i];
        s1 += v;
        s2 += v * v;
        mx = fmax(mx, fabs(v));
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
    __shared__ double shm[kThreads / 64];
    if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = mx;
    const double t1 = block_sum(s1, sh), t2 = block_sum(s2, sh);
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) m = fmax(m, shm[w]);
        part[3 * blockIdx.x] = t1;
        part[3 * blockIdx.x + 1] = t2;
        part[3 * blockIdx.x + 2] = m;
    }
}

// the diagnostics' view of the fields: wave E + pump, wave B + B0, and ELx = Ex
// (tools/arcsim.cu:121-142)
__global__ void diag_fields_kernel(Mesh m, aw_run_params rp, double t, int pump,
                                   float* __restrict__ out) {
    const int c = blockIdx.x * kThreads + threadIdx.x;
    const int nc = m.nx * m.ny;
    if (c >= nc) return;
    float ex = m.ex[c], ey = m.ey[c], ez = m.ez[c];
    if (pump) {
        const int i = c % m.nx;
        const double th = rp.pump_k0 * (i * (double)m.dx) - rp.pump_w0 * t;
        const double r = ramp_of(t, rp.pump_trmp, rp.pump_toff);
        ex += (float)(rp.pump_ex * cos(th) * r);
        ey += (float)(-rp.pump_ey * sin(th) * r);
        ez += (float)(rp.pump_ez * cos(th) * r);
    }
    out[c] = ex;
    out[nc + c] = ey;
    out[2 * nc + c] = ez;
    out[3 * nc + c] = m.bx[c] + rp.B0[0];
    out[4 * nc + c] = m.by[c] + rp.B0[1];
    out[5 * nc + c] = m.bz[c] + rp.B0[2];
    out[6 * nc + c] = m.ex[c];
}

struct MirrorLoad {
    const int32_t* off;  // nx + 1 marker offsets per cell
    int nx;
    float dx;
    double uth_par, uth_perp_eq, lc_rho, lc_kappa;
    float weight;
    uint64_t seed;
    int dist;
    int64_t n;
};

// mirror_init_kernel (include/pic/particles.hpp:390-449): equal-weight
// markers of the (E, mu)-mapped bi-Maxwellian (or loss-cone) in the profile,
// marker t in cell c with off[c] <= t < off[c+1], hashed draws of index t.
__global__ void mirror_load_kernel(MirrorLoad L, aw_run_params rp, float* __restrict__ x,
                                   float* __restrict__ y, float* __restrict__ ux,
                                   float* __restrict__ uy, float* __restrict__ uz,
                                   float* __restrict__ w) {
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= L.n) return;
    int lo = 0, hi = L.nx;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (L.off[mid] <= t) lo = mid;
        else hi = mid;
    }
    const int ti = (int)t;
    float xv = (float)lo + (float)dev::rng_uniform(ti, 0, L.seed);
    if (xv >= (float)L.nx) xv = nextafterf((float)L.nx, 0.f);
    const double b = (double)b0_profile(rp, xv * L.dx) / (double)rp.B0[0];
    const double tpa = L.uth_par * L.uth_par, tpe = L.uth_perp_eq * L.uth_perp_eq;
    const double tp = 1.0 / ((1.0 - 1.0 / b) / tpa + (1.0 / b) / tpe);
    const double r1 = fmax(dev::rng_uniform(ti, 1, L.seed), 1e-12);
    const double r2 = dev::rng_uniform(ti, 2, L.seed);
    const double r4 = dev::rng_uniform(ti, 4, L.seed);
    const double two_pi = 6.28318530717958647692;
    const double upar = L.uth_par * sqrt(-2.0 * log(r1)) * cos(two_pi * r2);
    double uperp;
    if (L.dist == 1) {  // rejection off the Rayleigh(Tp) envelope
        const double t2 = 1.0 / ((1.0 - 1.0 / b) / tpa + (1.0 / b) / (L.lc_kappa * tpe));
        const double dinv = 0.5 * (1.0 / t2 - 1.0 / tp);
        double u2 = 0.0;
        for (int k = 0; k < 64; ++k) {
            const double rm = fmax(dev::rng_uniform(ti, 10 + 2 * k, L.seed), 1e-12);
            const double ra = dev::rng_uniform(ti, 11 + 2 * k, L.seed);
            u2 = -2.0 * tp * log(rm);
            if (ra < 1.0 - L.lc_rho * exp(-u2 * dinv)) break;
        }
        uperp = sqrt(u2);
    } else {
        const double r3 = fmax(dev::rng_uniform(ti, 3, L.seed), 1e-12);
        uperp = sqrt(tp) * sqrt(-2.0 * log(r3));
    }
    x[t] = xv;
    y[t] = 0.5f;
    ux[t] = (float)upar;
    uy[t] = (float)(uperp * cos(two_pi * r4));
    uz[t] = (float)(uperp * sin(two_pi * r4));
    if (w) w[t] = L.weight;
}

// Per-cell marker offsets of the mirror load (include/pic/particles.hpp:573-597):
// count(c) = lround(ppc * n(x_c)/n_eq) at the cell centre.
std::vector<int32_t> mirror_offsets(const aw_grid& g, const aw_run_params& rp,
                                    const aw_species& q) {
    auto prof = [&](float xph) {  // the host twin of b0_profile
        const float d = xph - (float)rp.b0_xc;
        if (rp.b0_prof == 2) {
            const float sn = d / (float)rp.b0_sref;
            const float u = sn * sn;
            float p = (float)rp.b0_c[5];
            for (int k = 4; k >= 0; --k) p = p * u + (float)rp.b0_c[k];
            return rp.B0[0] * p;
        }
        return rp.B0[0] * (1.f + (float)rp.b0_a * d * d);
    };
    const double tpa = q.uth[0] * q.uth[0], tpe = q.uth[1] * q.uth[1];
    std::vector<int32_t> off((size_t)g.nx + 1, 0);
    int64_t tot = 0;
    for (int c = 0; c < g.nx; ++c) {
        const double b = (double)prof((c + 0.5f) * (float)g.dx) / (double)rp.B0[0];
        const double t1 = 1.0 / ((1.0 - 1.0 / b) / tpa + (1.0 / b) / tpe);
        double nfac = t1 / tpe;
        if (q.dist == 1) {
            const double t2 = 1.0 / ((1.0 - 1.0 / b) / tpa + (1.0 / b) / (q.lc_kappa * tpe));
            nfac = (t1 - q.lc_rho * t2) / (tpe * (1.0 - q.lc_rho * q.lc_kappa));
        }
        off[c] = (int32_t)tot;
        tot += std::lround(q.ppc * nfac);
        AW_REQUIRE(tot < (int64_t)INT32_MAX, "mirror load: marker count must fit int32");
    }
    off[g.nx] = (int32_t)tot;
    return off;
}

void check_mirror(const aw_grid& g, const aw_run_params& rp, const aw_species& q) {
    AW_REQUIRE(g.ny == 1, "initialize_mirror: ny must be 1");
    AW_REQUIRE(rp.b0_prof != 0, "initialize_mirror: rp.b0_prof required");
    AW_REQUIRE(q.uth[1] == q.uth[2], "initialize_mirror: gyrotropy uth[1] == uth[2] required");
    AW_REQUIRE(q.dist != 1 || (q.lc_rho >= 0.0 && q.lc_rho <= 1.0 && q.lc_kappa > 0.0 &&
                               q.lc_kappa < 1.0),
               "initialize_mirror: losscone needs 0<=rho<=1, 0<kappa<1");
    AW_REQUIRE(q.ppc > 0, "species ppc must be positive");
}

}  // namespace
}  // namespace aw

using namespace aw;

struct aw_maxwell {
    aw_grid g{};
    aw_run_params rp{};
    std::unique_ptr<aw_ctx> ctx;  // Esirkepov binning and the rho accumulator
    hipStream_t stream = nullptr;
    DevBuf<float> F;              // ex ey ez bx by bz jx jy jz
    DevBuf<float> jtmp;           // smoothing scratch
    DevBuf<float> mask;           // [node | half-node] x damping masks
    DevBuf<float> vc;             // [vcy | vcz] cold fluid
    DevBuf<float> x0, y0, vz, wdep;
    DevBuf<unsigned int> jguard;  // set when a particle leaves the current's fixed scale
    DevBuf<float> wpart;          // per-workgroup max |w|
    float wmax_val = 0.f;         // max |w| of the weights at (wmax_ptr, wmax_n)
    const float* wmax_ptr = nullptr;
    int64_t wmax_n = -1;
    DevBuf<float> rho, rres, r0, dfields;
    DevBuf<double> part;
    bool have_ref = false;
    int64_t nstep = 0;
    // rp.tile_sort > 0: particles kept in 16x16-tile order, re-sorted every
    // tile_sort steps; the current deposit walks the tile slices of the last sort
    DevBuf<float> sbuf;           // sort staging: x y ux uy uz [w] [wd]
    DevBuf<int32_t> soff, sperm;  // tile offsets of the last sort; its permutation
    int64_t next_sort = 0, sorted_n = -1;
    const float* sorted_x = nullptr;

    int64_t nc() const { return (int64_t)g.nx * g.ny; }

    Mesh mesh() const {
        Mesh m{};
        float* f = F.get();
        const int64_t n = nc();
        m.ex = f;
        m.ey = f + n;
        m.ez = f + 2 * n;
        m.bx = f + 3 * n;
        m.by = f + 4 * n;
        m.bz = f + 5 * n;
        m.jx = f + 6 * n;
        m.jy = f + 7 * n;
        m.jz = f + 8 * n;
        m.nx = g.nx;
        m.ny = g.ny;
        m.dx = (float)g.dx;
        m.dy = (float)g.dy;
        m.c2 = (float)(rp.c * rp.c);
        m.dt = (float)rp.dt;
        return m;
    }

    ~aw_maxwell() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void zero(float* p, int64_t n) {
        if (n <= 0) return;
        zero_kernel<<<(int)((n + kThreads - 1) / kThreads), kThreads, 0, stream>>>(p, n);
        AW_LAUNCH_CHECK();
    }

    void build_masks() {
        std::vector<float> h(2 * (size_t)g.nx);
        const int nd = rp.bnd_nd;
        auto val = [&](double xi) {
            double d = 0.0;
            if (xi < nd) d = (nd - xi) / (double)nd;
            else if (xi > g.nx - nd) d = (xi - (g.nx - nd)) / (double)nd;
            return d <= 0.0 ? 1.0 : std::exp(-rp.bnd_numax * d * d * rp.dt);
        };
        for (int i = 0; i < g.nx; ++i) {
            h[i] = (float)val(i);
            h[g.nx + i] = (float)val(i + 0.5);
        }
        mask.resize(h.size());
        AW_HIP(hipMemcpy(mask.get(), h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    }

    // rp.jfilter binomial passes, in place, over ncomp consecutive grids
    void smooth(float* a, int ncomp) {
        const int J = rp.jfilter;
        if (J <= 0) return;
        const Mesh m = mesh();
        const int64_t n = nc();
        const size_t lds = (size_t)(2 * n) * sizeof(float);
        if (n <= kSmoothLdsCells && lds <= ctx->lds_per_block) {
            binomial_lds_kernel<<<ncomp, kSmoothThreads, lds, stream>>>(a, m, J);
            AW_LAUNCH_CHECK();
            return;
        }
        jtmp.reserve((size_t)(2 * ncomp * n));
        float* t[2] = {jtmp.get(), jtmp.get() + ncomp * n};
        const dim3 grid(blocks_for(n), ncomp);
        if (J == 1) {
            binomial_kernel<<<grid, kThreads, 0, stream>>>(a, t[0], m);
            AW_LAUNCH_CHECK();
            AW_HIP(hipMemcpyAsync(a, t[0], (size_t)(ncomp * n) * sizeof(float),
                                  hipMemcpyDeviceToDevice, stream));
            return;
        }
        // the first pass reads a, the last writes it, the others ping-pong
        for (int k = 0; k < J; ++k) {
            const float* src = k == 0 ? a : t[(k - 1) & 1];
            float* dst = k == J - 1 ? a : t[k & 1];
            binomial_kernel<<<grid, kThreads, 0, stream>>>(src, dst, m);
            AW_LAUNCH_CHECK();
        }
    }

    // max |w| over the particle weights (|w_uniform| without them), taken
    // once per weight array; a later change of a weight past it sets the guard
    float weight_max(const aw_particles& p) {
        if (!p.w) return std::fabs((float)p.w_uniform);
        if (p.w != wmax_ptr || p.n != wmax_n) {
            const int nbk = std::min(blocks_for(p.n), kRedBlocks);
            wpart.reserve((size_t)nbk);
            max_abs_kernel<<<nbk, kThreads, 0, stream>>>(p.w, p.n, wpart.get());
            AW_LAUNCH_CHECK();
            std::vector<float> h((size_t)nbk);
            AW_HIP(hipMemcpyAsync(h.data(), wpart.get(), h.size() * sizeof(float),
                                  hipMemcpyDeviceToHost, stream));
            AW_HIP(hipStreamSynchronize(stream));
            float v = 0.f;
            for (float x : h) v = std::max(v, x);
            AW_REQUIRE(std::isfinite(v), "a particle weight is not finite");
            wmax_val = v;
            wmax_ptr = p.w;
            wmax_n = p.n;
        }
        return wmax_val;
    }

    // Reports (once) a step whose current left its fixed scale.
    void check_guard() {
        unsigned int f = 0;
        AW_HIP(hipStreamSynchronize(stream));
        AW_HIP(hipMemcpy(&f, jguard.get(), sizeof(f), hipMemcpyDeviceToHost));
        if (f) {
            AW_HIP(hipMemset(jguard.get(), 0, sizeof(unsigned int)));
            fail(AW_ERR_INVALID,
                 "a particle's |uz/gamma| exceeded 2 max(1, c), or a weight exceeded the "
                 "run's largest: the current of that step left its fixed-point scale",
                 __FILE__, __LINE__);
        }
    }

    // J += the accumulated current, then rp.jfilter smoothing passes: one LDS
    // launch for the three components of a small grid
    void convert_and_smooth(const esk::EskArgs& e) {
        const Mesh m = mesh();
        const int64_t n = nc();
        const size_t lds = (size_t)(2 * n) * sizeof(float);
        if (rp.jfilter > 0 && n <= kSmoothLdsCells && lds <= ctx->lds_per_block) {
            convert_smooth_kernel<<<3, kSmoothThreads, lds, stream>>>(m.jx, m, rp.jfilter,
                                                                     ctx->jacc.get(), e);
            AW_LAUNCH_CHECK();
            return;
        }
        esk::convert(*ctx, e, m.jx, m.jy, m.jz, stream);
        smooth(m.jx, 3);
    }

    // Stable physical sort of the particles into tile order, in place in the
    // caller's arrays (Particles::sort_by_tile, include/pic/particles.hpp:677-726)
    void tile_sort(aw_particles& p, const aw_deltaf_state* df) {
        AW_REQUIRE(!rp.deltaf || rp.df_wprec == 0,
                   "tile_sort needs df_wprec = 0 (the weight-precision accumulators are not "
                   "re-sorted); set tile_sort = 0");
        const size_t n = (size_t)p.n;
        const bool has_w = p.w != nullptr;
        const int nf = 5 + (has_w ? 1 : 0) + (rp.deltaf ? 1 : 0);
        sbuf.reserve((size_t)nf * n);
        float* b = sbuf.get();
        float* const from[7] = {p.x, p.y, p.ux, p.uy, p.uz, has_w ? p.w : nullptr,
                                rp.deltaf ? df->wd : nullptr};
        aw_particles src = p;
        float** to[6] = {&src.x, &src.y, &src.ux, &src.uy, &src.uz, &src.w};
        int k = 0;
        for (int f = 0; f < 7; ++f) {
            if (!from[f]) continue;
            float* d = b + (size_t)k * n;
            AW_HIP(hipMemcpyAsync(d, from[f], n * sizeof(float), hipMemcpyDeviceToDevice,
                                  stream));
            if (f < 6) *to[f] = d;
            ++k;
        }
        const float* wd_src = rp.deltaf ? b + (size_t)(nf - 1) * n : nullptr;
        soff.reserve((size_t)ctx->ntiles + 1);
        int32_t* perm = nullptr;
        if (rp.deltaf) {
            sperm.reserve(n);
            perm = sperm.get();
        }
        if (!tile_sort_onepass(*ctx, src, p, soff.get(), perm, stream)) {
            sperm.reserve(n);
            perm = sperm.get();
            bin_tiles(*ctx, src, perm, soff.get(), stream);
            permute_particles(src, p, perm, stream);
        }
        if (rp.deltaf) {
            gather_kernel<<<blocks_for(p.n), kThreads, 0, stream>>>(wd_src, perm, df->wd,
                                                                  (int)p.n);
            AW_LAUNCH_CHECK();
        }
        sorted_n = p.n;
        sorted_x = p.x;
    }

    // MaxwellSimulation::step_at (include/pic/simulation_maxwell.hpp:52-109)
    void step(aw_particles& p, const aw_deltaf_state* df, double tnow) {
        const Mesh m = mesh();
        const int nb = blocks_for(nc());
        const float h = 0.5f * (float)rp.dt;
        // the first half step also clears J for the particles' current
        faraday_kernel<<<nb, kThreads, 0, stream>>>(m, h, p.n > 0 ? 1 : 0);
        AW_LAUNCH_CHECK();
        if (p.n > 0) {
            AW_REQUIRE(!rp.deltaf || (df && df->wd), "delta-f run without a weight array");
            AW_REQUIRE(!rp.deltaf || rp.df_wprec != 1 || df->wc,
                       "df_wprec = 1 needs the compensation array");
            AW_REQUIRE(!rp.deltaf || rp.df_wprec != 2 || df->wdd,
                       "df_wprec = 2 needs the FP64 weight array");
            const size_t n = (size_t)p.n;
            // the current's uz/gamma is taken before the walls damp u (bnd_x = 2)
            const bool own_vz = rp.rel || rp.bnd_x;
            // With tile_sort the particles are kept in tile order; without it,
            // where the current would bin the moves by tile anyway, the
            // particles are binned by tile before the push (a permutation,
            // the arrays keep their order) and the same one-kernel push +
            // current walks the bins.
            const int64_t ncl = nc();
            const bool binned = rp.tile_sort <= 0 && ncl > 0 &&
                                !((size_t)(3 * ncl) * 8 <= ctx->lds_per_block &&
                                  p.n >= 16 * ncl) &&
                                (double)p.n / (double)ncl >= 8.0;
            const bool fused = rp.tile_sort > 0 || binned;
            if (!fused) {
                x0.reserve(n);
                y0.reserve(n);
                if (own_vz) vz.reserve(n);
                if (rp.deltaf) wdep.reserve(n);
            }
            PushArgs a{};
            a.x = p.x;
            a.y = p.y;
            a.ux = p.ux;
            a.uy = p.uy;
            a.uz = p.uz;
            a.w = p.w;
            a.w_uniform = (float)p.w_uniform;
            a.wd = rp.deltaf ? df->wd : nullptr;
            a.wc = rp.deltaf ? df->wc : nullptr;
            a.wdd = rp.deltaf ? df->wdd : nullptr;
            a.x0 = fused ? nullptr : x0.get();
            a.y0 = fused ? nullptr : y0.get();
            a.vz = !fused && own_vz ? vz.get() : nullptr;
            a.wdep = !fused && rp.deltaf ? wdep.get() : nullptr;
            a.n = (int)p.n;
            a.tnow = tnow;
            // The current's fixed-point scale comes from bounds fixed before the
            // push: |uz/gamma| <= 2 max(1, c) (below c for the relativistic
            // push), |w| <= the run's largest weight, and |w wd| <= 8 |w|
            // (wd is clamped to [-8, 1)).  Particles past them set the guard,
            // which synchronize reports.
            const float wm = weight_max(p);
            a.ucap = (float)(2.0 * std::max(1.0, rp.c));
            a.wmax = wm;
            a.guard = jguard.get();
            EskExtra ex;
            ex.fixed = true;
            ex.fixed_per_w = rp.deltaf || p.w ? 1 : 0;
            const float wb = rp.deltaf ? 8.f * wm : wm;
            ex.fixed_bound[0] = ex.fixed_per_w ? wb : 0.f;
            ex.fixed_bound[1] = ex.fixed_per_w
                                    ? std::nextafter((float)((double)wb * a.ucap), INFINITY)
                                    : a.ucap;
            if (fused) {
                PushTiles t{};
                if (binned) {
                    // the bins are rebuilt every kRebinEvery steps: between
                    // them the particles drift, and the windows' aprons (the
                    // mesh and global atomics beyond them) take the drift
                    if (nstep >= next_sort || sorted_n != p.n || sorted_x != p.x) {
                        soff.reserve((size_t)ctx->ntiles + 1);
                        sperm.reserve(n);
                        bin_tiles(*ctx, p, sperm.get(), soff.get(), stream);
                        sorted_n = p.n;
                        sorted_x = p.x;
                        next_sort = nstep + kRebinEvery;
                    }
                    t.idx = sperm.get();
                } else if (nstep >= next_sort || sorted_n != p.n || sorted_x != p.x) {
                    tile_sort(p, df);
                    next_sort = nstep + rp.tile_sort;
                }
                t.off = soff.get();
                t.ntx = ctx->ntx;
                t.tx = ctx->tx;
                t.ty = ctx->ty;
                t.D = kPushApron;
                t.W = t.tx + 2 * t.D + 2;
                t.H = t.ty + 2 * t.D + 2;
                t.JW = t.tx + 2 * esk::kSortedApron;
                t.JH = t.ty + 2 * esk::kSortedApron;
                t.bpt = tile_bpt(*ctx, p.n);
                const size_t sh = (size_t)3 * t.JW * t.JH * sizeof(unsigned long long) +
                                  (size_t)6 * t.W * t.H * sizeof(float);
                AW_REQUIRE(sh <= ctx->lds_per_block, "the tile windows do not fit LDS");
                const esk::EskArgs e = esk::setup(*ctx, nullptr, nullptr, p, rp.qm, rp.dt, stream,
                                                  ex);
                const bool plain =
                    !(rp.pump || rp.rel || rp.b0_prof || rp.deltaf || rp.bnd_x);
                auto kern = g.ny == 1 ? (plain ? yee_push_esk_tiles_kernel<true, true>
                                               : yee_push_esk_tiles_kernel<true, false>)
                                      : (plain ? yee_push_esk_tiles_kernel<false, true>
                                               : yee_push_esk_tiles_kernel<false, false>);
                hipLaunchKernelGGL(kern,
                                   dim3(ctx->ntiles * t.bpt), dim3(kThreads), sh, stream, a, m, rp,
                                   t, e, ctx->jacc.get());
                AW_LAUNCH_CHECK();
                convert_and_smooth(e);
            } else {
                const int pb =
                    std::min(blocks_for(p.n), kPushBlocksPerCu * std::max(1, ctx->num_cu));
                if (!(rp.pump || rp.rel || rp.b0_prof || rp.deltaf || rp.bnd_x))
                    yee_push_kernel<true><<<pb, kThreads, 0, stream>>>(a, m, rp);
                else if (!(rp.pump || rp.rel))
                    yee_push_kernel<false, true><<<pb, kThreads, 0, stream>>>(a, m, rp);
                else
                    yee_push_kernel<false><<<pb, kThreads, 0, stream>>>(a, m, rp);
                AW_LAUNCH_CHECK();
                aw_particles dep = p;
                if (own_vz) dep.uz = vz.get();
                if (rp.deltaf) dep.w = wdep.get();
                ex.xw = p.x;  // the periodic wrap of the move
                ex.yw = p.y;
                ex.no_convert = true;
                deposit_esirkepov(*ctx, x0.get(), y0.get(), dep, rp.qm, rp.dt, AW_DEPOSIT_AUTO,
                                  m.jx, m.jy, m.jz, stream, ex);
                convert_and_smooth(esk::args(*ctx, x0.get(), y0.get(), dep, rp.qm, rp.dt, ex));
            }
        }
        // without particles the antenna and the cold fluid start from J = 0
        const bool ant = rp.ant_amp != 0.0, cold = rp.cold_nc > 0.0;
        const bool clear = p.n == 0 && (ant || cold);
        float* vcy = cold ? vc.get() : nullptr;
        float* vcz = cold ? vc.get() + nc() : nullptr;
        field_mid_kernel<<<nb, kThreads, 0, stream>>>(m, rp, tnow, h, clear ? 1 : 0, ant ? 1 : 0,
                                                      vcy, vcz);
        AW_LAUNCH_CHECK();
        ampere_kernel<<<nb, kThreads, 0, stream>>>(m);
        AW_LAUNCH_CHECK();
        if (rp.bnd_x) {
            damp_kernel<<<nb, kThreads, 0, stream>>>(m, mask.get(), mask.get() + g.nx, vcy, vcz);
            AW_LAUNCH_CHECK();
        }
        ++nstep;
    }

    // sum of the per-block pairs of mesh_stats_kernel, in order
    void mesh_stats(int what, double out[2]) {
        part.reserve(2 * (size_t)kRedBlocks);
        const int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (nc() + 1023) / 1024));
        mesh_stats_kernel<<<nb, kThreads, 0, stream>>>(mesh(), what, rres.get(), r0.get(),
                                                       part.get());
        AW_LAUNCH_CHECK();
        std::vector<double> h(2 * (size_t)nb);
        AW_HIP(hipMemcpyAsync(h.data(), part.get(), h.size() * sizeof(double),
                              hipMemcpyDeviceToHost, stream));
        AW_HIP(hipStreamSynchronize(stream));
        out[0] = out[1] = 0.0;
        for (int b = 0; b < nb; ++b) {
            out[0] += h[2 * b];
            out[1] += h[2 * b + 1];
        }
    }

    // MaxwellSimulation::residuals (include/pic/simulation_maxwell.hpp:125-160)
    void residuals(const aw_particles& p, const float* wd, double out[2]) {
        const int64_t n = nc();
        rho.reserve((size_t)n);
        rres.reserve((size_t)n);
        r0.reserve((size_t)n);
        if (p.n > 0) {
            aw_particles q = p;
            if (wd) {  // delta-f: the perturbation charge, weights w wd
                wdep.reserve((size_t)p.n);
                scale_weights(p, wd, wdep.get());
                q.w = wdep.get();
            }
            ctx->acc.zero(stream);
            deposit_rho_fixed(*ctx, q, rp.qm, AW_DEPOSIT_AUTO, nullptr, stream);
            fix_to_float_add(ctx->acc.get(), rho.get(), n, ctx->fix_log2, true, stream, true);
        } else {
            zero(rho.get(), n);
        }
        smooth(rho.get(), 1);
        gauss_kernel<<<blocks_for(n), kThreads, 0, stream>>>(mesh(), rho.get(), rres.get());
        AW_LAUNCH_CHECK();
        if (!have_ref) {
            AW_HIP(hipMemcpyAsync(r0.get(), rres.get(), (size_t)n * sizeof(float),
                                  hipMemcpyDeviceToDevice, stream));
            have_ref = true;
        }
        double s[2];
        mesh_stats(1, s);
        out[0] = std::sqrt(s[0] / (double)n);
        out[1] = std::sqrt(s[1] / (double)n);
    }

    void scale_weights(const aw_particles& p, const float* wd, float* out);
};

namespace aw {
namespace {

__global__ void scale_weights_kernel(const float* __restrict__ w, float wu,
                                     const float* __restrict__ wd, float* __restrict__ out,
                                     int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t < n) out[t] = (w ? w[t] : wu) * wd[t];
}

}  // namespace
}  // namespace aw

void aw_maxwell::scale_weights(const aw_particles& p, const float* wd, float* out) {
    scale_weights_kernel<<<blocks_for(p.n), kThreads, 0, stream>>>(p.w, (float)p.w_uniform, wd,
                                                                   out, p.n);
    AW_LAUNCH_CHECK();
}

extern "C" {

aw_status aw_maxwell_create(const aw_grid* g, const aw_run_params* rp, aw_maxwell** out) {
    return guard("aw_maxwell_create", [&] {
        check_grid(g);
        AW_REQUIRE(rp && out, "NULL argument");
        AW_REQUIRE(rp->dt > 0 && rp->c > 0, "dt and c must be positive");
        AW_REQUIRE(!rp->bnd_x || rp->bnd_nd > 0, "bnd_x needs bnd_nd > 0");
        AW_REQUIRE(rp->cold_nc <= 0.0 || g->ny == 1,
                   "cold_nc: the linearized cold fluid is ny = 1 only (Ey/Ez stagger "
                   "degeneracy)");
        *out = nullptr;
        auto m = std::make_unique<aw_maxwell>();
        m->g = *g;
        if (m->g.Lx <= 0) m->g.Lx = g->nx * g->dx;
        if (m->g.Ly <= 0) m->g.Ly = g->ny * g->dy;
        m->rp = *rp;
        m->ctx.reset(make_ctx(&m->g, 0, 0));
        AW_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
        const int64_t nc = m->nc();
        m->F.resize((size_t)(9 * nc));
        m->zero(m->F.get(), 9 * nc);
        m->jguard.resize(1);
        AW_HIP(hipMemsetAsync(m->jguard.get(), 0, sizeof(unsigned int), m->stream));
        if (rp->jfilter > 0 && nc > kSmoothLdsCells) m->jtmp.resize((size_t)(6 * nc));
        if (rp->bnd_x) m->build_masks();
        if (rp->cold_nc > 0.0) {
            m->vc.resize((size_t)(2 * nc));
            m->zero(m->vc.get(), 2 * nc);
        }
        AW_HIP(hipStreamSynchronize(m->stream));
        *out = m.release();
    });
}

aw_status aw_maxwell_destroy(aw_maxwell* m) {
    return guard("aw_maxwell_destroy", [&] { delete m; });
}

aw_status aw_maxwell_stream(aw_maxwell* m, aw_stream* out) {
    return guard("aw_maxwell_stream", [&] {
        AW_REQUIRE(m && out, "NULL argument");
        *out = (aw_stream)m->stream;
    });
}

aw_status aw_maxwell_synchronize(aw_maxwell* m) {
    return guard("aw_maxwell_synchronize", [&] {
        AW_REQUIRE(m, "NULL argument");
        m->check_guard();
    });
}

aw_status aw_maxwell_mesh(aw_maxwell* m, aw_yee_mesh* out) {
    return guard("aw_maxwell_mesh", [&] {
        AW_REQUIRE(m && out, "NULL argument");
        const Mesh v = m->mesh();
        out->ex = v.ex;
        out->ey = v.ey;
        out->ez = v.ez;
        out->bx = v.bx;
        out->by = v.by;
        out->bz = v.bz;
        out->jx = v.jx;
        out->jy = v.jy;
        out->jz = v.jz;
        out->vcy = m->rp.cold_nc > 0.0 ? m->vc.get() : nullptr;
        out->vcz = m->rp.cold_nc > 0.0 ? m->vc.get() + m->nc() : nullptr;
    });
}

aw_status aw_maxwell_step(aw_maxwell* m, aw_particles* p, const aw_deltaf_state* df) {
    return guard("aw_maxwell_step", [&] {
        AW_REQUIRE(m && p, "NULL argument");
        check_particles(p);
        m->step(*p, df, (double)m->nstep * m->rp.dt);
    });
}

aw_status aw_maxwell_step_at(aw_maxwell* m, aw_particles* p, const aw_deltaf_state* df,
                             double tnow) {
    return guard("aw_maxwell_step_at", [&] {
        AW_REQUIRE(m && p, "NULL argument");
        check_particles(p);
        m->step(*p, df, tnow);
    });
}

aw_status aw_maxwell_field_energy(aw_maxwell* m, double* we, double* wb) {
    return guard("aw_maxwell_field_energy", [&] {
        AW_REQUIRE(m && we && wb, "NULL argument");
        double s[2];
        m->mesh_stats(0, s);
        *we = s[0];
        *wb = s[1];
    });
}

aw_status aw_maxwell_residuals(aw_maxwell* m, const aw_particles* p, const float* wd,
                               double* divb_rms, double* gauss_drift_rms) {
    return guard("aw_maxwell_residuals", [&] {
        AW_REQUIRE(m && p && divb_rms && gauss_drift_rms, "NULL argument");
        check_particles(p);
        double out[2];
        m->residuals(*p, wd, out);
        *divb_rms = out[0];
        *gauss_drift_rms = out[1];
    });
}

aw_status aw_maxwell_diag_fields(aw_maxwell* m, double t, aw_fields* out) {
    return guard("aw_maxwell_diag_fields", [&] {
        AW_REQUIRE(m && out, "NULL argument");
        const int64_t nc = m->nc();
        m->dfields.reserve((size_t)(7 * nc));
        const int pump = m->rp.pump && t < m->rp.pump_toff;
        diag_fields_kernel<<<blocks_for(nc), kThreads, 0, m->stream>>>(m->mesh(), m->rp, t, pump,
                                                                     m->dfields.get());
        AW_LAUNCH_CHECK();
        float* f = m->dfields.get();
        *out = aw_fields{};
        out->Ex = f;
        out->Ey = f + nc;
        out->Ez = f + 2 * nc;
        out->Bx = f + 3 * nc;
        out->By = f + 4 * nc;
        out->Bz = f + 5 * nc;
        out->ELx = f + 6 * nc;
    });
}

aw_status aw_wd_stats(const float* wd, int64_t n, double* sum, double* rms, double* max_abs,
                      aw_stream s) {
    return guard("aw_wd_stats", [&] {
        AW_REQUIRE(sum && rms && max_abs, "NULL argument");
        *sum = *rms = *max_abs = 0.0;
        if (n <= 0) return;
        AW_REQUIRE(wd, "wd is NULL");
        const int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (n + 4095) / 4096));
        DevBuf<double> part((size_t)(3 * nb));
        wd_stats_kernel<<<nb, kThreads, 0, S(s)>>>(wd, n, part.get());
        AW_LAUNCH_CHECK();
        std::vector<double> h(3 * (size_t)nb);
        AW_HIP(hipMemcpyAsync(h.data(), part.get(), h.size() * sizeof(double),
                              hipMemcpyDeviceToHost, S(s)));
        AW_HIP(hipStreamSynchronize(S(s)));
        double s1 = 0.0, s2 = 0.0, mx = 0.0;
        for (int b = 0; b < nb; ++b) {
            s1 += h[3 * b];
            s2 += h[3 * b + 1];
            mx = std::max(mx, h[3 * b + 2]);
        }
        *sum = s1;
        *rms = std::sqrt(s2 / (double)n);
        *max_abs = mx;
    });
}

aw_status aw_mirror_count(const aw_grid* g, const aw_run_params* rp, const aw_species* q,
                          int64_t* n) {
    return guard("aw_mirror_count", [&] {
        check_grid(g);
        AW_REQUIRE(rp && q && n, "NULL argument");
        check_mirror(*g, *rp, *q);
        *n = mirror_offsets(*g, *rp, *q).back();
    });
}

aw_status aw_load_mirror(const aw_grid* g, const aw_run_params* rp, const aw_species* q,
                         aw_particles* p, aw_stream s) {
    return guard("aw_load_mirror", [&] {
        check_grid(g);
        AW_REQUIRE(rp && q && p, "NULL argument");
        check_mirror(*g, *rp, *q);
        check_particles(p);
        const std::vector<int32_t> off = mirror_offsets(*g, *rp, *q);
        AW_REQUIRE(p->n == off.back(), "particle count must equal aw_mirror_count");
        if (p->n == 0) return;
        DevBuf<int32_t> doff(off.size());
        AW_HIP(hipMemcpyAsync(doff.get(), off.data(), off.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, S(s)));
        MirrorLoad L{};
        L.off = doff.get();
        L.nx = g->nx;
        L.dx = (float)g->dx;
        L.uth_par = q->uth[0];
        L.uth_perp_eq = q->uth[1];
        L.lc_rho = q->lc_rho;
        L.lc_kappa = q->lc_kappa;
        L.weight = (float)(q->density * g->dx * g->dy / q->ppc);
        L.seed = rp->rng_seed;
        L.dist = q->dist;
        L.n = p->n;
        mirror_load_kernel<<<blocks_for(p->n), kThreads, 0, S(s)>>>(L, *rp, p->x, p->y, p->ux,
                                                                  p->uy, p->uz, p->w);
        AW_LAUNCH_CHECK();
        AW_HIP(hipStreamSynchronize(S(s)));  // doff lives in this call
    });
}

}  // extern "C"
