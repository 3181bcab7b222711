// push.hip -- electrostatic Boris push kernels.
//
//   push_es_kernel        Pusher<Cfg>::boris (include/pic/pusher.hpp:112-129)
//   half_back_kernel      Pusher<Cfg>::half_step_back (:133-145)
//   push_deposit_kernel   the MI355X hot path: one pass per particle that
//                         gathers E^n from an LDS field tile, pushes, moves,
//                         wraps and deposits rho(x^{n+1}) into an LDS tile
//                         (fixed-point, flushed once per workgroup).  It does
//                         the work of boris_push_tiled_kernel (:154-202) +
//                         particle_migrate_kernel (particles.hpp:293-315) +
//                         the next step's deposit_charge_tiled_kernel
//                         (depositor.hpp:112-153) for 40 B of particle traffic
//                         (32 B when B0 is along z: uz is then invariant).
#include "aw_internal.hpp"

#include <cstring>

namespace aw {

namespace {

constexpr int kThreads = 256;

inline int blocks_for(int64_t n) { return (int)((n + kThreads - 1) / kThreads); }

struct PushArgs {
    float *x, *y, *ux, *uy, *uz;
    int n, nx, ny;
    const float* Ex;
    const float* Ey;
    BorisConsts bc;
    float cx, cy;  // dt/dx, dt/dy in cell units
    int has_b0, move;
};

__device__ __forceinline__ void gather_global(const float* __restrict__ Ex,
                                              const float* __restrict__ Ey, const dev::CicF& s,
                                              int nx, int ny, float& ex, float& ey) {
    int i0, i1, j0, j1;
    dev::wrap_pair(s.i0, nx, i0, i1);
    dev::wrap_pair(s.j0, ny, j0, j1);
    const int c0 = j0 * nx + i0, c1 = j0 * nx + i1, c2 = j1 * nx + i0, c3 = j1 * nx + i1;
    ex = dev::gather4(s, Ex[c0], Ex[c1], Ex[c2], Ex[c3]);
    ey = dev::gather4(s, Ey[c0], Ey[c1], Ey[c2], Ey[c3]);
}

__global__ void __launch_bounds__(kThreads) push_es_kernel(PushArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float x = a.x[i], y = a.y[i];
    float ux = a.ux[i], uy = a.uy[i], uz = a.uz[i];
    const dev::CicF s = dev::cic_f(x, y);
    float ex, ey;
    gather_global(a.Ex, a.Ey, s, a.nx, a.ny, ex, ey);
    dev::boris_es(a.bc, a.has_b0 != 0, ex, ey, ux, uy, uz);
    a.ux[i] = ux;
    a.uy[i] = uy;
    a.uz[i] = uz;
    if (a.move) {
        a.x[i] = fmaf(ux, a.cx, x);
        a.y[i] = fmaf(uy, a.cy, y);
    }
}

struct FusedArgs {
    float *x, *y, *ux, *uy, *uz;
    const float* w;  // nullptr: uniform weight
    const int32_t* off;
    int ntx, tx, ty, apron, W, H, bpt;
    int nx, ny;
    const float* Ex;
    const float* Ey;
    unsigned long long* acc;
    BorisConsts bc;
    float cx, cy;
    int has_b0;
    double qm, area, coef_u, scale;
    unsigned long long* strays;
    double* ke_part;  // per-workgroup sum of w*u^2 (nullptr: skip)
    float *sx, *sy;   // non-null: receive the positions before the move
};

__device__ __forceinline__ int local_index(int i, int org, int n) {
    int l = i - org;
    if (l < 0) l += n;
    else if (l >= n) l -= n;
    return l;
}

struct TileFrame {
    unsigned long long* s_rho;
    const float2* s_e;  // (Ex, Ey) interleaved
    int W, H, orgx, orgy;
};

// Everything after the loads for one particle: gather E^n, Boris, move,
// wrap, store, deposit rho(x^{n+1}).  UZ = false: B0 is along z (or absent),
// so uz is a constant of the motion; it is then neither rotated nor stored,
// and read only for the kinetic energy (KE).
template <bool PER_W, bool KE, bool UZ, bool SAVE = false>
__device__ __forceinline__ void push_one(const FusedArgs& a, const TileFrame& f, int k, float x,
                                         float y, float ux, float uy, float uz, float wk,
                                         unsigned int& nstray, double& ke) {
    const dev::CicF s = dev::cic_f(x, y);
    const int li = local_index(s.i0, f.orgx, a.nx), lj = local_index(s.j0, f.orgy, a.ny);
    float ex, ey;
    if (dev::in_tile(li, f.W - 2) && dev::in_tile(lj, f.H - 2)) {
        const int b = lj * f.W + li;
        const float2 e0 = f.s_e[b], e1 = f.s_e[b + 1], e2 = f.s_e[b + f.W],
                     e3 = f.s_e[b + f.W + 1];
        ex = dev::gather4(s, e0.x, e1.x, e2.x, e3.x);
        ey = dev::gather4(s, e0.y, e1.y, e2.y, e3.y);
    } else {
        gather_global(a.Ex, a.Ey, s, a.nx, a.ny, ex, ey);
        ++nstray;
    }

    if constexpr (UZ) dev::boris_es(a.bc, a.has_b0 != 0, ex, ey, ux, uy, uz);
    else dev::boris_es_bz(a.bc, a.has_b0 != 0, ex, ey, ux, uy);
    const float xn = dev::wrap_coord(fmaf(ux, a.cx, x), (float)a.nx);
    const float yn = dev::wrap_coord(fmaf(uy, a.cy, y), (float)a.ny);
    if constexpr (SAVE) {
        a.sx[k] = x;
        a.sy[k] = y;
    }
    a.x[k] = xn;
    a.y[k] = yn;
    a.ux[k] = ux;
    a.uy[k] = uy;
    if constexpr (UZ) a.uz[k] = uz;

    double coef;
    if constexpr (PER_W) {
        coef = a.qm * (double)wk / a.area;
        if constexpr (KE) ke += (double)wk * ((double)ux * ux + (double)uy * uy + (double)uz * uz);
    } else {
        coef = a.coef_u;
        if constexpr (KE) ke += (double)ux * ux + (double)uy * uy + (double)uz * uz;
    }

    const dev::CicD d = dev::cic_d(xn, yn);
    const int mi = local_index(d.i0, f.orgx, a.nx), mj = local_index(d.j0, f.orgy, a.ny);
    const unsigned long long q0 = dev::fixq(coef, d.w0, a.scale);
    const unsigned long long q1 = dev::fixq(coef, d.w1, a.scale);
    const unsigned long long q2 = dev::fixq(coef, d.w2, a.scale);
    const unsigned long long q3 = dev::fixq(coef, d.w3, a.scale);
    if (dev::in_tile(mi, f.W - 2) && dev::in_tile(mj, f.H - 2)) {
        const int b = mj * f.W + mi;
        atomicAdd(&f.s_rho[b], q0);
        atomicAdd(&f.s_rho[b + 1], q1);
        atomicAdd(&f.s_rho[b + f.W], q2);
        atomicAdd(&f.s_rho[b + f.W + 1], q3);
    } else {
        int i0, i1, j0, j1;
        dev::wrap_pair(d.i0, a.nx, i0, i1);
        dev::wrap_pair(d.j0, a.ny, j0, j1);
        atomicAdd(&a.acc[j0 * a.nx + i0], q0);
        atomicAdd(&a.acc[j0 * a.nx + i1], q1);
        atomicAdd(&a.acc[j1 * a.nx + i0], q2);
        atomicAdd(&a.acc[j1 * a.nx + i1], q3);
        ++nstray;
    }
}

// One workgroup's share of the fused step: blk is the workgroup index within
// the run (tile blk / bpt, part blk % bpt).
template <bool PER_W, bool KE, bool UZ, bool SAVE = false>
__device__ __forceinline__ void push_deposit_block(const FusedArgs& a, int blk) {
    constexpr bool kLoadUz = UZ || KE;
    extern __shared__ unsigned long long smem[];
    const int W = a.W, H = a.H, R = W * H;
    TileFrame f;
    f.s_rho = smem;
    float2* s_e = reinterpret_cast<float2*>(smem + R);
    f.s_e = s_e;
    f.W = W;
    f.H = H;

    const int tile = blk / a.bpt, part = blk % a.bpt;
    f.orgx = (tile % a.ntx) * a.tx - a.apron;
    f.orgy = (tile / a.ntx) * a.ty - a.apron;
    int r0, r1;
    dev::split_range(a.off[tile], a.off[tile + 1], part, a.bpt, r0, r1);

    // a thread's particles k, k + kThreads, ...: loaded two at a time, every
    // load of a pair issued before either particle is consumed
    struct Part {
        float x, y, ux, uy, uz, w;
    };
    auto fetch = [&](int i) {
        Part p{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (i < r1) {
            p.x = a.x[i];
            p.y = a.y[i];
            p.ux = a.ux[i];
            p.uy = a.uy[i];
            if constexpr (kLoadUz) p.uz = a.uz[i];
            if constexpr (PER_W) p.w = a.w[i];
        }
        return p;
    };
    // the first pair is requested before the field tile, so that the
    // workgroup's particle reads overlap its tile load
    int k = r0 + (int)threadIdx.x;
    Part c0 = fetch(k), c1 = fetch(k + kThreads);

    for (int c = threadIdx.x; c < R; c += kThreads) {
        const int gx = dev::wrapfar(f.orgx + c % W, a.nx);
        const int gy = dev::wrapfar(f.orgy + c / W, a.ny);
        s_e[c] = make_float2(a.Ex[gy * a.nx + gx], a.Ey[gy * a.nx + gx]);
        f.s_rho[c] = 0ull;
    }
    __syncthreads();

    unsigned int nstray = 0;
    double ke = 0.0;
    for (; k + kThreads < r1; k += 2 * kThreads) {
        push_one<PER_W, KE, UZ, SAVE>(a, f, k, c0.x, c0.y, c0.ux, c0.uy, c0.uz, c0.w, nstray,
                                      ke);
        push_one<PER_W, KE, UZ, SAVE>(a, f, k + kThreads, c1.x, c1.y, c1.ux, c1.uy, c1.uz, c1.w,
                                      nstray, ke);
        c0 = fetch(k + 2 * kThreads);
        c1 = fetch(k + 3 * kThreads);
    }
    if (k < r1)
        push_one<PER_W, KE, UZ, SAVE>(a, f, k, c0.x, c0.y, c0.ux, c0.uy, c0.uz, c0.w, nstray,
                                      ke);

    // per-workgroup stray count and kinetic-energy partial
    __shared__ unsigned int s_stray;
    __shared__ double s_ke[kThreads / 64];
    if (threadIdx.x == 0) s_stray = 0;
    __syncthreads();
    if (nstray) atomicAdd(&s_stray, nstray);
    if constexpr (KE) {
        for (int off = 32; off > 0; off >>= 1) ke += __shfl_down(ke, off, 64);
        if ((threadIdx.x & 63) == 0) s_ke[threadIdx.x >> 6] = ke;
    }
    __syncthreads();

    // flush the LDS charge tile (every cell once)
    for (int c = threadIdx.x; c < R; c += kThreads) {
        const unsigned long long v = f.s_rho[c];
        if (v) {
            const int gx = dev::wrapfar(f.orgx + c % W, a.nx);
            const int gy = dev::wrapfar(f.orgy + c / W, a.ny);
            atomicAdd(&a.acc[gy * a.nx + gx], v);
        }
    }
    if (threadIdx.x == 0) {
        if (s_stray && a.strays) atomicAdd(a.strays, (unsigned long long)s_stray);
        if constexpr (KE) {
            double t = 0.0;
            for (int w = 0; w < kThreads / 64; ++w) t += s_ke[w];
            a.ke_part[blk] = t;
        }
    }
}

template <bool PER_W, bool KE, bool UZ, bool SAVE = false>
__global__ void __launch_bounds__(kThreads) push_deposit_kernel(FusedArgs a) {
    push_deposit_block<PER_W, KE, UZ, SAVE>(a, dev::xcd_block(blockIdx.x, gridDim.x));
}

// Ensembles: one launch over every run.  runs[r] holds run r's arguments
// (pointers into its slices of the concatenated particle / grid arrays) and
// wg_begin its first workgroup; wg_run[blockIdx.x] names a workgroup's run
// (one load where a search over the runs took several dependent ones).
struct RunSlot {
    FusedArgs a;
    int wg_begin;
};

template <bool KE, bool UZ>
__global__ void __launch_bounds__(kThreads)
push_deposit_multi_kernel(const RunSlot* __restrict__ runs, const int32_t* __restrict__ wg_run) {
    // dispatch order kept: workgroups of different runs differ widely in
    // cost, and the round-robin over the XCDs balances them
    const int w = blockIdx.x;
    const RunSlot& r = runs[wg_run[w]];
    const FusedArgs a = r.a;
    push_deposit_block<false, KE, UZ>(a, w - r.wg_begin);
}

}  // namespace

int tile_bpt(const aw_ctx& c, int64_t n);
bool uz_evolves(const aw_run_params& rp, bool has_b0);

void push_es(aw_ctx& c, const aw_particles& p, const float* Ex, const float* Ey,
             const aw_run_params& rp, bool has_b0, bool half_back, hipStream_t s) {
    if (p.n == 0) return;
    AW_REQUIRE(p.n < (int64_t)INT32_MAX, "particle count must fit int32");
    PushArgs a{};
    a.x = p.x;
    a.y = p.y;
    a.ux = p.ux;
    a.uy = p.uy;
    a.uz = p.uz;
    a.n = (int)p.n;
    a.nx = c.g.nx;
    a.ny = c.g.ny;
    a.Ex = Ex;
    a.Ey = Ey;
    a.bc = boris_make(rp.qm, half_back ? -0.5 * rp.dt : rp.dt, rp.B0, has_b0);
    a.cx = (float)(rp.dt / c.g.dx);
    a.cy = (float)(rp.dt / c.g.dy);
    a.has_b0 = has_b0 ? 1 : 0;
    a.move = half_back ? 0 : 1;
    push_es_kernel<<<blocks_for(p.n), kThreads, 0, s>>>(a);
    AW_LAUNCH_CHECK();
}

static FusedArgs make_fused(aw_ctx& c, const aw_particles& p, const int32_t* off,
                            const float* Ex, const float* Ey, const aw_run_params& rp,
                            bool has_b0, unsigned long long* acc, unsigned long long* strays,
                            double* ke_part, int bpt) {
    FusedArgs a{};
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
    a.bpt = bpt > 0 ? bpt : tile_bpt(c, p.n);
    a.nx = c.g.nx;
    a.ny = c.g.ny;
    a.Ex = Ex;
    a.Ey = Ey;
    a.acc = acc;
    a.bc = boris_make(rp.qm, rp.dt, rp.B0, has_b0);
    a.cx = (float)(rp.dt / c.g.dx);
    a.cy = (float)(rp.dt / c.g.dy);
    a.has_b0 = has_b0 ? 1 : 0;
    a.qm = rp.qm;
    a.area = c.g.dx * c.g.dy;
    a.coef_u = rp.qm * p.w_uniform / a.area;
    a.scale = std::ldexp(1.0, c.fix_log2);
    a.strays = strays;
    a.ke_part = ke_part;
    return a;
}

// Returns the number of workgroups launched (size of ke_part).  save_xy
// (2n floats, may be null) receives [x | y] before the move.
int push_deposit_es_save(aw_ctx& c, const aw_particles& p, const int32_t* off, const float* Ex,
                         const float* Ey, const aw_run_params& rp, bool has_b0,
                         unsigned long long* strays, double* ke_part, int bpt, float* save_xy,
                         hipStream_t s) {
    if (p.n == 0) return 0;
    AW_REQUIRE(p.n < (int64_t)INT32_MAX, "particle count must fit int32");
    FusedArgs a = make_fused(c, p, off, Ex, Ey, rp, has_b0, c.acc.get(), strays, ke_part, bpt);
    const size_t sh = (size_t)a.W * a.H * (8 + 4 + 4);
    const int nblk = c.ntiles * a.bpt;
    const bool pw = p.w != nullptr, ke = ke_part != nullptr;
    const bool uz = uz_evolves(rp, has_b0);
    auto launch = [&](auto kernel) {
        kernel<<<nblk, kThreads, sh, s>>>(a);
        AW_LAUNCH_CHECK();
    };
    if (save_xy) {  // the adjoint's recompute: uniform weights, B0-free, no diagnostics
        AW_REQUIRE(!uz && !pw && !ke, "saving the positions needs the uniform-weight ES push");
        a.sx = save_xy;
        a.sy = save_xy + p.n;
        launch(push_deposit_kernel<false, false, false, true>);
        return nblk;
    }
    if (uz) {
        if (pw && ke) launch(push_deposit_kernel<true, true, true>);
        else if (pw) launch(push_deposit_kernel<true, false, true>);
        else if (ke) launch(push_deposit_kernel<false, true, true>);
        else launch(push_deposit_kernel<false, false, true>);
    } else {
        if (pw && ke) launch(push_deposit_kernel<true, true, false>);
        else if (pw) launch(push_deposit_kernel<true, false, false>);
        else if (ke) launch(push_deposit_kernel<false, true, false>);
        else launch(push_deposit_kernel<false, false, false>);
    }
    return nblk;
}

int push_deposit_es(aw_ctx& c, const aw_particles& p, const int32_t* off, const float* Ex,
                    const float* Ey, const aw_run_params& rp, bool has_b0,
                    unsigned long long* strays, double* ke_part, int bpt, hipStream_t s) {
    return push_deposit_es_save(c, p, off, Ex, Ey, rp, has_b0, strays, ke_part, bpt, nullptr, s);
}

// uz changes in the ES push only through a B0 component in the plane
bool uz_evolves(const aw_run_params& rp, bool has_b0) {
    return has_b0 && (rp.B0[0] != 0.0f || rp.B0[1] != 0.0f);
}

// ---- ensembles ----------------------------------------------------------------------

size_t run_slot_bytes() { return sizeof(RunSlot); }

// Writes run r's slot (host memory) for push_deposit_multi; returns its
// workgroup count.  acc, strays and ke_part point at the run's slices.
int fill_run_slot(void* slot, aw_ctx& c, const aw_particles& p, const int32_t* off,
                  const float* Ex, const float* Ey, const aw_run_params& rp, bool has_b0,
                  unsigned long long* acc, unsigned long long* strays, double* ke_part, int bpt,
                  int wg_begin) {
    AW_REQUIRE(p.w == nullptr, "ensemble runs carry one weight per run");
    RunSlot r{};
    r.a = make_fused(c, p, off, Ex, Ey, rp, has_b0, acc, strays, ke_part, bpt);
    r.wg_begin = wg_begin;
    std::memcpy(slot, &r, sizeof(r));
    return c.ntiles * r.a.bpt;
}

size_t push_lds_bytes(const aw_ctx& c) {
    return (size_t)(c.tx + 2 * c.apron + 1) * (c.ty + 2 * c.apron + 1) * (8 + 4 + 4);
}

void push_deposit_multi(const void* slots, const int32_t* wg_run, int total_wg, size_t lds,
                        bool ke, bool uz, hipStream_t s) {
    if (total_wg == 0) return;
    const RunSlot* r = static_cast<const RunSlot*>(slots);
    if (uz) {
        if (ke) push_deposit_multi_kernel<true, true><<<total_wg, kThreads, lds, s>>>(r, wg_run);
        else push_deposit_multi_kernel<false, true><<<total_wg, kThreads, lds, s>>>(r, wg_run);
    } else {
        if (ke) push_deposit_multi_kernel<true, false><<<total_wg, kThreads, lds, s>>>(r, wg_run);
        else push_deposit_multi_kernel<false, false><<<total_wg, kThreads, lds, s>>>(r, wg_run);
    }
    AW_LAUNCH_CHECK();
}

}  // namespace aw
