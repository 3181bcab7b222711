// deposit.hip -- CIC charge deposit in three policies, all accumulating
// 2^-S fixed-point integers so every policy returns identical bits.
//
// Reference: Depositor<Cfg> (include/pic/depositor.hpp:34-58 global atomics,
// :75-97 whole-grid shared, :112-153 tiled+binned, :440-496 dispatch).
#include "aw_internal.hpp"

namespace aw {

namespace {

constexpr int kThreads = 256;

inline int blocks_for(int64_t n) { return (int)((n + kThreads - 1) / kThreads); }

struct DepArgs {
    const float* x;
    const float* y;
    const float* w;
    int n, nx, ny;
    double qm, area, coef_u, scale;
};

__device__ __forceinline__ double coef_of(const DepArgs& a, int i) {
    return a.w ? a.qm * (double)a.w[i] / a.area : a.coef_u;
}

__device__ __forceinline__ void deposit_global4(unsigned long long* acc, const DepArgs& a,
                                                const dev::CicD& s, double coef) {
    int i0, i1, j0, j1;
    dev::wrap_pair(s.i0, a.nx, i0, i1);
    dev::wrap_pair(s.j0, a.ny, j0, j1);
    atomicAdd(&acc[j0 * a.nx + i0], dev::fixq(coef, s.w0, a.scale));
    atomicAdd(&acc[j0 * a.nx + i1], dev::fixq(coef, s.w1, a.scale));
    atomicAdd(&acc[j1 * a.nx + i0], dev::fixq(coef, s.w2, a.scale));
    atomicAdd(&acc[j1 * a.nx + i1], dev::fixq(coef, s.w3, a.scale));
}

// AtomicGlobalDeposit: one particle per thread, 4 global integer atomics.
__global__ void __launch_bounds__(kThreads)
deposit_global_kernel(DepArgs a, unsigned long long* __restrict__ acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    deposit_global4(acc, a, dev::cic_d(a.x[i], a.y[i]), coef_of(a, i));
}

// SharedTileDeposit: the whole grid privatised in LDS per resident block.
__global__ void __launch_bounds__(kThreads)
deposit_ldsgrid_kernel(DepArgs a, unsigned long long* __restrict__ acc) {
    extern __shared__ unsigned long long s_rho[];
    const int nc = a.nx * a.ny;
    for (int c = threadIdx.x; c < nc; c += blockDim.x) s_rho[c] = 0ull;
    __syncthreads();
    const int stride = blockDim.x * gridDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        const dev::CicD s = dev::cic_d(a.x[i], a.y[i]);
        const double coef = coef_of(a, i);
        int i0, i1, j0, j1;
        dev::wrap_pair(s.i0, a.nx, i0, i1);
        dev::wrap_pair(s.j0, a.ny, j0, j1);
        atomicAdd(&s_rho[j0 * a.nx + i0], dev::fixq(coef, s.w0, a.scale));
        atomicAdd(&s_rho[j0 * a.nx + i1], dev::fixq(coef, s.w1, a.scale));
        atomicAdd(&s_rho[j1 * a.nx + i0], dev::fixq(coef, s.w2, a.scale));
        atomicAdd(&s_rho[j1 * a.nx + i1], dev::fixq(coef, s.w3, a.scale));
    }
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
        const unsigned long long v = s_rho[c];
        if (v) atomicAdd(&acc[c], v);
    }
}

struct TileGeom {
    const int32_t* off;  // ntiles+1
    const int32_t* idx;  // permutation (binned) or nullptr (physically sorted)
    int ntx, tx, ty, apron, W, H, bpt;
};

// TiledBinnedDeposit: workgroup per (tile, part); tile + apron in LDS;
// particles outside the apron (stale sort) go straight to global atomics.
__global__ void __launch_bounds__(kThreads)
deposit_tiles_kernel(DepArgs a, TileGeom t, unsigned long long* __restrict__ acc) {
    extern __shared__ unsigned long long s_rho[];
    const int R = t.W * t.H;
    const int tile = blockIdx.x / t.bpt, part = blockIdx.x % t.bpt;
    const int orgx = (tile % t.ntx) * t.tx - t.apron, orgy = (tile / t.ntx) * t.ty - t.apron;
    for (int c = threadIdx.x; c < R; c += blockDim.x) s_rho[c] = 0ull;
    __syncthreads();
    const int beg = t.off[tile], end = t.off[tile + 1];
    const int chunk = (end - beg + t.bpt - 1) / t.bpt;
    const int r0 = beg + part * chunk, r1 = min(end, r0 + chunk);
    for (int k = r0 + (int)threadIdx.x; k < r1; k += blockDim.x) {
        const int i = t.idx ? t.idx[k] : k;
        const dev::CicD s = dev::cic_d(a.x[i], a.y[i]);
        const double coef = coef_of(a, i);
        int li = s.i0 - orgx, lj = s.j0 - orgy;
        if (li < 0) li += a.nx; else if (li >= a.nx) li -= a.nx;
        if (lj < 0) lj += a.ny; else if (lj >= a.ny) lj -= a.ny;
        if (dev::in_tile(li, t.W - 2) && dev::in_tile(lj, t.H - 2)) {
            const int b = lj * t.W + li;
            atomicAdd(&s_rho[b], dev::fixq(coef, s.w0, a.scale));
            atomicAdd(&s_rho[b + 1], dev::fixq(coef, s.w1, a.scale));
            atomicAdd(&s_rho[b + t.W], dev::fixq(coef, s.w2, a.scale));
            atomicAdd(&s_rho[b + t.W + 1], dev::fixq(coef, s.w3, a.scale));
        } else {
            deposit_global4(acc, a, s, coef);
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < R; c += blockDim.x) {
        const unsigned long long v = s_rho[c];
        if (v) {
            const int gx = dev::wrapfar(orgx + c % t.W, a.nx);
            const int gy = dev::wrapfar(orgy + c / t.W, a.ny);
            atomicAdd(&acc[gy * a.nx + gx], v);
        }
    }
}

// out = (assign ? 0 : out) + acc * 2^-S; optionally clears acc for the next
// accumulation and one 64-bit counter (saves separate memset launches).
// out2 (the R2C input) receives the value less the integer baseline
// (base_arr[c] when given, else base_val): the caller passes the mean charge
// density, which the solve drops anyway (k = 0), so the FP32 transform
// rounds relative to the fluctuation instead of the mean.
__global__ void fix_to_float_kernel(unsigned long long* __restrict__ acc, float* __restrict__ out,
                                    float* __restrict__ out2, int64_t n, double inv, int zero_acc,
                                    int assign, unsigned long long* __restrict__ counter,
                                    int ncounter, const long long* __restrict__ base_arr,
                                    long long base_val) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (counter && c < ncounter) counter[c] = 0ull;
    if (c >= n) return;
    const long long v = (long long)acc[c];
    const float add = (float)((double)v * inv);
    const float r = assign ? add : out[c] + add;
    out[c] = r;
    if (out2) {
        const long long b = base_arr ? base_arr[c] : base_val;
        out2[c] = b ? (float)((double)(v - b) * inv) : r;
    }
    if (zero_acc) acc[c] = 0ull;
}

}  // namespace

void fix_to_float_add(const unsigned long long* acc, float* out, int64_t n, int scale_log2,
                      bool zero_acc, hipStream_t s, bool assign, unsigned long long* counter,
                      float* out2, int ncounter, const long long* base_arr, long long base_val) {
    if (n == 0) return;
    AW_REQUIRE(ncounter >= 0, "ncounter must be non-negative");
    AW_REQUIRE(assign || (!base_arr && !base_val), "a baseline needs assign");
    fix_to_float_kernel<<<blocks_for(std::max<int64_t>(n, ncounter)), kThreads, 0, s>>>(
        const_cast<unsigned long long*>(acc), out, out2, n, std::ldexp(1.0, -scale_log2),
        zero_acc ? 1 : 0, assign ? 1 : 0, counter, ncounter, base_arr, base_val);
    AW_LAUNCH_CHECK();
}

long long rho_baseline(double qm, double density, int scale_log2) {
    const double v = std::ldexp(qm * density, scale_log2);
    // the accumulators' range is far below 2^62; out-of-range input means no baseline
    return std::fabs(v) < 4e18 ? (long long)std::llrint(v) : 0ll;
}

// rho of p into the caller's fixed-point accumulator acc (nx*ny cells, scale
// 2^scale_log2) by the global-atomic or whole-grid LDS policy, with no aw_ctx
// and no allocation; AUTO picks LDS_GRID when the grid fits LDS.
void deposit_rho_acc(const aw_grid& g, const aw_particles& p, double qm, int policy,
                     int scale_log2, unsigned long long* acc, int num_cu, size_t lds_per_block,
                     hipStream_t s) {
    if (p.n == 0) return;
    DepArgs a{};
    a.x = p.x;
    a.y = p.y;
    a.w = p.w;
    a.n = (int)p.n;
    a.nx = g.nx;
    a.ny = g.ny;
    a.qm = qm;
    a.area = g.dx * g.dy;
    a.coef_u = qm * p.w_uniform / a.area;
    a.scale = std::ldexp(1.0, scale_log2);
    const size_t sh = (size_t)g.nx * g.ny * 8;
    if (policy == AW_DEPOSIT_AUTO) policy = sh <= lds_per_block ? AW_DEPOSIT_LDS_GRID
                                                                : AW_DEPOSIT_GLOBAL;
    if (policy == AW_DEPOSIT_GLOBAL) {
        deposit_global_kernel<<<blocks_for(p.n), kThreads, 0, s>>>(a, acc);
        AW_LAUNCH_CHECK();
    } else if (policy == AW_DEPOSIT_LDS_GRID) {
        AW_REQUIRE(sh <= lds_per_block, "grid does not fit LDS for AW_DEPOSIT_LDS_GRID");
        int per_cu = 0;
        AW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(deposit_ldsgrid_kernel), kThreads, sh));
        int nb = std::max(1, per_cu) * std::max(1, num_cu);
        nb = std::min(nb, blocks_for(p.n));
        deposit_ldsgrid_kernel<<<nb, kThreads, sh, s>>>(a, acc);
        AW_LAUNCH_CHECK();
    } else {
        fail(AW_ERR_UNSUPPORTED, "policy " + std::to_string(policy) +
                                     " needs binning scratch: use AW_DEPOSIT_GLOBAL, "
                                     "AW_DEPOSIT_LDS_GRID or AW_DEPOSIT_AUTO",
             __FILE__, __LINE__);
    }
}

int tile_bpt(const aw_ctx& c, int64_t n) {
    // about 2048 particles (8 per lane) per workgroup: enough to amortise the
    // tile load and flush, few enough that the grid holds many workgroups per
    // CU and the tail of the launch stays short
    const int64_t per_tile = n / std::max(1, c.ntiles);
    int bpt = (int)std::clamp<int64_t>((per_tile + 2047) / 2048, 1, 16);
    // small grids: keep >= 4 workgroups per CU while each keeps >= 1 particle
    // per lane
    const int64_t want = 4ll * std::max(1, c.num_cu);
    while (bpt < 64 && (int64_t)c.ntiles * bpt < want && per_tile / (2 * bpt) >= kThreads)
        bpt *= 2;
    return bpt;
}

// Accumulates rho of p into c.acc (fixed point).  sorted_off != nullptr:
// particles are physically tile-sorted with those offsets (tiles policy,
// direct reads).  policy BINNED without offsets bins first (stable).
void deposit_rho_fixed(aw_ctx& c, const aw_particles& p, double qm, int policy,
                       const int32_t* sorted_off, hipStream_t s, unsigned long long* acc) {
    if (p.n == 0) return;
    if (!acc) acc = c.acc.get();
    AW_REQUIRE(p.n < (int64_t)INT32_MAX, "particle count must fit int32");
    DepArgs a{};
    a.x = p.x;
    a.y = p.y;
    a.w = p.w;
    a.n = (int)p.n;
    a.nx = c.g.nx;
    a.ny = c.g.ny;
    a.qm = qm;
    a.area = c.g.dx * c.g.dy;
    a.coef_u = qm * p.w_uniform / a.area;
    a.scale = std::ldexp(1.0, c.fix_log2);
    const int64_t nc = ncell(c);
    if (policy == AW_DEPOSIT_AUTO) {
        const double avg_ppc = (double)p.n / (double)nc;
        if (sorted_off) policy = AW_DEPOSIT_BINNED;
        else if ((size_t)nc * 8 <= c.lds_per_block) policy = AW_DEPOSIT_LDS_GRID;
        else policy = avg_ppc >= 24.0 ? AW_DEPOSIT_BINNED : AW_DEPOSIT_GLOBAL;
    }
    if (policy == AW_DEPOSIT_GLOBAL) {
        deposit_global_kernel<<<blocks_for(p.n), kThreads, 0, s>>>(a, acc);
        AW_LAUNCH_CHECK();
    } else if (policy == AW_DEPOSIT_LDS_GRID) {
        const size_t sh = (size_t)nc * 8;
        AW_REQUIRE(sh <= c.lds_per_block, "grid does not fit LDS for AW_DEPOSIT_LDS_GRID");
        int per_cu = 0;
        AW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(deposit_ldsgrid_kernel), kThreads, sh));
        int nb = std::max(1, per_cu) * std::max(1, c.num_cu);
        nb = std::min(nb, blocks_for(p.n));
        deposit_ldsgrid_kernel<<<nb, kThreads, sh, s>>>(a, acc);
        AW_LAUNCH_CHECK();
    } else if (policy == AW_DEPOSIT_BINNED) {
        TileGeom t{};
        t.ntx = c.ntx;
        t.tx = c.tx;
        t.ty = c.ty;
        if (sorted_off) {
            t.off = sorted_off;
            t.idx = nullptr;
            t.apron = c.apron;
        } else {
            // Particles::build_tile_bins (include/pic/particles.hpp:632-671)
            c.sort.bin_off.reserve((size_t)c.ntiles + 1);
            c.sort.bin_perm.reserve((size_t)std::max<int64_t>(1, p.n));
            bin_tiles(c, p, c.sort.bin_perm.get(), c.sort.bin_off.get(), s);
            t.off = c.sort.bin_off.get();
            t.idx = c.sort.bin_perm.get();
            t.apron = 0;
        }
        t.W = c.tx + 2 * t.apron + 1;
        t.H = c.ty + 2 * t.apron + 1;
        t.bpt = tile_bpt(c, p.n);
        const size_t sh = (size_t)t.W * t.H * 8;
        deposit_tiles_kernel<<<c.ntiles * t.bpt, kThreads, sh, s>>>(a, t, acc);
        AW_LAUNCH_CHECK();
    } else {
        fail(AW_ERR_INVALID, "unknown deposit policy " + std::to_string(policy), __FILE__,
             __LINE__);
    }
}

}  // namespace aw
