// esirkepov.hip -- charge-conserving (Esirkepov) current deposit.
//
// Reference: s1_shape4, esirkepov_scatter, GlobalJSink and the deposit half
// of k_push_esirkepov / k_push_esirkepov_tiled (include/pic/yee2d.hpp:124-132,
// 310-389, 401-469).  For a move (x0, y0) -> (x1, y1) of less than one cell the
// 4x4 union stencil {ib-1..ib+2} x {jb-1..jb+2}, ib = floor(min(x0, x1)),
// carries
//   Jx(i+1/2, j) = -q w/(dt dy) * sum_{a<=i} DSx_a (Sy0_j + DSy_j/2)
//   Jy(i, j+1/2) = -q w/(dt dx) * sum_{b<=j} DSy_b (Sx0_i + DSx_i/2)
//   Jz(i, j)     =  q w uz/(dx dy) * (Sx0 Sy0 + (DSx Sy0 + Sx0 DSy)/2 + DSx DSy/3)
// with S the CIC node shape max(0, 1 - |x - node|) and DS = S(x1) - S(x0).
//
// MI355X design: shapes and products in FP64, contributions summed as int64
// fixed point (order-free, so every policy returns the same bits).  The
// per-component scale comes from a bound over the particles: max |w| and
// max |w uz| give each component's largest single contribution, which maps to
// at most 2^44 (2^19 such contributions of headroom per node).  The bound is
// a max over per-workgroup partial maxima (no atomics), taken by a bound pass
// here or handed in by a push that already holds the values.  Policies:
// global atomics, whole grid in LDS, tiles binned by the old position with a
// 2-node apron in LDS, or -- for particles the caller keeps in tile order --
// the tile slices of the last sort with a 4-node apron (moves that leave the
// apron fall back to global atomics).  Each kernel can write the wrapped new
// position back, which folds the periodic wrap of the caller's move into the
// pass that already reads it.
#include "esk_device.hpp"

namespace aw {

namespace {

using namespace esk;

constexpr int kThreads = 256;
constexpr int kApron = 2;        // particles binned by this step's old position

inline int blocks_for(int64_t n) { return (int)((n + kThreads - 1) / kThreads); }

// per-workgroup max |w| and max |w uz| (max |uz| for a uniform weight) into
// part[2 b], part[2 b + 1]
__global__ void __launch_bounds__(kThreads) esk_bound_kernel(EskArgs a, float* __restrict__ part) {
    float mw = 0.f, mwu = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += blockDim.x * gridDim.x) {
        const float uz = fabsf(a.uz[i]);
        if (a.w) {
            const float w = fabsf(a.w[i]);
            mw = fmaxf(mw, w);
            mwu = fmaxf(mwu, w * uz);
        } else {
            mwu = fmaxf(mwu, uz);
        }
    }
    dev::block_max2(mw, mwu);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = mw;
        part[2 * blockIdx.x + 1] = mwu;
    }
}

// bound[0..1] = the maxima of (f0, f1) and the npart partial pairs, as
// positive-float bit patterns
constexpr int kFinalThreads = 1024;
__global__ void __launch_bounds__(kFinalThreads) esk_bound_final_kernel(const float* __restrict__ part,
                                                                   int npart, float f0, float f1,
                                                                   unsigned int* __restrict__ bound) {
    float mw = f0, mwu = f1;
    for (int b = threadIdx.x; b < npart; b += blockDim.x) {
        mw = fmaxf(mw, part[2 * b]);
        mwu = fmaxf(mwu, part[2 * b + 1]);
    }
    dev::block_max2(mw, mwu);
    if (threadIdx.x == 0) {
        bound[0] = __float_as_uint(mw);
        bound[1] = __float_as_uint(mwu);
    }
}

// the accumulators cleared and the fixed bounds written, in one launch
__global__ void __launch_bounds__(kThreads) esk_begin_kernel(unsigned long long* __restrict__ acc,
                                                             int64_t n, float f0, float f1,
                                                             unsigned int* __restrict__ bound) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (int64_t)kThreads * gridDim.x)
        acc[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        bound[0] = __float_as_uint(f0);
        bound[1] = __float_as_uint(f1);
    }
}

__device__ __forceinline__ void write_wrapped(const EskArgs& a, int i) {
    if (a.xw) {
        a.xw[i] = dev::wrap_coord(a.x1[i], (float)a.nx);
        a.yw[i] = dev::wrap_coord(a.y1[i], (float)a.ny);
    }
}

__device__ __forceinline__ void global_add(unsigned long long* acc, int nx, int ny, int64_t nc,
                                           int q, int i, int j, unsigned long long v) {
    atomicAdd(&acc[q * nc + (int64_t)dev::wrapfar(j, ny) * nx + dev::wrapfar(i, nx)], v);
}

template <bool ONE_ROW>
__global__ void __launch_bounds__(kThreads) esk_global_kernel(EskArgs a,
                                                              unsigned long long* __restrict__ acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double kc[3];
    esk_coefs(a, kc);
    const int64_t nc = (int64_t)a.nx * a.ny;
    EskMove m;
    esk_prepare(a, i, kc, m);
    esk_emit<ONE_ROW>(m, [&](int q, int k, int b, unsigned long long v) {
        global_add(acc, a.nx, a.ny, nc, q, m.sh.ib - 1 + k, m.sh.jb - 1 + b, v);
    });
    write_wrapped(a, i);
}

// whole grid (3 components) privatised in LDS per resident workgroup; the
// grid takes most of a CU's LDS, so a workgroup is 16 waves to keep the CU
// busy
constexpr int kGridThreads = 1024;
template <bool ONE_ROW>
__global__ void __launch_bounds__(kGridThreads) esk_ldsgrid_kernel(EskArgs a,
                                                               unsigned long long* __restrict__ acc) {
    extern __shared__ unsigned long long s_j[];
    const int nc = a.nx * a.ny;
    for (int c = threadIdx.x; c < 3 * nc; c += blockDim.x) s_j[c] = 0ull;
    __syncthreads();
    double kc[3];
    esk_coefs(a, kc);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += blockDim.x * gridDim.x) {
        EskMove m;
        esk_prepare(a, i, kc, m);
        esk_emit<ONE_ROW>(m, [&](int q, int k, int b, unsigned long long v) {
            const int gi = dev::wrapfar(m.sh.ib - 1 + k, a.nx);
            const int gj = dev::wrapfar(m.sh.jb - 1 + b, a.ny);
            atomicAdd(&s_j[q * nc + gj * a.nx + gi], v);
        });
        write_wrapped(a, i);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < 3 * nc; c += blockDim.x) {
        const unsigned long long v = s_j[c];
        if (v) atomicAdd(&acc[c], v);
    }
}

// The whole one-row grid in LDS (ny = 1), for particles that may arrive in
// cell order: when every particle of a wavefront starts its stencil at the
// same column, the wavefront sums its eight column quanta across the lanes
// (exact int64 butterflies) and one lane adds them, instead of 64 lanes
// adding to the same LDS words one after another.  Otherwise each lane adds
// its own.
__global__ void __launch_bounds__(kGridThreads) esk_ldsgrid_row_kernel(EskArgs a,
                                                                   unsigned long long* __restrict__ acc) {
    extern __shared__ unsigned long long s_j[];
    const int nc = a.nx;
    for (int c = threadIdx.x; c < 3 * nc; c += blockDim.x) s_j[c] = 0ull;
    __syncthreads();
    double kc[3];
    esk_coefs(a, kc);
    const int stride = blockDim.x * gridDim.x;
    // one iteration's particle loads are in flight while the previous one is
    // deposited
    float p0[6] = {};
    auto load = [&](int i, float* v) {
        if (i < a.n) {
            v[0] = a.x0[i];
            v[1] = a.y0[i];
            v[2] = a.x1[i];
            v[3] = a.y1[i];
            v[4] = a.uz[i];
            v[5] = a.w ? a.w[i] : 0.f;
        }
    };
    load(blockIdx.x * blockDim.x + (int)threadIdx.x, p0);
    for (int base = blockIdx.x * blockDim.x; base < a.n; base += stride) {
        const int i = base + (int)threadIdx.x;
        const bool valid = i < a.n;
        float p[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) p[k] = p0[k];
        load(i + stride, p0);
        unsigned long long col[3][4] = {};
        int key = 0;
        if (valid) {
            EskMove m;
            esk_prepare_v(a, p[0], p[1], p[2], p[3], p[4], a.w ? (double)p[5] : a.w_uniform,
                          kc, m);
            dev::esk_emit(m.sh, m.cJx, m.cJy, m.cJz,
                          [&](int q, int k, int, unsigned long long v) { col[q][k] += v; });
            key = dev::wrapfar(m.sh.ib - 1, a.nx);
            if (a.xw) {
                a.xw[i] = dev::wrap_coord(p[2], (float)a.nx);
                a.yw[i] = dev::wrap_coord(p[3], (float)a.ny);
            }
        }
        // the first valid lane's column base; invalid lanes carry zeros
        const unsigned long long vmask = __ballot(valid);
        const int lead = __ffsll((long long)vmask) - 1;
        const int k0 = __shfl(key, lead, 64);
        const bool uniform = __ballot(valid && key != k0) == 0ull;
        if (uniform) {
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int k = 1; k < 4; ++k) {
                    unsigned long long v = col[q][k];
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                    col[q][k] = v;
                }
            if ((int)(threadIdx.x & 63) == lead) {
#pragma unroll
                for (int q = 0; q < 3; ++q)
#pragma unroll
                    for (int k = 1; k < 4; ++k)
                        if (col[q][k]) {
                            atomicAdd(&s_j[q * nc + dev::wrapfar(k0 + k, a.nx)], col[q][k]);
                        }
            }
        } else if (valid) {
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int k = 1; k < 4; ++k)
                    if (col[q][k]) {
                        const int gi = dev::wrapfar(key + k, a.nx);
                        atomicAdd(&s_j[q * nc + gi], col[q][k]);
                    }
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < 3 * nc; c += blockDim.x) {
        const unsigned long long v = s_j[c];
        if (v) atomicAdd(&acc[c], v);
    }
}

struct EskTiles {
    const int32_t* off;  // ntiles+1, tile bins of the old positions (or of the last sort)
    const int32_t* idx;  // bin permutation (NULL: the particles are in tile order)
    int ntx, tx, ty, A, W, H, bpt;  // A: apron nodes on each side, W = tx + 2A, H = ty + 2A
};

// workgroup per (tile, part): the tile's J plus an A-node apron in LDS
template <bool ONE_ROW>
__global__ void __launch_bounds__(kThreads) esk_tiles_kernel(EskArgs a, EskTiles t,
                                                             unsigned long long* __restrict__ acc) {
    extern __shared__ unsigned long long s_j[];
    const int R = t.W * t.H;
    // a tile's workgroups on one XCD: their flushes meet in one L2
    const int lb = dev::xcd_block(blockIdx.x, gridDim.x);
    const int tile = lb / t.bpt, part = lb % t.bpt;
    const int orgx = (tile % t.ntx) * t.tx - t.A, orgy = (tile / t.ntx) * t.ty - t.A;
    for (int c = threadIdx.x; c < 3 * R; c += blockDim.x) s_j[c] = 0ull;
    __syncthreads();
    double kc[3];
    esk_coefs(a, kc);
    const int64_t nc = (int64_t)a.nx * a.ny;
    const int beg = t.off[tile], end = t.off[tile + 1];
    const int chunk = (end - beg + t.bpt - 1) / t.bpt;
    const int r0 = beg + part * chunk, r1 = min(end, r0 + chunk);
    for (int r = r0 + (int)threadIdx.x; r < r1; r += blockDim.x) {
        const int i = t.idx ? t.idx[r] : r;
        EskMove m;
        esk_prepare(a, i, kc, m);
        const int li = wrap_local(m.sh.ib - 1 - orgx, a.nx), lj = wrap_local(m.sh.jb - 1 - orgy, a.ny);
        if (dev::in_tile(li, t.W - 4) && dev::in_tile(lj, t.H - 4)) {
            const int base = lj * t.W + li;
            esk_emit<ONE_ROW>(m, [&](int q, int k, int b, unsigned long long v) {
                atomicAdd(&s_j[q * R + base + b * t.W + k], v);
            });
        } else {
            esk_emit<ONE_ROW>(m, [&](int q, int k, int b, unsigned long long v) {
                global_add(acc, a.nx, a.ny, nc, q, m.sh.ib - 1 + k, m.sh.jb - 1 + b, v);
            });
        }
        write_wrapped(a, i);
    }
    __syncthreads();
    // (row, column) over the three components' 3 H rows, advanced by the
    // stride without dividing
    const int dr = (int)blockDim.x / t.W, dc = (int)blockDim.x % t.W;
    int row = (int)threadIdx.x / t.W, col = (int)threadIdx.x % t.W;
    for (int c = threadIdx.x; c < 3 * R; c += blockDim.x) {
        const unsigned long long v = s_j[c];
        if (v) {
            const int q = row >= 2 * t.H ? 2 : (row >= t.H ? 1 : 0);
            const int gx = dev::wrapfar(orgx + col, a.nx);
            const int gy = dev::wrapfar(orgy + row - q * t.H, a.ny);
            atomicAdd(&acc[q * nc + (int64_t)gy * a.nx + gx], v);
        }
        row += dr;
        col += dc;
        if (col >= t.W) {
            col -= t.W;
            ++row;
        }
    }
}

// J_q[c] += acc_q[c] * 2^-S_q
__global__ void __launch_bounds__(kThreads) esk_convert_kernel(EskArgs a,
                                                               const unsigned long long* __restrict__ acc,
                                                               float* jx, float* jy, float* jz) {
    const int64_t nc = (int64_t)a.nx * a.ny;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int S[3];
    esk_exponents(a, S);
    float* out[3] = {jx, jy, jz};
    for (int q = 0; q < 3; ++q) {
        out[q][c] += esk_value((long long)acc[q * nc + c], S[q]);
    }
}

}  // namespace

int tile_bpt(const aw_ctx& c, int64_t n);

namespace {
// the one-row (ny = 1) instance of a kernel, or the general one
template <class K, class... Args>
void launch_one_row(K one, K general, bool one_row, int grid, int threads, size_t sh,
                    hipStream_t s, Args... args) {
    hipLaunchKernelGGL(one_row ? one : general, dim3(grid), dim3(threads), sh, s, args...);
}
}  // namespace

void deposit_esirkepov(aw_ctx& c, const float* x0, const float* y0, const aw_particles& p,
                       double qm, double dt, int policy, float* jx, float* jy, float* jz,
                       hipStream_t s, const EskExtra& ex) {
    AW_REQUIRE(p.n < (int64_t)INT32_MAX, "particle count must fit int32");
    EskArgs a = esk::setup(c, x0, y0, p, qm, dt, s, ex);
    const int64_t nc = ncell(c);
    unsigned long long* acc = c.jacc.get();
    const bool one_row = c.g.ny == 1;
    if (p.n > 0) {
        if (ex.slice_off) {
            AW_REQUIRE(policy == AW_DEPOSIT_AUTO || policy == AW_DEPOSIT_BINNED,
                       "tile-ordered particles take the binned deposit");
            EskTiles t{};
            t.off = ex.slice_off;
            t.idx = nullptr;
            t.ntx = c.ntx;
            t.tx = c.tx;
            t.ty = c.ty;
            t.A = kSortedApron;
            t.W = c.tx + 2 * t.A;
            t.H = c.ty + 2 * t.A;
            t.bpt = tile_bpt(c, p.n);
            const size_t sh = (size_t)3 * t.W * t.H * 8;
            AW_REQUIRE(sh <= c.lds_per_block, "tile apron does not fit LDS");
            launch_one_row(esk_tiles_kernel<true>, esk_tiles_kernel<false>, one_row, c.ntiles * t.bpt, kThreads, sh, s, a, t, acc);
            AW_LAUNCH_CHECK();
            policy = -1;  // done
        } else if (policy == AW_DEPOSIT_AUTO) {
            if ((size_t)(3 * nc) * 8 <= c.lds_per_block && p.n >= 16 * nc)
                policy = AW_DEPOSIT_LDS_GRID;
            else
                policy = (double)p.n / (double)nc >= 8.0 ? AW_DEPOSIT_BINNED : AW_DEPOSIT_GLOBAL;
        }
        if (policy == -1) {
        } else if (policy == AW_DEPOSIT_GLOBAL) {
            launch_one_row(esk_global_kernel<true>, esk_global_kernel<false>, one_row, blocks_for(p.n), kThreads, 0, s, a, acc);
            AW_LAUNCH_CHECK();
        } else if (policy == AW_DEPOSIT_LDS_GRID) {
            const size_t sh = (size_t)(3 * nc) * 8;
            AW_REQUIRE(sh <= c.lds_per_block, "grid does not fit LDS for AW_DEPOSIT_LDS_GRID");
            int per_cu = 0;
            AW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void*>(esk_ldsgrid_kernel<false>), kGridThreads, sh));
            const int nb = (int)std::min<int64_t>(std::max(1, per_cu) * std::max(1, c.num_cu),
                                                  (p.n + kGridThreads - 1) / kGridThreads);
            if (one_row)
                esk_ldsgrid_row_kernel<<<nb, kGridThreads, sh, s>>>(a, acc);
            else
                esk_ldsgrid_kernel<false><<<nb, kGridThreads, sh, s>>>(a, acc);
            AW_LAUNCH_CHECK();
        } else if (policy == AW_DEPOSIT_BINNED) {
            // bin by the old position (stable), as deposit_rho_fixed does
            aw_particles old = p;
            old.x = const_cast<float*>(x0);
            old.y = const_cast<float*>(y0);
            c.sort.bin_off.reserve((size_t)c.ntiles + 1);
            c.sort.bin_perm.reserve((size_t)std::max<int64_t>(1, p.n));
            bin_tiles(c, old, c.sort.bin_perm.get(), c.sort.bin_off.get(), s);
            EskTiles t{};
            t.off = c.sort.bin_off.get();
            t.idx = c.sort.bin_perm.get();
            t.ntx = c.ntx;
            t.tx = c.tx;
            t.ty = c.ty;
            t.A = kApron;
            t.W = c.tx + 2 * kApron;
            t.H = c.ty + 2 * kApron;
            t.bpt = tile_bpt(c, p.n);
            const size_t sh = (size_t)3 * t.W * t.H * 8;
            launch_one_row(esk_tiles_kernel<true>, esk_tiles_kernel<false>, one_row, c.ntiles * t.bpt, kThreads, sh, s, a, t, acc);
            AW_LAUNCH_CHECK();
        } else {
            fail(AW_ERR_INVALID, "unknown deposit policy " + std::to_string(policy), __FILE__,
                 __LINE__);
        }
    }
    if (!ex.no_convert) esk::convert(c, a, jx, jy, jz, s);
}

namespace esk {

EskArgs args(aw_ctx& c, const float* x0, const float* y0, const aw_particles& p, double qm,
             double dt, const EskExtra& ex) {
    c.jbound.reserve(2);
    EskArgs a{};
    a.x0 = x0;
    a.y0 = y0;
    a.x1 = p.x;
    a.y1 = p.y;
    a.uz = p.uz;
    a.w = p.w;
    a.per_w = ex.fixed ? ex.fixed_per_w : (p.w != nullptr);
    a.xw = ex.xw;
    a.yw = ex.yw;
    a.n = (int)p.n;
    a.nx = c.g.nx;
    a.ny = c.g.ny;
    a.qm = qm;
    a.w_uniform = p.w_uniform;
    a.dt = dt;
    a.dx = c.g.dx;
    a.dy = c.g.dy;
    a.bound = c.jbound.get();
    return a;
}

EskArgs setup(aw_ctx& c, const float* x0, const float* y0, const aw_particles& p, double qm,
              double dt, hipStream_t s, const EskExtra& ex) {
    AW_REQUIRE(p.n < (int64_t)INT32_MAX, "particle count must fit int32");
    AW_REQUIRE(dt > 0, "dt must be positive");
    AW_REQUIRE(!ex.xw == !ex.yw, "wrapped x and y outputs go together");
    const int64_t nc = ncell(c);
    c.jacc.reserve((size_t)(3 * nc));
    if (!ex.fixed)
        AW_HIP(hipMemsetAsync(c.jacc.get(), 0, (size_t)(3 * nc) * sizeof(unsigned long long), s));
    const EskArgs a = args(c, x0, y0, p, qm, dt, ex);
    if (ex.fixed) {
        AW_REQUIRE(ex.fixed_bound[0] >= 0.f && ex.fixed_bound[1] >= 0.f &&
                       std::isfinite(ex.fixed_bound[0]) && std::isfinite(ex.fixed_bound[1]),
                   "the current's fixed scale bounds must be finite and non-negative");
        const int nb = (int)std::min<int64_t>((3 * nc + kThreads - 1) / kThreads,
                                              8 * std::max(1, c.num_cu));
        esk_begin_kernel<<<nb, kThreads, 0, s>>>(c.jacc.get(), 3 * nc, ex.fixed_bound[0],
                                                 ex.fixed_bound[1], c.jbound.get());
        AW_LAUNCH_CHECK();
        return a;
    }
    const int npart = p.n > 0 ? std::min(blocks_for(p.n), 4 * std::max(1, c.num_cu)) : 0;
    c.jpart.reserve((size_t)std::max(2, 2 * npart));
    if (npart > 0) {
        esk_bound_kernel<<<npart, kThreads, 0, s>>>(a, c.jpart.get());
        AW_LAUNCH_CHECK();
    }
    esk_bound_final_kernel<<<1, kFinalThreads, 0, s>>>(c.jpart.get(), npart, 0.f, 0.f,
                                                       c.jbound.get());
    AW_LAUNCH_CHECK();
    return a;
}

void convert(aw_ctx& c, const EskArgs& a, float* jx, float* jy, float* jz, hipStream_t s) {
    esk_convert_kernel<<<blocks_for(ncell(c)), kThreads, 0, s>>>(a, c.jacc.get(), jx, jy, jz);
    AW_LAUNCH_CHECK();
}

}  // namespace esk

}  // namespace aw
