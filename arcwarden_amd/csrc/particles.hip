// particles.hip -- integer layer: cell index, tile keys, stable LSD radix
// sort (wavefront ballot ranking), SoA permutation, migrate, loaders.
//
// Reference behaviour: include/pic/particles.hpp (tile_of :74-77, counting
// sort :84-153, physical sort :162-185/:678-726, migrate :293-315, loaders
// :190-290/:318-379).  The reference hands out in-tile slots in atomic
// order; here every sort is a stable radix sort, so the permutation is a
// pure function of the keys and matches a scalar host counting sort.
#include "aw_internal.hpp"

namespace aw {

namespace {

using dev::radical_inverse;
using dev::rng_uniform;

constexpr int kThreads = 256;
constexpr int kRsRounds = 32;                       // 256 x 32 = 8192 keys / chunk
constexpr int kRsChunk = kThreads * kRsRounds;
constexpr int kScanSeg = kThreads * 8;

inline int blocks_for(int64_t n, int per_block = kThreads) {
    return (int)((n + per_block - 1) / per_block);
}

// ---- keys --------------------------------------------------------------------

__global__ void tile_key_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                int n, int nx, int ny, int tx, int ty, int ntx, int cell_key,
                                uint32_t* __restrict__ key) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = dev::cell_of(x[i], y[i], nx, ny);
    const int ci = c % nx, cj = c / nx;
    const uint32_t t = (uint32_t)((cj / ty) * ntx + ci / tx);
    key[i] = cell_key ? t * (uint32_t)(tx * ty) + (uint32_t)((cj % ty) * tx + (ci % tx)) : t;
}

__global__ void cell_index_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                  int n, int nx, int ny, int32_t* __restrict__ cell) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cell[i] = dev::cell_of(x[i], y[i], nx, ny);
}

// ---- radix sort: per-chunk digit histogram -------------------------------------

__global__ void __launch_bounds__(kThreads)
rs_hist_kernel(const uint32_t* __restrict__ keys, int n, int shift, int nchunks,
               uint32_t* __restrict__ table) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRsChunk;
    const int end = min(n, base + kRsChunk);
    for (int i = base + (int)threadIdx.x; i < end; i += kThreads)
        atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    __syncthreads();
    table[(size_t)threadIdx.x * nchunks + blockIdx.x] = h[threadIdx.x];
}

// ---- device-wide exclusive scan (three phases) ----------------------------------

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh,
                                                         uint32_t* total) {
    // sh: kThreads entries; Hillis-Steele inclusive then shift
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < kThreads; d <<= 1) {
        const uint32_t a = threadIdx.x >= (unsigned)d ? sh[threadIdx.x - d] : 0u;
        __syncthreads();
        sh[threadIdx.x] += a;
        __syncthreads();
    }
    const uint32_t incl = sh[threadIdx.x];
    *total = sh[kThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(kThreads)
scan_reduce_kernel(const uint32_t* __restrict__ in, int64_t n, uint32_t* __restrict__ sums) {
    __shared__ uint32_t sh[kThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanSeg + (int64_t)threadIdx.x * 8;
    uint32_t v = 0;
    for (int k = 0; k < 8; ++k)
        if (base + k < n) v += in[base + k];
    uint32_t tot;
    (void)block_exclusive_scan(v, sh, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kThreads)
scan_sums_kernel(uint32_t* __restrict__ sums, int nsums) {
    __shared__ uint32_t sh[kThreads];
    uint32_t carry = 0;
    for (int base = 0; base < nsums; base += kThreads) {
        const int i = base + (int)threadIdx.x;
        const uint32_t v = i < nsums ? sums[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, sh, &tot);
        if (i < nsums) sums[i] = carry + ex;
        carry += tot;
    }
}

__global__ void __launch_bounds__(kThreads)
scan_apply_kernel(uint32_t* __restrict__ data, int64_t n, const uint32_t* __restrict__ sums) {
    __shared__ uint32_t sh[kThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanSeg + (int64_t)threadIdx.x * 8;
    uint32_t loc[8];
    uint32_t v = 0;
    for (int k = 0; k < 8; ++k) {
        loc[k] = base + k < n ? data[base + k] : 0u;
        v += loc[k];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(v, sh, &tot) + sums[blockIdx.x];
    for (int k = 0; k < 8; ++k) {
        if (base + k < n) data[base + k] = run;
        run += loc[k];
    }
}

void exclusive_scan(aw_ctx& c, uint32_t* data, int64_t n, hipStream_t s) {
    const int nseg = (int)((n + kScanSeg - 1) / kScanSeg);
    c.sort.sums.reserve((size_t)nseg + 1);
    scan_reduce_kernel<<<nseg, kThreads, 0, s>>>(data, n, c.sort.sums.get());
    AW_LAUNCH_CHECK();
    scan_sums_kernel<<<1, kThreads, 0, s>>>(c.sort.sums.get(), nseg);
    AW_LAUNCH_CHECK();
    scan_apply_kernel<<<nseg, kThreads, 0, s>>>(data, n, c.sort.sums.get());
    AW_LAUNCH_CHECK();
}

// ---- radix sort: stable scatter with wavefront ballot ranking ------------------

__global__ void __launch_bounds__(kThreads)
rs_scatter_kernel(const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin,
                  uint32_t* __restrict__ kout, int32_t* __restrict__ vout, int n, int shift,
                  int nchunks, const uint32_t* __restrict__ table) {
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[kThreads / 64][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    base[tid] = table[(size_t)tid * nchunks + blockIdx.x];
    for (int w = 0; w < kThreads / 64; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    const int cbase = blockIdx.x * kRsChunk;
    const int rounds = min(kRsRounds, (n - cbase + kThreads - 1) / kThreads);
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int r = 0; r < rounds; ++r) {
        const int i = cbase + r * kThreads + tid;
        const bool valid = i < n;
        const uint32_t key = valid ? kin[i] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
        #pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
            kout[pos] = key;
            vout[pos] = vin ? vin[i] : i;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kThreads / 64; ++w) {
            add += wcnt[w][tid];
            wcnt[w][tid] = 0;
        }
        base[tid] += add;
        __syncthreads();
    }
}

__global__ void key_offsets_kernel(const uint32_t* __restrict__ skeys, int n, uint32_t nkeys,
                                   int32_t* __restrict__ off) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const int64_t prev = (i == 0) ? -1 : (int64_t)skeys[i - 1];
    const int64_t cur = (i == n) ? (int64_t)nkeys : (int64_t)skeys[i];
    for (int64_t t = prev + 1; t <= cur; ++t) off[t] = i;
}

// ---- one-pass stable tile sort that moves the SoA ------------------------------------
//
// For particles that are already nearly tile-sorted (the previous sort plus a
// few steps of drift) a single stable counting sort by tile id beats the LSD
// radix + permute path: one histogram pass over (x, y), one scan of the
// chunk x tile count table, and one scatter that ranks each particle inside
// its chunk with wavefront ballots and writes it straight to its final slot.
// In the scatter a chunk belongs to one single-wave workgroup: ranking needs
// no cross-wave bookkeeping, and the LDS table is one 32-bit counter per
// tile, so many workgroups stay resident per CU; each wave loads kTsAhead
// rounds ahead of the one it ranks to keep enough reads in flight.  The
// histogram only counts, so four waves share a chunk and its LDS table.

constexpr int kTsWave = 64;
constexpr int kTsHistThreads = 256;
constexpr int kTsLoads = 4;  // particles per thread per histogram iteration
constexpr int kTsAhead = 8;  // rounds of 64 particles in flight per scatter wave
constexpr int kTsMatch = 4;  // tiles matched one by one per scatter round

// rounds of 64 particles per chunk: at least 16, more when the chunk x tile
// count table would exceed 2^24 entries
inline int ts_rounds(int64_t n, int ntiles) {
    int64_t r = 16;
    while (r < 4096 && ((n + kTsWave * r - 1) / (kTsWave * r)) * (int64_t)ntiles > (1ll << 24))
        r *= 2;
    return (int)r;
}

// Tile geometry for tile ids; rtx, rty = 1/tx, 1/ty in FP32.
struct TileMap {
    int nx, ny, tx, ty, ntx;
    float rtx, rty;
};

// a / b for 0 <= a < 2^24, b >= 1: the FP32 estimate a * (1/b) is within one
// of the quotient, and one integer check each way makes it exact
__device__ __forceinline__ int div_small(int a, int b, float rb) {
    int q = (int)((float)a * rb);
    if (q * b > a) --q;
    else if ((q + 1) * b <= a) ++q;
    return q;
}

// tile of the cell holding (x, y), cells clamped to the grid as cell_of does
__device__ __forceinline__ int tile_id(float x, float y, const TileMap& m) {
    const int ci = min(max((int)floorf(x), 0), m.nx - 1);
    const int cj = min(max((int)floorf(y), 0), m.ny - 1);
    return div_small(cj, m.ty, m.rty) * m.ntx + div_small(ci, m.tx, m.rtx);
}

__global__ void __launch_bounds__(kTsHistThreads)
ts_hist_kernel(const float* __restrict__ x, const float* __restrict__ y, int n, TileMap tm,
               int ntiles, int rounds, uint32_t* __restrict__ table) {
    extern __shared__ uint32_t h[];
    for (int t = threadIdx.x; t < ntiles; t += kTsHistThreads) h[t] = 0;
    __syncthreads();
    const int lane = threadIdx.x & (kTsWave - 1);
    const int base = blockIdx.x * kTsWave * rounds;
    const int end = min(n, base + kTsWave * rounds);
    // kTsLoads particles per thread per iteration, all loaded before any is
    // counted (the counts do not depend on the order)
    for (int i0 = base; i0 < end; i0 += kTsHistThreads * kTsLoads) {
        float px[kTsLoads], py[kTsLoads];
#pragma unroll
        for (int k = 0; k < kTsLoads; ++k) {
            const int i = i0 + k * kTsHistThreads + (int)threadIdx.x;
            px[k] = i < end ? x[i] : 0.f;
            py[k] = i < end ? y[i] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < kTsLoads; ++k) {
            const bool valid = i0 + k * kTsHistThreads + (int)threadIdx.x < end;
            const int t = valid ? tile_id(px[k], py[k], tm) : -1;
            // one LDS atomic per distinct tile of the wave
            unsigned long long todo = __ballot(valid);
            while (todo) {
                const int leader = __ffsll((long long)todo) - 1;
                const int lt = __builtin_amdgcn_readlane(t, leader);
                const unsigned long long mk = __ballot(t == lt) & todo;
                if (lane == leader) atomicAdd(&h[lt], (uint32_t)__popcll(mk));
                todo &= ~mk;
            }
        }
    }
    __syncthreads();
    // chunk-major table: this chunk's counts are one contiguous row
    uint32_t* row = table + (size_t)blockIdx.x * ntiles;
    for (int t = threadIdx.x; t < ntiles; t += kTsHistThreads) row[t] = h[t];
}

// Offsets of the chunk-major count table T[chunk][tile] in tile-major order:
// off(t, c) = base(t) + sum_{c' < c} T[c'][t], base(t) = sum_{t' < t} total(t').
// Chunks are taken kTsColBlock at a time: ts_colsum sums each (block, tile)
// column segment, ts_colscan scans the block sums per tile (and yields the
// tile totals), the totals are scanned into base(t), and ts_colapply rewrites
// T in place with the offsets.  Every pass reads rows of T, coalesced across
// the tiles.
constexpr int kTsColBlock = 64;

__global__ void ts_colsum_kernel(const uint32_t* __restrict__ table, int nchunks, int ntiles,
                                 uint32_t* __restrict__ bsum) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int c0 = blockIdx.y * kTsColBlock, c1 = min(nchunks, c0 + kTsColBlock);
    uint32_t v = 0;
    for (int c = c0; c < c1; ++c) v += table[(size_t)c * ntiles + t];
    bsum[(size_t)blockIdx.y * ntiles + t] = v;
}

// bsum[b][t] -> exclusive prefix over b; total[t] = column total
__global__ void ts_colscan_kernel(uint32_t* __restrict__ bsum, int nblocks, int ntiles,
                                  uint32_t* __restrict__ total) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    uint32_t run = 0;
    for (int b = 0; b < nblocks; ++b) {
        const uint32_t v = bsum[(size_t)b * ntiles + t];
        bsum[(size_t)b * ntiles + t] = run;
        run += v;
    }
    total[t] = run;
}

__global__ void ts_colapply_kernel(uint32_t* __restrict__ table, int nchunks, int ntiles,
                                   const uint32_t* __restrict__ bpre,
                                   const uint32_t* __restrict__ base) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int c0 = blockIdx.y * kTsColBlock, c1 = min(nchunks, c0 + kTsColBlock);
    uint32_t run = base[t] + bpre[(size_t)blockIdx.y * ntiles + t];
    for (int c = c0; c < c1; ++c) {
        uint32_t* e = table + (size_t)c * ntiles + t;
        const uint32_t v = *e;
        *e = run;
        run += v;
    }
}

struct TsMove {
    const float *x, *y, *ux, *uy, *uz, *w;
    float *xo, *yo, *uxo, *uyo, *uzo, *wo;
    int32_t* perm;  // optional: perm[pos] = source index
};

// MOVE = false: only the permutation (m.perm) is written; the particles are
// neither moved nor read beyond x, y.
template <bool MOVE>
__global__ void __launch_bounds__(kTsWave)
ts_scatter_kernel(TsMove m, int n, TileMap tm, int ntiles, int rounds, int tbits,
                  const uint32_t* __restrict__ table) {
    extern __shared__ uint32_t base[];  // running output offset per tile
    const int lane = threadIdx.x;
    const uint32_t* row = table + (size_t)blockIdx.x * ntiles;
    for (int t = lane; t < ntiles; t += kTsWave) base[t] = row[t];
    __syncthreads();
    const int cbase = blockIdx.x * kTsWave * rounds;
    const int nr = min(rounds, (n - cbase + kTsWave - 1) / kTsWave);
    const unsigned long long below = (1ull << lane) - 1ull;
    // software pipeline: rounds r+1 .. r+kTsAhead are loaded before round r
    // is ranked and stored (the rounds still complete in order, so the
    // permutation is the stable one)
    float ring[kTsAhead][6];
    auto load = [&](int r, float* v) {
        const int i = cbase + r * kTsWave + lane;
        if (r < nr && i < n) {
            v[0] = m.x[i];
            v[1] = m.y[i];
            if constexpr (MOVE) {
                v[2] = m.ux[i];
                v[3] = m.uy[i];
                v[4] = m.uz[i];
                v[5] = m.w ? m.w[i] : 0.f;
            } else {
                v[2] = v[3] = v[4] = v[5] = 0.f;
            }
        } else {
            for (int q = 0; q < 6; ++q) v[q] = 0.f;
        }
    };
#pragma unroll
    for (int p = 0; p < kTsAhead; ++p) load(p, ring[p]);
    for (int r0 = 0; r0 < nr; r0 += kTsAhead) {
#pragma unroll
        for (int p = 0; p < kTsAhead; ++p) {
            const int r = r0 + p;
            if (r >= nr) break;
            const int i = cbase + r * kTsWave + lane;
            const bool valid = i < n;
            const float x = ring[p][0], y = ring[p][1], ux = ring[p][2], uy = ring[p][3],
                        uz = ring[p][4], w = ring[p][5];
            load(r + kTsAhead, ring[p]);
            const int t = valid ? tile_id(x, y, tm) : -1;
            // peers = the lanes with this lane's tile.  A nearly sorted wave
            // spans few tiles: match them one at a time (the lowest pending
            // lane leads); past kTsMatch tiles, the rest split on tile-id
            // bits (a tile's lanes are all pending or all matched).
            unsigned long long todo = __ballot(valid), peers = 0;
            for (int it = 0; it < kTsMatch && todo; ++it) {
                const int lt = __builtin_amdgcn_readlane(t, __ffsll((long long)todo) - 1);
                const unsigned long long mk = __ballot(t == lt);
                if (t == lt) peers = mk;
                todo &= ~mk;
            }
            if (todo) {
                const bool mine = (todo >> lane) & 1ull;
                unsigned long long pk = todo;
                for (int b = 0; b < tbits; ++b) {
                    const bool bit = (t >> b) & 1;
                    const unsigned long long bal = __ballot(bit);
                    pk &= bit ? bal : ~bal;
                }
                if (mine) peers = pk;
            }
            const uint32_t rank = (uint32_t)__popcll(peers & below);
            // one wave per workgroup: its LDS accesses complete in program
            // order, so every lane reads base[t] before its leader advances it
            if (valid) {
                const uint32_t pos = base[t] + rank;
                if constexpr (MOVE) {
                    m.xo[pos] = x;
                    m.yo[pos] = y;
                    m.uxo[pos] = ux;
                    m.uyo[pos] = uy;
                    m.uzo[pos] = uz;
                    if (m.wo) m.wo[pos] = w;
                }
                if (m.perm) m.perm[pos] = i;
                if (rank == 0) base[t] += (uint32_t)__popcll(peers);
            }
        }
    }
}

__global__ void ts_offsets_kernel(const uint32_t* __restrict__ base, int ntiles, int n,
                                  int32_t* __restrict__ off) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ntiles) off[t] = (int32_t)base[t];
    if (t == ntiles) off[t] = n;
}

// ---- SoA permutation ---------------------------------------------------------------

__global__ void permute_kernel(const float* __restrict__ x, const float* __restrict__ y,
                               const float* __restrict__ ux, const float* __restrict__ uy,
                               const float* __restrict__ uz, const float* __restrict__ w,
                               float* __restrict__ xo, float* __restrict__ yo,
                               float* __restrict__ uxo, float* __restrict__ uyo,
                               float* __restrict__ uzo, float* __restrict__ wo,
                               const int32_t* __restrict__ perm, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = perm[k];
    xo[k] = x[i];
    yo[k] = y[i];
    uxo[k] = ux[i];
    uyo[k] = uy[i];
    uzo[k] = uz[i];
    if (w) wo[k] = w[i];
}

// ---- migrate ---------------------------------------------------------------------

__global__ void migrate_kernel(float* __restrict__ x, float* __restrict__ y, int n, int nx,
                               int ny, int32_t* __restrict__ cell) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float xv = dev::wrap_coord(x[i], (float)nx);
    const float yv = dev::wrap_coord(y[i], (float)ny);
    x[i] = xv;
    y[i] = yv;
    if (cell) cell[i] = dev::cell_of(xv, yv, nx, ny);
}

// ---- loaders ---------------------------------------------------------------------

struct SpeciesLoad {
    int ppc, index;
    int64_t base, count;
    double uth[3], ufl[3];
    int noisy;
    uint64_t seed;
    int nmodes;
    double mode_amp_cells[32];  // amplitude in cells
};

__device__ __forceinline__ float apply_modes(float x, const SpeciesLoad& sp, int nx) {
    if (sp.nmodes == 0) return x;
    const double k1 = 6.28318530717958647692 / (double)nx;
    double d = 0.0;
    for (int m = 0; m < sp.nmodes; ++m) d += sp.mode_amp_cells[m] * sin((m + 1) * k1 * (double)x);
    return dev::wrap_coord((float)((double)x + d), (float)nx);
}

// species_init_kernel (include/pic/particles.hpp:333-379); storage slot
// base + s*ncell + c (rank-major) so tile-sorted order interleaves cells.
__global__ void species_load_kernel(float* __restrict__ x, float* __restrict__ y,
                                    float* __restrict__ ux, float* __restrict__ uy,
                                    float* __restrict__ uz, int nx, int ny, SpeciesLoad sp) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= sp.count) return;
    const int64_t t = sp.base + l;
    const int c = (int)(l / sp.ppc), s = (int)(l % sp.ppc);
    const int i = c % nx, j = c / nx;
    float fx, fy;
    if (sp.noisy) {
        fx = (float)rng_uniform(t, 0, sp.seed);
        fy = (float)rng_uniform(t, 1, sp.seed);
    } else {
        const float px = (float)sp.index * 0.6180339887498949f;
        const float py = (float)sp.index * 0.7548776662466927f;
        fx = (float)radical_inverse((uint32_t)s + 1, 2) + px;
        fx -= floorf(fx);
        fy = (float)radical_inverse((uint32_t)s + 1, 3) + py;
        fy -= floorf(fy);
    }
    float xv = (float)i + fx, yv = (float)j + fy;
    if (xv >= (float)nx) xv = nextafterf((float)nx, 0.0f);
    if (yv >= (float)ny) yv = nextafterf((float)ny, 0.0f);
    xv = apply_modes(xv, sp, nx);
    const double qx = sp.noisy ? rng_uniform(t, 2, sp.seed) : radical_inverse((uint32_t)t + 1, 5);
    const double qy = sp.noisy ? rng_uniform(t, 3, sp.seed) : radical_inverse((uint32_t)t + 1, 7);
    const double qz = sp.noisy ? rng_uniform(t, 4, sp.seed) : radical_inverse((uint32_t)t + 1, 11);
    const double r2 = 1.41421356237309515;
    const int64_t slot = sp.base + (int64_t)s * ((int64_t)nx * ny) + c;
    x[slot] = xv;
    y[slot] = yv;
    ux[slot] = (float)(sp.ufl[0] + r2 * sp.uth[0] * erfinv(2.0 * qx - 1.0));
    uy[slot] = (float)(sp.ufl[1] + r2 * sp.uth[1] * erfinv(2.0 * qy - 1.0));
    uz[slot] = (float)(sp.ufl[2] + r2 * sp.uth[2] * erfinv(2.0 * qz - 1.0));
}

struct LegacyLoad {
    int ppc, nx, ny;
    double vth, vd, perturb_amp;
    int perturb_kx, two_stream, bump_on_tail, noisy;
    double beam_frac, beam_vd, beam_vth;
    uint64_t seed;
};

// particle_init_kernel (include/pic/particles.hpp:213-290), rank-major slot.
__global__ void legacy_load_kernel(float* __restrict__ x, float* __restrict__ y,
                                   float* __restrict__ ux, float* __restrict__ uy,
                                   float* __restrict__ uz, int64_t n, LegacyLoad lp) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int ppc = lp.ppc;
    const int c = (int)(t / ppc), s = (int)(t % ppc);
    const int i = c % lp.nx, j = c / lp.nx;
    const float fx = lp.noisy ? (float)rng_uniform(t, 0, lp.seed) : (float)radical_inverse(s + 1, 2);
    const float fy = lp.noisy ? (float)rng_uniform(t, 1, lp.seed) : (float)radical_inverse(s + 1, 3);
    float xv = (float)i + fx;
    const float yv0 = (float)j + fy;
    if (lp.perturb_amp != 0.0) {
        const float kx = 6.28318530717958648f * (float)lp.perturb_kx / (float)lp.nx;
        xv += (float)lp.perturb_amp * sinf(kx * xv);
        xv = fmodf(xv, (float)lp.nx);
        if (xv < 0.0f) xv += (float)lp.nx;
    }
    float yv = yv0;
    if (xv >= (float)lp.nx) xv = nextafterf((float)lp.nx, 0.0f);
    if (yv >= (float)lp.ny) yv = nextafterf((float)lp.ny, 0.0f);
    const double qx = lp.noisy ? rng_uniform(t, 2, lp.seed) : radical_inverse((uint32_t)t + 1, 5);
    const double qy = lp.noisy ? rng_uniform(t, 3, lp.seed) : radical_inverse((uint32_t)t + 1, 7);
    const double qz = lp.noisy ? rng_uniform(t, 4, lp.seed) : radical_inverse((uint32_t)t + 1, 11);
    double drift_x = lp.vd, width_x = lp.vth;
    if (lp.bump_on_tail) {
        const int n_beam = (int)(lp.beam_frac * ppc + 0.5);
        const bool is_beam = s < n_beam;
        drift_x = is_beam ? lp.beam_vd : 0.0;
        width_x = is_beam ? lp.beam_vth : lp.vth;
    } else if (lp.two_stream) {
        drift_x = (s & 1) ? -lp.vd : lp.vd;
    }
    const double r2 = 1.41421356237309515;
    const int64_t slot = (int64_t)s * ((int64_t)lp.nx * lp.ny) + c;
    x[slot] = xv;
    y[slot] = yv;
    ux[slot] = (float)(drift_x + r2 * width_x * erfinv(2.0 * qx - 1.0));
    uy[slot] = (float)(r2 * lp.vth * erfinv(2.0 * qy - 1.0));
    uz[slot] = (float)(r2 * lp.vth * erfinv(2.0 * qz - 1.0));
}

__global__ void fill_kernel(float* __restrict__ a, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

int key_bits(uint32_t nkeys) {
    int b = 0;
    while (b < 32 && (nkeys - 1) >> b) ++b;
    return b == 0 ? 1 : b;
}

}  // namespace

// ---------------------------------------------------------------------------
// host API
// ---------------------------------------------------------------------------

void compute_tile_keys(aw_ctx& c, const aw_particles& p, uint32_t* key, bool cell_key,
                       hipStream_t s) {
    if (p.n == 0) return;
    tile_key_kernel<<<blocks_for(p.n), kThreads, 0, s>>>(p.x, p.y, (int)p.n, c.g.nx, c.g.ny, c.tx,
                                                         c.ty, c.ntx, cell_key ? 1 : 0, key);
    AW_LAUNCH_CHECK();
}

// Keys must already be in c.sort.keyA[0..n).  Result: perm (n), off (nkeys+1).
void stable_sort_keys(aw_ctx& c, int64_t n, uint32_t nkeys, int32_t* perm, int32_t* off,
                      hipStream_t s) {
    AW_REQUIRE(n < (int64_t)INT32_MAX, "particle count must fit int32");
    const int nn = (int)n;
    const int bits = key_bits(nkeys);
    const int passes = (bits + 7) / 8;
    const int nchunks = std::max(1, (nn + kRsChunk - 1) / kRsChunk);
    c.sort.keyB.reserve((size_t)std::max(nn, 1));
    c.sort.valA.reserve((size_t)std::max(nn, 1));
    c.sort.valB.reserve((size_t)std::max(nn, 1));
    c.sort.table.reserve((size_t)256 * nchunks);
    uint32_t* kin = c.sort.keyA.get();
    uint32_t* kout = c.sort.keyB.get();
    int32_t* vin = nullptr;  // first pass ranks the implicit index
    // Values ping-pong through valA/valB; the final pass lands in `perm`.
    for (int pass = 0; pass < passes && nn > 0; ++pass) {
        const bool last = pass == passes - 1;
        int32_t* vdst = last ? perm : (vin == c.sort.valA.get() ? c.sort.valB.get()
                                                                 : c.sort.valA.get());
        rs_hist_kernel<<<nchunks, kThreads, 0, s>>>(kin, nn, pass * 8, nchunks, c.sort.table.get());
        AW_LAUNCH_CHECK();
        exclusive_scan(c, c.sort.table.get(), (int64_t)256 * nchunks, s);
        rs_scatter_kernel<<<nchunks, kThreads, 0, s>>>(kin, vin, kout, vdst, nn, pass * 8,
                                                       nchunks, c.sort.table.get());
        AW_LAUNCH_CHECK();
        std::swap(kin, kout);
        vin = vdst;
    }
    if (off) {
        key_offsets_kernel<<<blocks_for((int64_t)nn + 1), kThreads, 0, s>>>(kin, nn, nkeys, off);
        AW_LAUNCH_CHECK();
    }
}

// Stable tile sort src -> dst in one pass (see ts_* kernels).  Returns false
// when the per-tile LDS tables do not fit (caller uses the radix path).
bool tile_sort_onepass(aw_ctx& c, const aw_particles& src, const aw_particles& dst, int32_t* off,
                       int32_t* perm, hipStream_t s) {
    const size_t lds = (size_t)c.ntiles * 4;
    if (lds > c.lds_per_block) return false;
    AW_REQUIRE(src.n < (int64_t)INT32_MAX, "particle count must fit int32");
    const int n = (int)src.n;
    if (n == 0) {
        AW_HIP(hipMemsetAsync(off, 0, ((size_t)c.ntiles + 1) * sizeof(int32_t), s));
        return true;
    }
    const int rounds = ts_rounds(n, c.ntiles);
    const int nchunks = (int)(((int64_t)n + kTsWave * rounds - 1) / (kTsWave * rounds));
    int tbits = 0;
    while ((1 << tbits) < c.ntiles) ++tbits;
    c.sort.table.reserve((size_t)c.ntiles * nchunks);
    AW_REQUIRE(c.g.nx < (1 << 24) && c.g.ny < (1 << 24), "grid sides must be below 2^24");
    const TileMap tm{c.g.nx, c.g.ny, c.tx, c.ty, c.ntx, 1.0f / (float)c.tx, 1.0f / (float)c.ty};
    ts_hist_kernel<<<nchunks, kTsHistThreads, lds, s>>>(src.x, src.y, n, tm, c.ntiles, rounds,
                                                        c.sort.table.get());
    AW_LAUNCH_CHECK();
    // offsets of the chunk-major table (see ts_colsum_kernel)
    const int nblk = (nchunks + kTsColBlock - 1) / kTsColBlock;
    const size_t nt = (size_t)c.ntiles;
    c.sort.colsum.reserve((size_t)nblk * nt + nt + 1);
    uint32_t* bsum = c.sort.colsum.get();
    uint32_t* tbase = bsum + (size_t)nblk * nt;  // tile totals, scanned into tile bases
    const dim3 cg(blocks_for((int64_t)c.ntiles), nblk);
    ts_colsum_kernel<<<cg, kThreads, 0, s>>>(c.sort.table.get(), nchunks, c.ntiles, bsum);
    AW_LAUNCH_CHECK();
    ts_colscan_kernel<<<blocks_for((int64_t)c.ntiles), kThreads, 0, s>>>(bsum, nblk, c.ntiles,
                                                                         tbase);
    AW_LAUNCH_CHECK();
    exclusive_scan(c, tbase, (int64_t)c.ntiles, s);
    ts_colapply_kernel<<<cg, kThreads, 0, s>>>(c.sort.table.get(), nchunks, c.ntiles, bsum,
                                               tbase);
    AW_LAUNCH_CHECK();
    ts_offsets_kernel<<<blocks_for((int64_t)c.ntiles + 1), kThreads, 0, s>>>(tbase, c.ntiles, n,
                                                                           off);
    AW_LAUNCH_CHECK();
    TsMove m{src.x, src.y, src.ux, src.uy, src.uz, src.w,
             dst.x, dst.y, dst.ux, dst.uy, dst.uz, dst.w, perm};
    if (dst.x)
        ts_scatter_kernel<true><<<nchunks, kTsWave, lds, s>>>(m, n, tm, c.ntiles, rounds, tbits,
                                                              c.sort.table.get());
    else
        ts_scatter_kernel<false><<<nchunks, kTsWave, lds, s>>>(m, n, tm, c.ntiles, rounds,
                                                               tbits, c.sort.table.get());
    AW_LAUNCH_CHECK();
    return true;
}

void bin_tiles(aw_ctx& c, const aw_particles& p, int32_t* perm, int32_t* off, hipStream_t s) {
    AW_REQUIRE(perm && off, "bin_tiles needs perm and off");
    aw_particles none{};
    if (tile_sort_onepass(c, p, none, off, perm, s)) return;
    c.sort.keyA.reserve((size_t)std::max<int64_t>(1, p.n));
    compute_tile_keys(c, p, c.sort.keyA.get(), false, s);
    stable_sort_keys(c, p.n, (uint32_t)c.ntiles, perm, off, s);
}

void permute_particles(const aw_particles& src, const aw_particles& dst, const int32_t* perm,
                       hipStream_t s) {
    if (src.n == 0) return;
    permute_kernel<<<blocks_for(src.n), kThreads, 0, s>>>(src.x, src.y, src.ux, src.uy, src.uz,
                                                          src.w, dst.x, dst.y, dst.ux, dst.uy,
                                                          dst.uz, dst.w, perm, (int)src.n);
    AW_LAUNCH_CHECK();
}

void cell_index(aw_ctx& c, const aw_particles& p, int32_t* cell, hipStream_t s) {
    if (p.n == 0) return;
    cell_index_kernel<<<blocks_for(p.n), kThreads, 0, s>>>(p.x, p.y, (int)p.n, c.g.nx, c.g.ny,
                                                           cell);
    AW_LAUNCH_CHECK();
}

void migrate(aw_ctx& c, const aw_particles& p, int32_t* cell, hipStream_t s) {
    if (p.n == 0) return;
    migrate_kernel<<<blocks_for(p.n), kThreads, 0, s>>>(p.x, p.y, (int)p.n, c.g.nx, c.g.ny, cell);
    AW_LAUNCH_CHECK();
}

void load_species(aw_ctx& c, const aw_species* sp, int nsp, const aw_run_params& rp,
                  const double* mode_amp, int n_modes, const aw_particles& p,
                  double* weights_out, hipStream_t s) {
    const int64_t nc = ncell(c);
    int64_t total = 0;
    for (int k = 0; k < nsp; ++k) {
        AW_REQUIRE(sp[k].ppc > 0, "species ppc must be positive");
        total += (int64_t)sp[k].ppc * nc;
    }
    AW_REQUIRE(total == p.n, "aw_load_species: p->n must equal sum(ppc)*nx*ny");
    AW_REQUIRE(n_modes >= 0 && n_modes <= 32, "at most 32 perturbation modes");
    int64_t base = 0;
    for (int k = 0; k < nsp; ++k) {
        SpeciesLoad L{};
        L.ppc = sp[k].ppc;
        L.index = k;
        L.base = base;
        L.count = (int64_t)sp[k].ppc * nc;
        for (int d = 0; d < 3; ++d) {
            L.uth[d] = sp[k].uth[d];
            L.ufl[d] = sp[k].ufl[d];
        }
        L.noisy = rp.noisy_load ? 1 : 0;
        L.seed = rp.rng_seed;
        L.nmodes = n_modes;
        for (int m = 0; m < n_modes; ++m) L.mode_amp_cells[m] = mode_amp[m] / c.g.dx;
        const double wgt = sp[k].density * c.g.dx * c.g.dy / sp[k].ppc;
        if (weights_out) weights_out[k] = wgt;
        species_load_kernel<<<blocks_for(L.count), kThreads, 0, s>>>(p.x, p.y, p.ux, p.uy, p.uz,
                                                                      c.g.nx, c.g.ny, L);
        AW_LAUNCH_CHECK();
        if (p.w) {
            fill_kernel<<<blocks_for(L.count), kThreads, 0, s>>>(p.w + base, L.count, (float)wgt);
            AW_LAUNCH_CHECK();
        }
        base += L.count;
    }
}

void load_legacy(aw_ctx& c, const aw_run_params& rp, const aw_particles& p, hipStream_t s) {
    AW_REQUIRE(rp.ppc > 0, "rp.ppc must be positive");
    AW_REQUIRE(p.n == (int64_t)rp.ppc * ncell(c), "p->n must equal ppc*nx*ny");
    LegacyLoad L{};
    L.ppc = rp.ppc;
    L.nx = c.g.nx;
    L.ny = c.g.ny;
    L.vth = rp.vth;
    L.vd = rp.vd;
    L.perturb_amp = rp.perturb_amp;
    L.perturb_kx = rp.perturb_kx;
    L.two_stream = rp.two_stream;
    L.bump_on_tail = rp.bump_on_tail;
    L.noisy = rp.noisy_load;
    L.beam_frac = rp.beam_frac;
    L.beam_vd = rp.beam_vd;
    L.beam_vth = rp.beam_vth;
    L.seed = rp.rng_seed;
    legacy_load_kernel<<<blocks_for(p.n), kThreads, 0, s>>>(p.x, p.y, p.ux, p.uy, p.uz, p.n, L);
    AW_LAUNCH_CHECK();
    if (p.w) {
        fill_kernel<<<blocks_for(p.n), kThreads, 0, s>>>(p.w, p.n, (float)rp.weight);
        AW_LAUNCH_CHECK();
    }
}

}  // namespace aw
