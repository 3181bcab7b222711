// aw_internal.hpp -- shared host/device internals of libarcwarden_hip.
//
// Host side: error plumbing (AwError -> aw_status at the C ABI), RAII device
// buffers, the aw_ctx definition, launch helpers.  Device side: the inline
// arithmetic every kernel shares (wrap, CIC, Boris, fixed-point quantiser).
// The FP32 sequences here are the ones the CPU oracle restates with explicit
// fmaf; the library is compiled with -ffp-contract=off so nothing else fuses.
#pragma once

#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "aw.h"

namespace aw {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

struct AwError : std::runtime_error {
    aw_status code;
    AwError(aw_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);
const std::string& last_error();

[[noreturn]] inline void fail(aw_status code, const std::string& what, const char* file,
                              int line) {
    throw AwError(code, std::string(file) + ":" + std::to_string(line) + ": " + what);
}

#define AW_REQUIRE(cond, msg)                                                          \
    do {                                                                               \
        if (!(cond)) ::aw::fail(AW_ERR_INVALID, std::string(msg), __FILE__, __LINE__); \
    } while (0)

#define AW_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            ::aw::fail(e_ == hipErrorOutOfMemory ? AW_ERR_NOMEM : AW_ERR_HIP,          \
                       std::string(#expr) + ": " + hipGetErrorName(e_) + " (" +        \
                           hipGetErrorString(e_) + ")",                                \
                       __FILE__, __LINE__);                                            \
    } while (0)

#define AW_FFT(expr)                                                                   \
    do {                                                                               \
        rocfft_status e_ = (expr);                                                     \
        if (e_ != rocfft_status_success)                                               \
            ::aw::fail(AW_ERR_FFT, std::string(#expr) + ": rocfft status " +           \
                                       std::to_string((int)e_),                        \
                       __FILE__, __LINE__);                                            \
    } while (0)

#define AW_LAUNCH_CHECK() AW_HIP(hipGetLastError())

// Runs body, converting exceptions to a status + thread-local message.
template <class F>
aw_status guard(const char* fn, F&& body) {
    try {
        body();
        return AW_OK;
    } catch (const AwError& e) {
        set_last_error(std::string(fn) + ": " + e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error(std::string(fn) + ": host allocation failed");
        return AW_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(std::string(fn) + ": " + e.what());
        return AW_ERR_INVALID;
    }
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------

template <class T>
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { resize(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p_ = std::exchange(o.p_, nullptr);
            n_ = std::exchange(o.n_, 0);
        }
        return *this;
    }
    // Grows to at least n elements (contents not preserved).
    void reserve(size_t n) {
        if (n > n_) resize(n);
    }
    void resize(size_t n) {
        release();
        if (n) AW_HIP(hipMalloc(&p_, n * sizeof(T)));
        n_ = n;
    }
    void release() noexcept {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void zero(hipStream_t s) {
        if (n_) AW_HIP(hipMemsetAsync(p_, 0, n_ * sizeof(T), s));
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    void swap(DevBuf& o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
    }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// ---------------------------------------------------------------------------
// uniform-B0 Boris constants (host-computed, identical formula to the oracle)
// ---------------------------------------------------------------------------

struct BorisConsts {
    float qmh, t[3], s[3];
};

inline BorisConsts boris_make(double qm, double dt_eff, const float B0[3], bool has_b0) {
    BorisConsts b{};
    const double qmh = qm * 0.5 * dt_eff;
    double t[3], t2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        t[k] = has_b0 ? qmh * (double)B0[k] : 0.0;
        t2 += t[k] * t[k];
    }
    const double sf = 2.0 / (1.0 + t2);
    b.qmh = (float)qmh;
    for (int k = 0; k < 3; ++k) {
        b.t[k] = (float)t[k];
        b.s[k] = (float)(sf * t[k]);
    }
    return b;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

// Field-combination pointers: E = (E_L + E_T) + pump; ET[0] null means E_L + pump.
struct FormE {
    float* E[3];
    const float* EL[2];
    const float* ET[3];
    long long sE = 0, sL = 0, sT = 0;  // per-run strides of a batch (E, E_L, E_T)
};

// A uniform batch of Darwin runs (same grid, particle count and weights)
// stored run after run: the particle kernels take blockIdx.y as the run and
// offset each per-run pointer by run * stride.  runs = 1 is a single run.
struct EmBatch {
    int runs = 1;
    long long pstride = 0;  // particles
    long long estride = 0;  // E block (3 grids)
    long long bstride = 0;  // B block
    long long astride = 0;  // moment accumulators
    long long kstride = 0;  // KE partials
    int ostride = 0;        // tile offsets
};

// scratch of the standalone Darwin operators (aw_deposit_rho_j_amu, ...)
struct EmScratch {
    DevBuf<unsigned long long> acc;  // up to 8 fixed-point grids
    DevBuf<unsigned int> bound;      // max|w|, max|u| (float bits)
    DevBuf<float> cq, cdcu, fpart;   // moment / dcu coefficients, field maxima
    DevBuf<double> inv, dv;          // inverse scales (11), dcu bound inputs (2)
};

struct SortScratch {
    DevBuf<uint32_t> keyA, keyB;
    DevBuf<int32_t> valA, valB;
    DevBuf<uint32_t> table, sums;
    DevBuf<uint32_t> colsum;  // tile sort: per (chunk block, tile) sums, then tile bases
    DevBuf<int32_t> bin_off, bin_perm;  // tile bins of the last binned deposit
};

}  // namespace aw

struct aw_ctx {
    aw_grid g{};
    int tx = 16, ty = 16, ntx = 0, nty = 0, ntiles = 0;
    int apron = 2;
    int fix_log2 = 40;
    // the run's mean charge density in fixed point, subtracted from the
    // R2C input (the solve drops k = 0; 0: nothing subtracted)
    long long rho_base = 0;
    int device = 0, num_cu = 0;
    size_t lds_per_block = 0;
    // rocFFT
    rocfft_plan r2c = nullptr, c2r = nullptr, c2r2 = nullptr;
    rocfft_plan r2c2 = nullptr, r2c3 = nullptr, c2r3 = nullptr;
    rocfft_plan r2c4 = nullptr, r2c7 = nullptr, c2r5 = nullptr;  // Darwin batches
    rocfft_plan r2cB = nullptr, c2rB = nullptr;                  // ensemble batches
    int fft_batch = 0;
    rocfft_plan r2c4B = nullptr, c2r5B = nullptr, r2c7B = nullptr, c2r3B = nullptr;
    int em_batch = 0;                                            // Darwin batch size
    rocfft_execution_info info = nullptr;
    aw::DevBuf<unsigned char> fft_work;
    // workspace
    aw::DevBuf<float> ktmp;                  // complex scratch, 2 floats per mode
    aw::DevBuf<float> rin;                   // R2C input copy (rocFFT may clobber it)
    aw::DevBuf<unsigned long long> acc;      // fixed-point grid accumulator
    aw::DevBuf<double> red;                  // reduction partials
    aw::DevBuf<unsigned long long> jacc;     // Esirkepov Jx|Jy|Jz fixed-point accumulators
    aw::DevBuf<unsigned int> jbound;         // Esirkepov scale bounds (max|w|, max|w uz|)
    aw::DevBuf<float> jpart;                 // their per-workgroup partial maxima
    aw::SortScratch sort;
    aw::EmScratch em;
    ~aw_ctx();
};

namespace aw {

inline int64_t ncell(const aw_ctx& c) { return (int64_t)c.g.nx * c.g.ny; }
inline int64_t nkmodes(const aw_ctx& c) { return (int64_t)(c.g.nx / 2 + 1) * c.g.ny; }
inline hipStream_t S(aw_stream s) { return (hipStream_t)s; }

void fft_exec(aw_ctx& c, rocfft_plan plan, void* in, void* out, hipStream_t s);

// Sort / key machinery (particles.hip)
void compute_tile_keys(aw_ctx& c, const aw_particles& p, uint32_t* key, bool cell_key,
                       hipStream_t s);
// Stable LSD radix sort of (key, index) pairs; perm receives the permutation,
// off (nkeys+1) the key offsets.  keys are consumed.
void stable_sort_keys(aw_ctx& c, int64_t n, uint32_t nkeys, int32_t* perm, int32_t* off,
                      hipStream_t s);
void permute_particles(const aw_particles& src, const aw_particles& dst, const int32_t* perm,
                       hipStream_t s);
bool tile_sort_onepass(aw_ctx& c, const aw_particles& src, const aw_particles& dst, int32_t* off,
                       int32_t* perm, hipStream_t s);
// Stable bins by tile without moving the particles: perm[pos] = source index
// in tile order, off[ntiles + 1] (one-pass counting sort when the per-tile
// counters fit LDS, else the radix sort of tile keys).
void bin_tiles(aw_ctx& c, const aw_particles& p, int32_t* perm, int32_t* off, hipStream_t s);

// Esirkepov current (esirkepov.hip).  Optional inputs: fixed scale bounds the
// caller guarantees, in place of a bound pass over the particles; the tile
// offsets of particles the caller keeps in tile order (ntiles + 1, from an
// earlier stable tile sort of the same n particles); outputs for the wrapped
// new positions (may alias p.x, p.y).
struct EskExtra {
    // fixed bounds: (max |w|, max |w uz|) with per-particle deposit weights,
    // (unused, max |uz|) with w_uniform
    bool fixed = false;
    int fixed_per_w = 0;
    float fixed_bound[2] = {0.f, 0.f};
    const int32_t* slice_off = nullptr;
    float *xw = nullptr, *yw = nullptr;
    // leave the accumulated current in c.jacc for the caller to convert
    bool no_convert = false;
};
void deposit_esirkepov(aw_ctx& c, const float* x0, const float* y0, const aw_particles& p,
                       double qm, double dt, int policy, float* jx, float* jy, float* jz,
                       hipStream_t s, const EskExtra& ex = EskExtra{});

// Argument checks of the C ABI (capi.hip)
void check_grid(const aw_grid* g);
void check_particles(const aw_particles* p);

// Deposit helpers (deposit.hip)
void deposit_rho_acc(const aw_grid& g, const aw_particles& p, double qm, int policy,
                     int scale_log2, unsigned long long* acc, int num_cu, size_t lds_per_block,
                     hipStream_t s);
// counter: ncounter counters cleared by the same launch (stray counts);
// out2: a copy less the integer baseline (base_arr per cell, else base_val)
void fix_to_float_add(const unsigned long long* acc, float* out, int64_t n, int scale_log2,
                      bool zero_acc, hipStream_t s, bool assign = false,
                      unsigned long long* counter = nullptr, float* out2 = nullptr,
                      int ncounter = 1, const long long* base_arr = nullptr,
                      long long base_val = 0);
// qm * density in fixed point at 2^scale_log2: the mean charge density of a
// run whose species densities sum to `density`, removed from the R2C input
long long rho_baseline(double qm, double density, int scale_log2);
// acc == nullptr: the context's accumulator
void deposit_rho_fixed(aw_ctx& c, const aw_particles& p, double qm, int policy,
                       const int32_t* sorted_off, hipStream_t s,
                       unsigned long long* acc = nullptr);
int tile_bpt(const aw_ctx& c, int64_t n);

// Spectral (spectral.hip)
void solve_es_contig(aw_ctx& c, const float* rho, float* E2, double eps0, hipStream_t s);
void enable_batch_fft(aw_ctx& c, int batch);
void solve_es_batch(aw_ctx& c, int batch, float* rho_in, float* E2B, double eps0,
                    hipStream_t s);

// Particles, push (particles.hip, push.hip)
void load_species(aw_ctx& c, const aw_species* sp, int nsp, const aw_run_params& rp,
                  const double* mode_amp, int n_modes, const aw_particles& p,
                  double* weights_out, hipStream_t s);
void push_es(aw_ctx& c, const aw_particles& p, const float* Ex, const float* Ey,
             const aw_run_params& rp, bool has_b0, bool half_back, hipStream_t s);
bool uz_evolves(const aw_run_params& rp, bool has_b0);
size_t run_slot_bytes();
int fill_run_slot(void* slot, aw_ctx& c, const aw_particles& p, const int32_t* off,
                  const float* Ex, const float* Ey, const aw_run_params& rp, bool has_b0,
                  unsigned long long* acc, unsigned long long* strays, double* ke_part, int bpt,
                  int wg_begin);
size_t push_lds_bytes(const aw_ctx& c);
// wg_run[w] = the run (slot index) of workgroup w of the launch
void push_deposit_multi(const void* slots, const int32_t* wg_run, int total_wg, size_t lds,
                        bool ke, bool uz, hipStream_t s);

// Contexts (capi.hip)
aw_ctx* make_ctx(const aw_grid* g, int tile_x, int tile_y);
int fixed_scale_for(const std::vector<aw_species>& sp, const aw_run_params& rp, bool legacy);
int rho_scale_guarded(const aw_grid& g, double qm, double wmax, int64_t n, int S);
// Added to a step's stray count when a Darwin moment left the fixed-point
// range its static scale assumes (the count itself stays below 2^40).
constexpr unsigned long long kRangeFlag = 1ull << 40;
void autotune_run(const aw_grid& g, const aw_run_params& rp, const aw_species* sp, int nsp,
                  bool has_b0, int& tile, int& bpt);
void set_ctx_tiles(aw_ctx& c, int tile_x, int tile_y);
void sort_tiles(aw_ctx& c, const aw_particles& src, const aw_particles& dst, int32_t* perm,
                int32_t* off, hipStream_t s);

}  // namespace aw

// ===========================================================================
// device helpers
// ===========================================================================

namespace aw::dev {

__device__ __forceinline__ int wrapi(int i, int n) {
    if (i >= n) i -= n;
    if (i < 0) i += n;
    return i;
}

__device__ __forceinline__ int wrapfar(int i, int n) {
    // within one period of the box: no integer division
    if (i >= 0 && i < n) return i;
    if (i < 0 && i >= -n) return i + n;
    if (i >= n && i - n < n) return i - n;
    i %= n;
    if (i < 0) i += n;
    return i;
}

// Wrapped CIC pair (i, i+1) valid for ANY int i (a NaN / inf position gives
// an arbitrary but in-range index instead of an out-of-bounds address).
__device__ __forceinline__ void wrap_pair(int i, int n, int& a, int& b) {
    a = wrapfar(i, n);
    b = (a + 1 == n) ? 0 : a + 1;
}

// Max over the workgroup of (a, b), for workgroups of up to 1024 threads in
// which every thread calls it; thread 0 receives the result.
__device__ __forceinline__ void block_max2(float& a, float& b) {
    __shared__ float sa[16], sb[16];
    for (int off = 32; off > 0; off >>= 1) {
        a = fmaxf(a, __shfl_xor(a, off, 64));
        b = fmaxf(b, __shfl_xor(b, off, 64));
    }
    const int wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) {
        sa[wave] = a;
        sb[wave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < nw; ++k) {
            a = fmaxf(a, sa[k]);
            b = fmaxf(b, sb[k]);
        }
    }
}

// Tile-local LDS coordinate test: l in [0, lim] as an unsigned compare, so a
// garbage negative index fails and takes the (safe) global path.
__device__ __forceinline__ bool in_tile(int l, int lim) {
    return (unsigned)l <= (unsigned)lim;
}

// Periodic wrap with the reference migrate semantics; identical to the
// oracle's wrap_coord (one-period fast paths are exact rewrites of fmodf).
__device__ __forceinline__ float wrap_coord(float v, float n) {
    if (v >= n) {
        v = (v < 2.0f * n) ? v - n : fmodf(v, n);
    } else if (v < 0.0f) {
        if (!(v > -n)) v = fmodf(v, n);
        if (v < 0.0f) v += n;
    }
    if (v >= n) v = 0.0f;
    return v;
}

__device__ __forceinline__ int cell_of(float x, float y, int nx, int ny) {
    int ci = (int)floorf(x), cj = (int)floorf(y);
    if (ci >= nx) ci = nx - 1;
    if (cj >= ny) cj = ny - 1;
    if (ci < 0) ci = 0;
    if (cj < 0) cj = 0;
    return cj * nx + ci;
}

// The loaders' deterministic draws (include/pic/particles.hpp:190-211): the
// radical inverse of i in a base, and the hashed uniform in (0, 1) of draw
// index t on a stream.
__device__ __forceinline__ double radical_inverse(uint32_t i, uint32_t base) {
    const double inv = 1.0 / (double)base;
    double f = inv, r = 0.0;
    while (i > 0) {
        r += (double)(i % base) * f;
        i /= base;
        f *= inv;
    }
    return r;
}

__device__ __forceinline__ uint32_t hash_u32(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352dU;
    v ^= v >> 15;
    v *= 0x846ca68bU;
    v ^= v >> 16;
    return v;
}

__device__ __forceinline__ double rng_uniform(int64_t t, uint32_t stream, uint64_t seed) {
    const uint32_t h = hash_u32((uint32_t)t * 0x9e3779b9U + stream * 0x85ebca6bU +
                                (uint32_t)seed);
    return ((double)h + 0.5) * (1.0 / 4294967296.0);
}

struct CicF {
    int i0, j0;
    float w0, w1, w2, w3;
};

__device__ __forceinline__ CicF cic_f(float x, float y) {
    CicF s;
    s.i0 = (int)floorf(x);
    s.j0 = (int)floorf(y);
    const float fx = x - (float)s.i0, fy = y - (float)s.j0;
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    s.w0 = gx * gy;
    s.w1 = fx * gy;
    s.w2 = gx * fy;
    s.w3 = fx * fy;
    return s;
}

// gather as the oracle's fma chain: w3*f3 + (w2*f2 + (w1*f1 + w0*f0))
__device__ __forceinline__ float gather4(const CicF& s, float f0, float f1, float f2, float f3) {
    return fmaf(s.w3, f3, fmaf(s.w2, f2, fmaf(s.w1, f1, s.w0 * f0)));
}

// boris_velocity_update, ES form (no Ez kick)
__device__ __forceinline__ void boris_es(const BorisConsts& b, bool has_b0, float ex, float ey,
                                         float& ux, float& uy, float& uz) {
    ux = fmaf(b.qmh, ex, ux);
    uy = fmaf(b.qmh, ey, uy);
    if (has_b0) {
        const float upx = fmaf(uy, b.t[2], fmaf(-uz, b.t[1], ux));
        const float upy = fmaf(uz, b.t[0], fmaf(-ux, b.t[2], uy));
        const float upz = fmaf(ux, b.t[1], fmaf(-uy, b.t[0], uz));
        ux = fmaf(upy, b.s[2], fmaf(-upz, b.s[1], ux));
        uy = fmaf(upz, b.s[0], fmaf(-upx, b.s[2], uy));
        uz = fmaf(upx, b.s[1], fmaf(-upy, b.s[0], uz));
    }
    ux = fmaf(b.qmh, ex, ux);
    uy = fmaf(b.qmh, ey, uy);
}

// boris_es for B0 along z (t[0] = t[1] = s[0] = s[1] = 0): every term the
// general form takes from uz is an exact zero there, so this gives the same
// ux, uy without reading uz, and uz is left unchanged.
__device__ __forceinline__ void boris_es_bz(const BorisConsts& b, bool has_b0, float ex, float ey,
                                            float& ux, float& uy) {
    ux = fmaf(b.qmh, ex, ux);
    uy = fmaf(b.qmh, ey, uy);
    if (has_b0) {
        const float upx = fmaf(uy, b.t[2], ux);
        const float upy = fmaf(-ux, b.t[2], uy);
        ux = fmaf(upy, b.s[2], ux);
        uy = fmaf(-upx, b.s[2], uy);
    }
    ux = fmaf(b.qmh, ex, ux);
    uy = fmaf(b.qmh, ey, uy);
}

// Full Boris with a per-particle (gathered) B: boris_update_full
// (include/pic/pusher.hpp:103-109) with explicit fmaf; the oracle restates it.
__device__ __forceinline__ void boris_full(float qmh, float ex, float ey, float ez, float bx,
                                           float by, float bz, float& ux, float& uy, float& uz) {
    ux = fmaf(qmh, ex, ux);
    uy = fmaf(qmh, ey, uy);
    uz = fmaf(qmh, ez, uz);
    const float tx = qmh * bx, ty = qmh * by, tz = qmh * bz;
    const float t2 = fmaf(tz, tz, fmaf(ty, ty, tx * tx));
    const float sf = 2.0f / (1.0f + t2);
    const float sx = sf * tx, sy = sf * ty, sz = sf * tz;
    const float upx = fmaf(uy, tz, fmaf(-uz, ty, ux));
    const float upy = fmaf(uz, tx, fmaf(-ux, tz, uy));
    const float upz = fmaf(ux, ty, fmaf(-uy, tx, uz));
    ux = fmaf(upy, sz, fmaf(-upz, sy, ux));
    uy = fmaf(upz, sx, fmaf(-upx, sz, uy));
    uz = fmaf(upx, sy, fmaf(-upy, sx, uz));
    ux = fmaf(qmh, ex, ux);
    uy = fmaf(qmh, ey, uy);
    uz = fmaf(qmh, ez, uz);
}

// rint(v) (ties to even) as int64 for |v| < 2^51, with one FP64 add: v +
// 1.5*2^52 lies in [2^52, 2^53), where the spacing is 1, so the sum is rounded
// to the nearest integer and its bit pattern exceeds that of 1.5*2^52 by
// exactly rint(v).  Callers keep |v| < 2^51 by clamping a factor once (the
// clamp_* helpers) rather than each contribution.
__device__ __forceinline__ long long fix64_raw(double v) {
    constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52
    return __double_as_longlong(v + kMagic) - __double_as_longlong(kMagic);
}

// The fixed-point scales bound every contribution by 2^48 (J and amu assume
// |u| <= 2c, so a particle would need |u| > 16c to reach 2^51).  Anything
// larger, or NaN, saturates instead of wrapping.
__device__ __forceinline__ double clamp_fix_f64(double v) {
    constexpr double kLim = 2251799813685247.0;  // 2^51 - 1
    return fmin(fmax(v, -kLim), kLim);
}
__device__ __forceinline__ float clamp_fix_f32(float v) {
    constexpr float kLim = 2251799679467520.0f;  // 2^51 - 2^27, the largest float below 2^51
    return fminf(fmaxf(v, -kLim), kLim);
}

__device__ __forceinline__ long long fix64(double v) { return fix64_raw(clamp_fix_f64(v)); }


// Fixed-point contribution rint(v * w) with one FP64 fma: the product of the
// moment value v (carrying the quantity's 2^S scale) and the CIC weight w is
// formed exactly and rounded once, as the magic-number add of fix64_raw.  v
// must come through clamp_fix_f64 and |w| <= 1, so |v w| < 2^51.
__device__ __forceinline__ unsigned long long fix_mul(double v, double w) {
    constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52
    return (unsigned long long)(__double_as_longlong(fma(v, w, kMagic)) -
                                __double_as_longlong(kMagic));
}

// CIC weights in double (exact products of 24-bit factors) and the
// fixed-point quantiser: q = rint(coef * W * 2^S).
struct CicD {
    int i0, j0;
    double w0, w1, w2, w3;
};

__device__ __forceinline__ CicD cic_d(float x, float y) {
    CicD s;
    s.i0 = (int)floorf(x);
    s.j0 = (int)floorf(y);
    const double fx = (double)x - (double)s.i0, fy = (double)y - (double)s.j0;
    const double gx = 1.0 - fx, gy = 1.0 - fy;
    s.w0 = gx * gy;
    s.w1 = fx * gy;
    s.w2 = gx * fy;
    s.w3 = fx * fy;
    return s;
}

__device__ __forceinline__ unsigned long long fixq(double coef, double w, double scale) {
    return (unsigned long long)fix64(coef * w * scale);
}

// ---- Esirkepov move geometry (include/pic/yee2d.hpp:124-132,310-389) --------
// For a move (X0, Y0) -> (X1, Y1) of less than one cell (cell units; X1, Y1
// taken as the minimal image of the stored, possibly wrapped, new position)
// the 4x4 union stencil {ib-1..ib+2} x {jb-1..jb+2}, ib = floor(min(X0, X1)),
// carries the CIC node shapes S0 = S(X0), S1 = S(X1), S the node shape
// max(0, 1 - |x - node|).  Both end points lie at or beyond node ib (jb), so
// every shape at node ib - 1 (jb - 1) is zero; nodes ib..ib+2 keep the mean
// shape M = (S0 + S1)/2 and the difference D = S1 - S0.  Every one of these
// is exact in FP64: differences of float positions and small integers.
struct EskShape {
    int ib, jb;
    double mx[3], my[3], dx[3], dy[3];  // nodes ib..ib+2 (jb..jb+2)
};

__device__ __forceinline__ void esk_axis(double a, double b, int base, double m[3], double d[3]) {
    const double fa = a - (double)base, fb = b - (double)base;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double sa = fmax(0.0, 1.0 - fabs(fa - (double)k));
        const double sb = fmax(0.0, 1.0 - fabs(fb - (double)k));
        m[k] = 0.5 * (sa + sb);
        d[k] = sb - sa;
    }
}

__device__ __forceinline__ void esk_shape(double X0, double Y0, double X1, double Y1, int nx,
                                          int ny, EskShape& m) {
    if (X1 - X0 > 0.5 * nx) X1 -= nx;
    else if (X0 - X1 > 0.5 * nx) X1 += nx;
    if (Y1 - Y0 > 0.5 * ny) Y1 -= ny;
    else if (Y0 - Y1 > 0.5 * ny) Y1 += ny;
    m.ib = (int)floor(fmin(X0, X1));
    m.jb = (int)floor(fmin(Y0, Y1));
    esk_axis(X0, X1, m.ib, m.mx, m.dx);
    esk_axis(Y0, Y1, m.jb, m.my, m.dy);
}

// Calls add(q, k, b, v) for every non-zero fixed-point contribution v of
// component q (0 Jx, 1 Jy, 2 Jz) at stencil node (ib-1+k, jb-1+b), k, b in
// 1..3 (nodes ib..ib+2, jb..jb+2):
//   Jx = cJx * P_k My_b,   P_k = sum_{a<=k} Dx_a
//   Jy = cJy * Q_b Mx_k,   Q_b = sum_{c<=b} Dy_c
//   Jz = cJz * (Mx My + Dx Dy / 12)
// (cJx = -q w/(dt dy), cJy = -q w/(dt dx), cJz = q w uz/(dx dy), each times
// the fixed-point scale).  These are the Esirkepov weights: S0 + D/2 = M and
// S0 Sy0 + (Dx Sy0 + Sx0 Dy)/2 + Dx Dy/3 = Mx My + Dx Dy/12.  The full prefix
// sum at k = 3 (b = 3) is the particle's net charge change, zero, and is
// skipped: 6 + 6 + 9 contributions.  P My, Q Mx and (12 Mx My + Dx Dy)/16
// lie in [-1, 1] and are exact in FP64 wherever the float positions are
// spaced 2^-23 or wider (|x| >= 1); each contribution is then one rounding,
// the fma of fix_mul, of its exact value.
template <class Add>
__device__ __forceinline__ void esk_emit(const EskShape& m, double cJx, double cJy, double cJz,
                                         Add&& add) {
    // clamping the coefficients keeps each product inside the range of fix_mul
    cJx = clamp_fix_f64(cJx);
    cJy = clamp_fix_f64(cJy);
    const double cz = clamp_fix_f64(cJz * (16.0 / 12.0));
    const double P[2] = {m.dx[0], m.dx[0] + m.dx[1]};
    const double Q[2] = {m.dy[0], m.dy[0] + m.dy[1]};
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const unsigned long long v = fix_mul(cJx, P[k] * m.my[b]);
            if (v) add(0, k + 1, b + 1, v);
        }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const unsigned long long v = fix_mul(cJy, Q[b] * m.mx[k]);
            if (v) add(1, k + 1, b + 1, v);
        }
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const double my12 = (12.0 / 16.0) * m.my[b], dy16 = (1.0 / 16.0) * m.dy[b];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const unsigned long long v = fix_mul(cz, fma(m.dx[k], dy16, m.mx[k] * my12));
            if (v) add(2, k + 1, b + 1, v);
        }
    }
}

// The dispatcher hands consecutive workgroup ids to the 8 XCDs in turn.  This
// bijection of [0, n) gives XCD x a contiguous range of logical workgroups
// (x * (n/8) + min(x, n%8) onward), so the workgroups of one tile, which
// read the same field tile and flush into the same cells, share one L2.
__device__ __forceinline__ int xcd_block(int b, int n) {
    const int x = b & 7, k = b >> 3, q = n >> 3, r = n & 7;
    return x * q + min(x, r) + k;
}

// Particle range [r0, r1) of workgroup `part` of the bpt workgroups sharing
// the tile range [beg, end): the split points are rounded up to multiples of
// 32 particles (128 B of each FP32 array), so neighbouring workgroups never
// read or write the same cache line of a particle array (only tile ends stay
// unaligned).
__device__ __forceinline__ void split_range(int beg, int end, int part, int bpt, int& r0,
                                            int& r1) {
    auto at = [&](int p) {
        if (p <= 0) return beg;
        if (p >= bpt) return end;
        int v = beg + (int)((long long)(end - beg) * p / bpt);
        v = (v + 31) & ~31;
        return v < beg ? beg : (v > end ? end : v);
    };
    r0 = at(part);
    r1 = at(part + 1);
}

}  // namespace aw::dev
