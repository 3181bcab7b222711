// diag.hip -- on-device FP64 reductions for the scalar diagnostics.
//
// Reference: Diagnostics::compute (include/pic/diagnostics.hpp:58-108) copies
// every particle and grid array to the host and loops in double.  Here each
// workgroup reduces a fixed slice in double and a single-workgroup pass
// combines the partials in order, so the result is deterministic.
#include "aw_internal.hpp"

#include <map>
#include <mutex>
#include <vector>

namespace aw {

namespace {

constexpr int kThreads = 256;
constexpr int kRedBlocks = 512;
constexpr int kChunk = 1024;  // partials per workgroup of the first finalize stage

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

__device__ __forceinline__ double block_max(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kThreads / 64; ++w) t = fmax(t, sh[w]);
    __syncthreads();
    return t;
}

__global__ void __launch_bounds__(kThreads)
ke_kernel(const float* __restrict__ ux, const float* __restrict__ uy,
          const float* __restrict__ uz, const float* __restrict__ w, int64_t n,
          double* __restrict__ part) {
    __shared__ double sh[kThreads / 64];
    double acc = 0.0;
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b = (int64_t)blockIdx.x * per, e = min(n, b + per);
    for (int64_t i = b + threadIdx.x; i < e; i += kThreads) {
        const double u2 = (double)ux[i] * ux[i] + (double)uy[i] * uy[i] + (double)uz[i] * uz[i];
        acc += w ? (double)w[i] * u2 : u2;
    }
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kThreads)
grid_kernel(const float* __restrict__ Ex, const float* __restrict__ Ey,
            const float* __restrict__ rho, int64_t n, double* __restrict__ part) {
    __shared__ double sh[kThreads / 64];
    double se = 0.0, me = 0.0, sr = 0.0;
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b = (int64_t)blockIdx.x * per, e = min(n, b + per);
    for (int64_t c = b + threadIdx.x; c < e; c += kThreads) {
        const double e2 = (double)Ex[c] * Ex[c] + (double)Ey[c] * Ey[c];
        se += e2;
        me = fmax(me, e2);
        if (rho) sr += rho[c];
    }
    const double tse = block_sum(se, sh);
    const double tme = block_max(me, sh);
    const double tsr = block_sum(sr, sh);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = tse;
        part[3 * blockIdx.x + 1] = tme;
        part[3 * blockIdx.x + 2] = tsr;
    }
}

// out[c] = sum of in[c*kChunk .. (c+1)*kChunk) in a fixed order.
__global__ void __launch_bounds__(kThreads)
chunk_sum_kernel(const double* __restrict__ in, int n, double* __restrict__ out) {
    __shared__ double sh[kThreads / 64];
    const int b = blockIdx.x * kChunk, e = min(n, b + kChunk);
    double a = 0.0;
    for (int i = b + threadIdx.x; i < e; i += kThreads) a += in[i];
    const double t = block_sum(a, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// out[0] = sum(w u^2) (times w_scale), out[1] = sum |E|^2, out[2] = max |E|^2,
// out[3] = sum rho.
__global__ void __launch_bounds__(kThreads)
finalize_kernel(const double* __restrict__ kep, int nke, double w_scale,
                const double* __restrict__ gp, int ng, double* __restrict__ out) {
    __shared__ double sh[kThreads / 64];
    double a = 0.0, b = 0.0, m = 0.0, r = 0.0;
    // fixed assignment: thread t sums entries t, t+256, ... in order
    for (int i = threadIdx.x; i < nke; i += kThreads) a += kep[i];
    for (int i = threadIdx.x; i < ng; i += kThreads) {
        b += gp[3 * i];
        m = fmax(m, gp[3 * i + 1]);
        r += gp[3 * i + 2];
    }
    const double ta = block_sum(a, sh);
    const double tb = block_sum(b, sh);
    const double tm = block_max(m, sh);
    const double tr = block_sum(r, sh);
    if (threadIdx.x == 0) {
        out[0] = ta * w_scale;
        out[1] = tb;
        out[2] = tm;
        out[3] = tr;
    }
}

// ---- module diagnostics (include/pic/diag/*.hpp) ----------------------------

struct PhaseHistArgs {
    const float *x, *ux, *uz;
    int64_t n;
    int nx, nxb, nvb;
    double bx, bz, vth, vlo, vhi, dx, Lx;
};

// f(x, v_par / v_th) counts, one integer atomic per particle inside the
// window.  FP32 (the phase-video binning, include/pic/diag/module.hpp:58-70):
// x wrapped in cells, v_par = (ux bx + uz bz) / vth, ix = (int)(x / nx * nxb).
// FP64 (the snapshot binning, include/pic/diag/snapshot.hpp:55-62): x in
// physical units wrapped by Lx, the same formulas in double.  Both truncate
// toward zero, so values just below the window's low edge (in (-1, 0) bins)
// land in bin 0, as the reference's casts do.  -ffp-contract=off keeps every
// product and sum separately rounded.
template <bool F64>
__global__ void __launch_bounds__(kThreads)
phase_hist_kernel(PhaseHistArgs a, unsigned int* __restrict__ counts) {
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < a.n; i += stride) {
        int ix, iv;
        if constexpr (F64) {
            double xx = (double)a.x[i] * a.dx;
            if (xx < 0.0) xx += a.Lx;
            else if (xx >= a.Lx) xx -= a.Lx;
            const double vp = ((double)a.ux[i] * a.bx + (double)a.uz[i] * a.bz) / a.vth;
            ix = (int)(xx / a.Lx * (double)a.nxb);
            iv = (int)((vp - a.vlo) / (a.vhi - a.vlo) * (double)a.nvb);
        } else {
            const float fnx = (float)a.nx;
            float xx = a.x[i];
            if (xx < 0.f) xx += fnx;
            else if (xx >= fnx) xx -= fnx;
            const float vp = (a.ux[i] * (float)a.bx + a.uz[i] * (float)a.bz) / (float)a.vth;
            ix = (int)(xx / fnx * (float)a.nxb);
            iv = (int)((vp - (float)a.vlo) / ((float)a.vhi - (float)a.vlo) * (float)a.nvb);
        }
        if (ix >= 0 && ix < a.nxb && iv >= 0 && iv < a.nvb)
            atomicAdd(&counts[(int64_t)iv * a.nxb + ix], 1u);
    }
}

// per-block FP64 sums of (a - off)^2 over a fixed slice (deterministic)
__global__ void __launch_bounds__(kThreads)
sum_sq_kernel(const float* __restrict__ a, int64_t n, double off, double* __restrict__ part) {
    __shared__ double sh[kThreads / 64];
    double acc = 0.0;
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b = (int64_t)blockIdx.x * per, e = min(n, b + per);
    for (int64_t c = b + threadIdx.x; c < e; c += kThreads) {
        const double d = (double)a[c] - off;
        acc += d * d;
    }
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void gather_kernel(const float* __restrict__ src, const int32_t* __restrict__ idx,
                              int64_t m, float* __restrict__ dst) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) dst[k] = src[idx[k]];
}

}  // namespace

int ke_partials(const aw_particles& p, double* part, hipStream_t s) {
    const int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (p.n + 4095) / 4096));
    ke_kernel<<<nb, kThreads, 0, s>>>(p.ux, p.uy, p.uz, p.w, p.n, part);
    AW_LAUNCH_CHECK();
    return nb;
}

int grid_partials(const float* Ex, const float* Ey, const float* rho, int64_t nc, double* part,
                  hipStream_t s) {
    const int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (nc + 4095) / 4096));
    grid_kernel<<<nb, kThreads, 0, s>>>(Ex, Ey, rho, nc, part);
    AW_LAUNCH_CHECK();
    return nb;
}

// stage (may be NULL) holds ke_stage_capacity(nke) - nke doubles: with it, a
// long list of KE partials (one per fused workgroup) is first summed in
// chunks by many workgroups, so the single-workgroup pass stays short.
void finalize_diag(const double* kep, int nke, double w_scale, const double* gp, int ng,
                   double* out4, hipStream_t s, double* stage) {
    if (stage && nke > 2 * kChunk) {
        const int nb = (nke + kChunk - 1) / kChunk;
        chunk_sum_kernel<<<nb, kThreads, 0, s>>>(kep, nke, stage);
        AW_LAUNCH_CHECK();
        kep = stage;
        nke = nb;
    }
    finalize_kernel<<<1, kThreads, 0, s>>>(kep, nke, w_scale, gp, ng, out4);
    AW_LAUNCH_CHECK();
}

size_t ke_stage_capacity(size_t nke) { return nke + (nke + kChunk - 1) / kChunk + 1; }

// layout of a reduction scratch buffer: [ke partials | grid partials | out4]
int reduce_partial_capacity() { return 4 * kRedBlocks + 8; }
int reduce_grid_offset() { return kRedBlocks; }
int reduce_out_offset() { return 4 * kRedBlocks; }

}  // namespace aw

using namespace aw;

namespace {

// per-device partial-sum buffers of aw_sum_sq (allocated on first use)
std::mutex g_sumsq_mu;
std::map<int, DevBuf<double>> g_sumsq_part;

}  // namespace

extern "C" {

aw_status aw_phase_hist(const aw_particles* p, const aw_grid* g, int32_t nxb, int32_t nvb,
                        double bx, double bz, double vth, double vlo, double vhi, int32_t fp64,
                        uint32_t* counts, aw_stream s) {
    return guard("aw_phase_hist", [&] {
        check_grid(g);
        check_particles(p);
        AW_REQUIRE(counts, "counts is NULL");
        AW_REQUIRE(nxb > 0 && nvb > 0, "nxb, nvb must be positive");
        AW_REQUIRE(vhi > vlo, "vhi must exceed vlo");
        AW_REQUIRE(vth != 0.0, "vth must be non-zero");
        if (p->n == 0) return;
        PhaseHistArgs a{};
        a.x = p->x;
        a.ux = p->ux;
        a.uz = p->uz;
        a.n = p->n;
        a.nx = g->nx;
        a.nxb = nxb;
        a.nvb = nvb;
        a.bx = bx;
        a.bz = bz;
        a.vth = vth;
        a.vlo = vlo;
        a.vhi = vhi;
        a.dx = g->dx;
        a.Lx = g->Lx > 0 ? g->Lx : g->nx * g->dx;
        const int nb = (int)std::min<int64_t>(4096, (p->n + kThreads - 1) / kThreads);
        if (fp64) phase_hist_kernel<true><<<nb, kThreads, 0, S(s)>>>(a, counts);
        else phase_hist_kernel<false><<<nb, kThreads, 0, S(s)>>>(a, counts);
        AW_LAUNCH_CHECK();
    });
}

aw_status aw_sum_sq(const float* a, int64_t n, double offset, double* out, aw_stream s) {
    return guard("aw_sum_sq", [&] {
        AW_REQUIRE(out, "out is NULL");
        AW_REQUIRE(n >= 0 && (n == 0 || a), "a is NULL or n < 0");
        *out = 0.0;
        if (n == 0) return;
        int dev = 0;
        AW_HIP(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lk(g_sumsq_mu);
        DevBuf<double>& part = g_sumsq_part[dev];
        part.reserve((size_t)kRedBlocks);
        const int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (n + 4095) / 4096));
        sum_sq_kernel<<<nb, kThreads, 0, S(s)>>>(a, n, offset, part.get());
        AW_LAUNCH_CHECK();
        std::vector<double> h((size_t)nb);
        AW_HIP(hipMemcpyAsync(h.data(), part.get(), (size_t)nb * sizeof(double),
                              hipMemcpyDeviceToHost, S(s)));
        AW_HIP(hipStreamSynchronize(S(s)));
        double t = 0.0;
        for (double v : h) t += v;
        *out = t;
    });
}

aw_status aw_gather_f32(const float* src, const int32_t* idx, int64_t m, float* dst,
                        aw_stream s) {
    return guard("aw_gather_f32", [&] {
        AW_REQUIRE(m >= 0, "m must be non-negative");
        if (m == 0) return;
        AW_REQUIRE(src && idx && dst, "NULL argument");
        gather_kernel<<<(int)((m + kThreads - 1) / kThreads), kThreads, 0, S(s)>>>(src, idx, m,
                                                                                  dst);
        AW_LAUNCH_CHECK();
    });
}

}  // extern "C"
