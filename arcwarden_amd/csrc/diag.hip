// diag.hip -- on-device FP64 reductions for the scalar diagnostics.
//
// Reference: Diagnostics::compute (include/pic/diagnostics.hpp:58-108) copies
// every particle and grid array to the host and loops in double.  Here each
// workgroup reduces a fixed slice in double and a single-workgroup pass
// combines the partials in order, so the result is deterministic.
#include "aw_internal.hpp"

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
