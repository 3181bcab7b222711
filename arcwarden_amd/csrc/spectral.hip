// spectral.hip -- rocFFT plans and the pointwise k-space solves.
//
// Reference: CufftPlan2D (include/pic/fft.hpp:26-93), SpectralEngine
// (include/pic/spectral.hpp:106-149), es_poisson_field_kernel
// (include/pic/solver_es.hpp:39-78), SpectralFormFactor (spectral.hpp:68-92).
// rocFFT is the only library call.  The C2R plans carry the 1/(nx*ny)
// normalisation as a plan scale factor, so there is no separate scale pass
// (the reference runs scale_inplace_kernel after every C2R, spectral.hpp:97-101).
#include "aw_internal.hpp"

#include <mutex>

namespace aw {

namespace {

constexpr int kThreads = 256;

inline int blocks_for(int64_t n) { return (int)((n + kThreads - 1) / kThreads); }

struct KGeom {
    int nx, ny, nkx;
    double dkx, dky;
};

__device__ __forceinline__ void kvec(const KGeom& k, int i, int j, double& kx, double& ky) {
    const int m = (j <= k.ny / 2) ? j : j - k.ny;
    kx = i * k.dkx;
    ky = m * k.dky;
}

__device__ __forceinline__ bool nyquist(const KGeom& k, int i, int j) {
    return ((k.nx % 2 == 0) && i == k.nx / 2) || ((k.ny % 2 == 0) && j == k.ny / 2);
}

// phi_k = rho_k/(eps0 k^2); E_k = -i k phi_k (k-grid and Green's function in
// double, as the reference); DC (green = 0) and Nyquist force zeroed.
__global__ void es_kspace_kernel(const float2* __restrict__ rk, float2* __restrict__ exk,
                                 float2* __restrict__ eyk, KGeom kg, double eps0) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kg.nkx * kg.ny) return;
    const int i = t % kg.nkx, j = t / kg.nkx;
    double kx, ky;
    kvec(kg, i, j, kx, ky);
    const double k2 = kx * kx + ky * ky;
    const double g = k2 > 0.0 ? 1.0 / (eps0 * k2) : 0.0;
    const float2 r = rk[t];
    const double pr = (double)r.x * g, pi = (double)r.y * g;
    float2 ex, ey;
    if (nyquist(kg, i, j)) {
        ex = make_float2(0.f, 0.f);
        ey = make_float2(0.f, 0.f);
    } else {
        ex = make_float2((float)(kx * pi), (float)(-kx * pr));
        ey = make_float2((float)(ky * pi), (float)(-ky * pr));
    }
    exk[t] = ex;
    eyk[t] = ey;
}

std::once_flag g_rocfft_once;

rocfft_plan make_plan(const aw_ctx& c, bool forward, int batch, double scale) {
    rocfft_plan_description desc = nullptr;
    AW_FFT(rocfft_plan_description_create(&desc));
    const size_t nx = (size_t)c.g.nx, ny = (size_t)c.g.ny, nkx = nx / 2 + 1;
    const size_t real_stride[2] = {1, nx}, cplx_stride[2] = {1, nkx};
    const size_t real_dist = nx * ny, cplx_dist = nkx * ny;
    if (forward) {
        AW_FFT(rocfft_plan_description_set_data_layout(
            desc, rocfft_array_type_real, rocfft_array_type_hermitian_interleaved, nullptr,
            nullptr, 2, real_stride, real_dist, 2, cplx_stride, cplx_dist));
    } else {
        AW_FFT(rocfft_plan_description_set_data_layout(
            desc, rocfft_array_type_hermitian_interleaved, rocfft_array_type_real, nullptr,
            nullptr, 2, cplx_stride, cplx_dist, 2, real_stride, real_dist));
    }
    if (scale != 1.0) AW_FFT(rocfft_plan_description_set_scale_factor(desc, scale));
    const size_t lengths[2] = {nx, ny};  // innermost (x) first
    rocfft_plan plan = nullptr;
    const rocfft_status st = rocfft_plan_create(
        &plan, rocfft_placement_notinplace,
        forward ? rocfft_transform_type_real_forward : rocfft_transform_type_real_inverse,
        rocfft_precision_single, 2, lengths, (size_t)batch, desc);
    rocfft_plan_description_destroy(desc);
    AW_FFT(st);
    return plan;
}

}  // namespace

void init_fft(aw_ctx& c) {
    std::call_once(g_rocfft_once, [] { rocfft_setup(); });
    const double inv = 1.0 / ((double)c.g.nx * (double)c.g.ny);
    c.r2c = make_plan(c, true, 1, 1.0);
    c.c2r = make_plan(c, false, 1, inv);
    c.c2r2 = make_plan(c, false, 2, inv);
    c.r2c3 = make_plan(c, true, 3, 1.0);
    c.c2r3 = make_plan(c, false, 3, inv);
    size_t work = 0;
    for (rocfft_plan p : {c.r2c, c.c2r, c.c2r2, c.r2c3, c.c2r3}) {
        size_t w = 0;
        AW_FFT(rocfft_plan_get_work_buffer_size(p, &w));
        work = std::max(work, w);
    }
    if (work) c.fft_work.resize(work);
    AW_FFT(rocfft_execution_info_create(&c.info));
}

void destroy_fft(aw_ctx& c) noexcept {
    for (rocfft_plan* p : {&c.r2c, &c.c2r, &c.c2r2, &c.r2c3, &c.c2r3})
        if (*p) {
            rocfft_plan_destroy(*p);
            *p = nullptr;
        }
    if (c.info) {
        rocfft_execution_info_destroy(c.info);
        c.info = nullptr;
    }
}

void fft_exec(aw_ctx& c, rocfft_plan plan, void* in, void* out, hipStream_t s) {
    AW_FFT(rocfft_execution_info_set_stream(c.info, s));
    if (c.fft_work.size())
        AW_FFT(rocfft_execution_info_set_work_buffer(c.info, c.fft_work.get(), c.fft_work.size()));
    void* ib[1] = {in};
    void* ob[1] = {out};
    AW_FFT(rocfft_execute(plan, ib, ob, c.info));
}

// rho -> (Ex, Ey) with E2 = [Ex | Ey] contiguous (2*nx*ny floats): one R2C,
// one k-space kernel, one batched (2) C2R.
void solve_es_contig(aw_ctx& c, const float* rho, float* E2, double eps0, hipStream_t s) {
    const int64_t nk = nkmodes(c);
    c.ktmp.reserve((size_t)6 * nk);
    float2* rk = reinterpret_cast<float2*>(c.ktmp.get());
    float2* ek = rk + nk;  // [Ex_k | Ey_k]
    fft_exec(c, c.r2c, const_cast<float*>(rho), rk, s);
    KGeom kg{c.g.nx, c.g.ny, c.g.nx / 2 + 1, 6.28318530717958647692 / c.g.Lx,
             6.28318530717958647692 / c.g.Ly};
    es_kspace_kernel<<<blocks_for(nk), kThreads, 0, s>>>(rk, ek, ek + nk, kg, eps0);
    AW_LAUNCH_CHECK();
    fft_exec(c, c.c2r2, ek, E2, s);
}

// General form with separate output buffers.
void solve_es(aw_ctx& c, const float* rho, float* Ex, float* Ey, double eps0, hipStream_t s) {
    const int64_t nk = nkmodes(c);
    c.ktmp.reserve((size_t)6 * nk);
    float2* rk = reinterpret_cast<float2*>(c.ktmp.get());
    float2* exk = rk + nk;
    float2* eyk = exk + nk;
    fft_exec(c, c.r2c, const_cast<float*>(rho), rk, s);
    KGeom kg{c.g.nx, c.g.ny, c.g.nx / 2 + 1, 6.28318530717958647692 / c.g.Lx,
             6.28318530717958647692 / c.g.Ly};
    es_kspace_kernel<<<blocks_for(nk), kThreads, 0, s>>>(rk, exk, eyk, kg, eps0);
    AW_LAUNCH_CHECK();
    fft_exec(c, c.c2r, exk, Ex, s);
    fft_exec(c, c.c2r, eyk, Ey, s);
}

}  // namespace aw
