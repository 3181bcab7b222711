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
    c.r2c2 = make_plan(c, true, 2, 1.0);
    c.r2c3 = make_plan(c, true, 3, 1.0);
    c.c2r3 = make_plan(c, false, 3, inv);
    size_t work = 0;
    for (rocfft_plan p : {c.r2c, c.c2r, c.c2r2, c.r2c2, c.r2c3, c.c2r3}) {
        size_t w = 0;
        AW_FFT(rocfft_plan_get_work_buffer_size(p, &w));
        work = std::max(work, w);
    }
    if (work) c.fft_work.resize(work);
    AW_FFT(rocfft_execution_info_create(&c.info));
}

// Darwin batches: R2C x4 [rho|Jx|Jy|Jz], R2C x7 [dcu(3)|amu(4)], C2R x5
// [ELx|ELy|Bx|By|Bz] (C2R x3 [ETx|ETy|ETz] is shared with the base set).
void enable_em_fft(aw_ctx& c) {
    if (c.r2c4) return;
    const double inv = 1.0 / ((double)c.g.nx * (double)c.g.ny);
    c.r2c4 = make_plan(c, true, 4, 1.0);
    c.r2c7 = make_plan(c, true, 7, 1.0);
    c.c2r5 = make_plan(c, false, 5, inv);
    size_t work = c.fft_work.size();
    for (rocfft_plan p : {c.r2c4, c.r2c7, c.c2r5}) {
        size_t w = 0;
        AW_FFT(rocfft_plan_get_work_buffer_size(p, &w));
        work = std::max(work, w);
    }
    if (work > c.fft_work.size()) c.fft_work.resize(work);
}

// Darwin batches of R runs: R2C x4R, C2R x5R, R2C x7R, C2R x3R.
void enable_em_batch_fft(aw_ctx& c, int R) {
    if (c.em_batch == R) return;
    AW_REQUIRE(c.em_batch == 0, "Darwin batch plans already prepared for another size");
    enable_em_fft(c);
    const double inv = 1.0 / ((double)c.g.nx * (double)c.g.ny);
    c.r2c4B = make_plan(c, true, 4 * R, 1.0);
    c.c2r5B = make_plan(c, false, 5 * R, inv);
    c.r2c7B = make_plan(c, true, 7 * R, 1.0);
    c.c2r3B = make_plan(c, false, 3 * R, inv);
    size_t work = c.fft_work.size();
    for (rocfft_plan p : {c.r2c4B, c.c2r5B, c.r2c7B, c.c2r3B}) {
        size_t w = 0;
        AW_FFT(rocfft_plan_get_work_buffer_size(p, &w));
        work = std::max(work, w);
    }
    if (work > c.fft_work.size()) c.fft_work.resize(work);
    c.em_batch = R;
}

void destroy_fft(aw_ctx& c) noexcept {
    for (rocfft_plan* p : {&c.r2c, &c.c2r, &c.c2r2, &c.r2c2, &c.r2c3, &c.c2r3, &c.r2c4, &c.r2c7,
                           &c.c2r5, &c.r2cB, &c.c2rB, &c.r2c4B, &c.c2r5B, &c.r2c7B, &c.c2r3B})
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
// rocFFT may overwrite the input of an out-of-place transform: the R2C always
// reads the context's copy (c.rin), filled here unless the caller wrote it.
static float* r2c_input(aw_ctx& c, const float* rho, hipStream_t s) {
    c.rin.reserve((size_t)ncell(c));
    if (rho != c.rin.get())
        AW_HIP(hipMemcpyAsync(c.rin.get(), rho, (size_t)ncell(c) * sizeof(float),
                              hipMemcpyDeviceToDevice, s));
    return c.rin.get();
}

// ---- ensembles: B same-size grids per transform ------------------------------------

namespace {

// es_kspace_kernel over B runs: rk [B][nk] -> ek [B][Ex_k | Ey_k]
__global__ void es_kspace_batch_kernel(const float2* __restrict__ rk, float2* __restrict__ ek,
                                       KGeom kg, double eps0, int batch) {
    const int64_t nk = (int64_t)kg.nkx * kg.ny;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk * batch) return;
    const int r = (int)(t / nk);
    const int m = (int)(t - (int64_t)r * nk);
    const int i = m % kg.nkx, j = m / kg.nkx;
    double kx, ky;
    kvec(kg, i, j, kx, ky);
    const double k2 = kx * kx + ky * ky;
    const double g = k2 > 0.0 ? 1.0 / (eps0 * k2) : 0.0;
    const float2 v = rk[t];
    const double pr = (double)v.x * g, pi = (double)v.y * g;
    float2 ex = make_float2(0.f, 0.f), ey = make_float2(0.f, 0.f);
    if (!nyquist(kg, i, j)) {
        ex = make_float2((float)(kx * pi), (float)(-kx * pr));
        ey = make_float2((float)(ky * pi), (float)(-ky * pr));
    }
    ek[(int64_t)r * 2 * nk + m] = ex;
    ek[(int64_t)r * 2 * nk + nk + m] = ey;
}

}  // namespace

void enable_batch_fft(aw_ctx& c, int batch) {
    if (c.fft_batch == batch) return;
    for (rocfft_plan* p : {&c.r2cB, &c.c2rB})
        if (*p) {
            rocfft_plan_destroy(*p);
            *p = nullptr;
        }
    const double inv = 1.0 / ((double)c.g.nx * (double)c.g.ny);
    c.r2cB = make_plan(c, true, batch, 1.0);
    c.c2rB = make_plan(c, false, 2 * batch, inv);
    size_t work = c.fft_work.size();
    for (rocfft_plan p : {c.r2cB, c.c2rB}) {
        size_t w = 0;
        AW_FFT(rocfft_plan_get_work_buffer_size(p, &w));
        work = std::max(work, w);
    }
    if (work > c.fft_work.size()) c.fft_work.resize(work);
    c.fft_batch = batch;
}

// rho_in: B grids [B][ny][nx] (rocFFT may overwrite them); E2B: [B][Ex|Ey].
void solve_es_batch(aw_ctx& c, int batch, float* rho_in, float* E2B, double eps0,
                    hipStream_t s) {
    AW_REQUIRE(c.fft_batch == batch, "batched plans not prepared for this batch");
    const int64_t nk = nkmodes(c);
    c.ktmp.reserve((size_t)6 * nk * batch);
    float2* rk = reinterpret_cast<float2*>(c.ktmp.get());
    float2* ek = rk + nk * batch;
    fft_exec(c, c.r2cB, rho_in, rk, s);
    KGeom kg{c.g.nx, c.g.ny, c.g.nx / 2 + 1, 6.28318530717958647692 / c.g.Lx,
             6.28318530717958647692 / c.g.Ly};
    es_kspace_batch_kernel<<<blocks_for(nk * batch), kThreads, 0, s>>>(rk, ek, kg, eps0, batch);
    AW_LAUNCH_CHECK();
    fft_exec(c, c.c2rB, ek, E2B, s);
}

void solve_es_contig(aw_ctx& c, const float* rho, float* E2, double eps0, hipStream_t s) {
    const int64_t nk = nkmodes(c);
    c.ktmp.reserve((size_t)6 * nk);
    float2* rk = reinterpret_cast<float2*>(c.ktmp.get());
    float2* ek = rk + nk;  // [Ex_k | Ey_k]
    fft_exec(c, c.r2c, r2c_input(c, rho, s), rk, s);
    KGeom kg{c.g.nx, c.g.ny, c.g.nx / 2 + 1, 6.28318530717958647692 / c.g.Lx,
             6.28318530717958647692 / c.g.Ly};
    es_kspace_kernel<<<blocks_for(nk), kThreads, 0, s>>>(rk, ek, ek + nk, kg, eps0);
    AW_LAUNCH_CHECK();
    fft_exec(c, c.c2r2, ek, E2, s);
}

// ---- Darwin k-space ---------------------------------------------------------------

namespace {

// E_L (es_poisson_field_kernel, solver_es.hpp:39-78) and B_k = i mu0 (k x J)/k^2
// (darwin_b_field_kernel, solver_darwin.hpp:38-75) from [rho|Jx|Jy|Jz]_k in
// one pass.  The uniform background B0 (add_b0_kernel, pump.hpp:50-54) enters
// as the DC mode of B: the C2R scale 1/(nx*ny) turns B0*nx*ny back into B0.
// stag: Jx is sampled at (i+1/2, j) and Jy at (i, j+1/2) (Esirkepov current);
// the half-cell shift moves them onto the nodes: J_k *= exp(-i k dx/2).
__global__ void darwin_elb_kernel(const float2* __restrict__ src, float2* __restrict__ out,
                                  KGeom kg, double eps0, double c, float b0x, float b0y,
                                  float b0z, int stag) {
    const int nk = kg.nkx * kg.ny;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk) return;
    src += (size_t)blockIdx.y * 4 * nk;  // run blockIdx.y of a batch
    out += (size_t)blockIdx.y * 5 * nk;
    const int i = t % kg.nkx, j = t / kg.nkx;
    double kx, ky;
    kvec(kg, i, j, kx, ky);
    const double k2 = kx * kx + ky * ky;
    const bool nyq = nyquist(kg, i, j);
    const float2 r = src[t], jz = src[3 * nk + t];
    double jxr = src[nk + t].x, jxi = src[nk + t].y, jyr = src[2 * nk + t].x,
           jyi = src[2 * nk + t].y;
    if (stag) {
        const double kPi = 3.14159265358979323846;
        const double phx = kPi * (double)i / kg.nx;
        const double phy = kPi * (double)(j <= kg.ny / 2 ? j : j - kg.ny) / kg.ny;
        const double cx = cos(phx), sx = sin(phx), cy = cos(phy), sy = sin(phy);
        const double xr = jxr * cx + jxi * sx, xi = jxi * cx - jxr * sx;
        const double yr = jyr * cy + jyi * sy, yi = jyi * cy - jyr * sy;
        jxr = xr;
        jxi = xi;
        jyr = yr;
        jyi = yi;
    }
    // longitudinal E
    const double g = k2 > 0.0 ? 1.0 / (eps0 * k2) : 0.0;
    const double pr = (double)r.x * g, pi = (double)r.y * g;
    float2 ex = make_float2(0.f, 0.f), ey = make_float2(0.f, 0.f);
    if (!nyq) {
        ex = make_float2((float)(kx * pi), (float)(-kx * pr));
        ey = make_float2((float)(ky * pi), (float)(-ky * pr));
    }
    // magnetic field
    float2 bx = make_float2(0.f, 0.f), by = bx, bz = bx;
    if (!nyq && k2 > 0.0) {
        const double G = 1.0 / (eps0 * c * c * k2);
        bx = make_float2((float)(-G * ky * jz.y), (float)(G * ky * jz.x));
        by = make_float2((float)(G * kx * jz.y), (float)(-G * kx * jz.x));
        const double cr = kx * jyr - ky * jxr, ci = kx * jyi - ky * jxi;
        bz = make_float2((float)(-G * ci), (float)(G * cr));
    } else if (t == 0) {
        const float nn = (float)kg.nx * (float)kg.ny;
        bx = make_float2(b0x * nn, 0.f);
        by = make_float2(b0y * nn, 0.f);
        bz = make_float2(b0z * nn, 0.f);
    }
    out[t] = ex;
    out[nk + t] = ey;
    out[2 * nk + t] = bx;
    out[3 * nk + t] = by;
    out[4 * nk + t] = bz;
}

// darwin_et_kernel (solver_darwin.hpp:83-129): dJ/dt = dcu - i S(amu), transverse
// projection, E_T = -G_et (dJ/dt)_T with G_et = 1/(eps0 c^2 k^2 + n0).
__global__ void darwin_et_kernel(const float2* __restrict__ src, float2* __restrict__ out,
                                 KGeom kg, double eps0, double c, double n0) {
    const int nk = kg.nkx * kg.ny;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk) return;
    src += (size_t)blockIdx.y * 7 * nk;  // run blockIdx.y of a batch
    out += (size_t)blockIdx.y * 3 * nk;
    const int i = t % kg.nkx, j = t / kg.nkx;
    double kx, ky;
    kvec(kg, i, j, kx, ky);
    const double k2 = kx * kx + ky * ky;
    float2 ex = make_float2(0.f, 0.f), ey = ex, ez = ex;
    if (!nyquist(kg, i, j) && k2 > 0.0) {
        const double G = 1.0 / (eps0 * c * c * k2 + n0);
        const float2 dcx = src[t], dcy = src[nk + t], dcz = src[2 * nk + t];
        const float2 a0 = src[3 * nk + t], a1 = src[4 * nk + t], a2 = src[5 * nk + t],
                     a3 = src[6 * nk + t];
        const double Sxr = kx * a0.x + ky * a1.x, Sxi = kx * a0.y + ky * a1.y;
        const double Syr = kx * a1.x - ky * a0.x, Syi = kx * a1.y - ky * a0.y;
        const double Szr = kx * a2.x + ky * a3.x, Szi = kx * a2.y + ky * a3.y;
        double Jxr = dcx.x + Sxi, Jxi = dcx.y - Sxr;
        double Jyr = dcy.x + Syi, Jyi = dcy.y - Syr;
        const double Jzr = dcz.x + Szi, Jzi = dcz.y - Szr;
        const double inv = 1.0 / k2;
        const double kdr = (kx * Jxr + ky * Jyr) * inv, kdi = (kx * Jxi + ky * Jyi) * inv;
        Jxr -= kx * kdr;
        Jxi -= kx * kdi;
        Jyr -= ky * kdr;
        Jyi -= ky * kdi;
        ex = make_float2((float)(-G * Jxr), (float)(-G * Jxi));
        ey = make_float2((float)(-G * Jyr), (float)(-G * Jyi));
        ez = make_float2((float)(-G * Jzr), (float)(-G * Jzi));
    }
    out[t] = ex;
    out[nk + t] = ey;
    out[2 * nk + t] = ez;
}

}  // namespace

// src4 = [rho|Jx|Jy|Jz] (float, contiguous) -> out5 = [ELx|ELy|Bx|By|Bz];
// stag: Jx, Jy staggered half a cell (Esirkepov current).  runs > 1: a batch
// of runs, src4 = [run][4 grids], out5 = [run][5 grids] (enable_em_batch_fft).
void solve_darwin_elb(aw_ctx& c, const float* src4, float* out5, double eps0, double cl,
                      const float B0[3], hipStream_t s, bool stag, int runs) {
    AW_REQUIRE(runs == 1 || runs == c.em_batch, "Darwin batch plans not prepared");
    const int64_t nk = nkmodes(c);
    c.ktmp.reserve((size_t)2 * 12 * nk * runs);
    float2* k4 = reinterpret_cast<float2*>(c.ktmp.get());
    float2* k5 = k4 + 4 * nk * runs;
    fft_exec(c, runs == 1 ? c.r2c4 : c.r2c4B, const_cast<float*>(src4), k4, s);
    KGeom kg{c.g.nx, c.g.ny, c.g.nx / 2 + 1, 6.28318530717958647692 / c.g.Lx,
             6.28318530717958647692 / c.g.Ly};
    darwin_elb_kernel<<<dim3(blocks_for(nk), runs), kThreads, 0, s>>>(
        k4, k5, kg, eps0, cl, B0[0], B0[1], B0[2], stag ? 1 : 0);
    AW_LAUNCH_CHECK();
    fft_exec(c, runs == 1 ? c.c2r5 : c.c2r5B, k5, out5, s);
}

// src7 = [dcux|dcuy|dcuz|amu0..3] -> out3 = [ETx|ETy|ETz]; runs > 1 as above.
void solve_darwin_et(aw_ctx& c, const float* src7, float* out3, double eps0, double cl,
                     double n0, hipStream_t s, int runs) {
    AW_REQUIRE(runs == 1 || runs == c.em_batch, "Darwin batch plans not prepared");
    const int64_t nk = nkmodes(c);
    c.ktmp.reserve((size_t)2 * 12 * nk * runs);
    float2* k7 = reinterpret_cast<float2*>(c.ktmp.get());
    float2* k3 = k7 + 7 * nk * runs;
    fft_exec(c, runs == 1 ? c.r2c7 : c.r2c7B, const_cast<float*>(src7), k7, s);
    KGeom kg{c.g.nx, c.g.ny, c.g.nx / 2 + 1, 6.28318530717958647692 / c.g.Lx,
             6.28318530717958647692 / c.g.Ly};
    darwin_et_kernel<<<dim3(blocks_for(nk), runs), kThreads, 0, s>>>(k7, k3, kg, eps0, cl, n0);
    AW_LAUNCH_CHECK();
    fft_exec(c, runs == 1 ? c.c2r3 : c.c2r3B, k3, out3, s);
}

// General form with separate output buffers.
void solve_es(aw_ctx& c, const float* rho, float* Ex, float* Ey, double eps0, hipStream_t s) {
    const int64_t nk = nkmodes(c);
    c.ktmp.reserve((size_t)6 * nk);
    float2* rk = reinterpret_cast<float2*>(c.ktmp.get());
    float2* exk = rk + nk;
    float2* eyk = exk + nk;
    fft_exec(c, c.r2c, r2c_input(c, rho, s), rk, s);
    KGeom kg{c.g.nx, c.g.ny, c.g.nx / 2 + 1, 6.28318530717958647692 / c.g.Lx,
             6.28318530717958647692 / c.g.Ly};
    es_kspace_kernel<<<blocks_for(nk), kThreads, 0, s>>>(rk, exk, eyk, kg, eps0);
    AW_LAUNCH_CHECK();
    fft_exec(c, c.c2r, exk, Ex, s);
    fft_exec(c, c.c2r, eyk, Ey, s);
}

}  // namespace aw
