// capi.hip -- the C ABI (include/aw.h): context, low-level operators and the
// Simulation engine.
//
// Simulation mirrors Simulation<Cfg> (include/pic/simulation.hpp:38-140):
// init = load, t=0 deposit + field solve, leapfrog half step back; step(n) =
// deposit (already fused into the previous push) -> solve -> push -> wrap ->
// diagnostics.  The ES step runs as: [fixed->float rho^n] -> R2C -> k-space
// -> batched C2R -> fused push+deposit, with a stable tile re-sort on a
// cadence or when too many particles strayed from their tile apron.
#include "aw_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

namespace aw {

// --- other translation units ---------------------------------------------------
void init_fft(aw_ctx& c);
void destroy_fft(aw_ctx& c) noexcept;
void solve_es(aw_ctx& c, const float* rho, float* Ex, float* Ey, double eps0, hipStream_t s);
void push_es(aw_ctx& c, const aw_particles& p, const float* Ex, const float* Ey,
             const aw_run_params& rp, bool has_b0, bool half_back, hipStream_t s);
int push_deposit_es(aw_ctx& c, const aw_particles& p, const int32_t* off, const float* Ex,
                    const float* Ey, const aw_run_params& rp, bool has_b0,
                    unsigned long long* strays, double* ke_part, int bpt, hipStream_t s);
void cell_index(aw_ctx& c, const aw_particles& p, int32_t* cell, hipStream_t s);
void migrate(aw_ctx& c, const aw_particles& p, int32_t* cell, hipStream_t s);
void load_species(aw_ctx& c, const aw_species* sp, int nsp, const aw_run_params& rp,
                  const double* mode_amp, int n_modes, const aw_particles& p,
                  double* weights_out, hipStream_t s);
void load_legacy(aw_ctx& c, const aw_run_params& rp, const aw_particles& p, hipStream_t s);
int ke_partials(const aw_particles& p, double* part, hipStream_t s);
int grid_partials(const float* Ex, const float* Ey, const float* rho, int64_t nc, double* part,
                  hipStream_t s);
void finalize_diag(const double* kep, int nke, double w_scale, const double* gp, int ng,
                   double* out4, hipStream_t s, double* stage);
size_t ke_stage_capacity(size_t nke);
int reduce_partial_capacity();
int reduce_grid_offset();
int reduce_out_offset();
int tile_bpt(const aw_ctx& c, int64_t n);
bool uz_evolves(const aw_run_params& rp, bool has_b0);
void enable_em_fft(aw_ctx& c);
void solve_darwin_elb(aw_ctx& c, const float* src4, float* out5, double eps0, double cl,
                      const float B0[3], hipStream_t s, bool stag, int runs);
void solve_darwin_et(aw_ctx& c, const float* src7, float* out3, double eps0, double cl,
                     double n0, hipStream_t s, int runs);
void enable_em_batch_fft(aw_ctx& c, int R);
int dcu_part_capacity();
void em_moments(aw_ctx& c, const aw_particles& p, const int32_t* off, const int32_t* idx,
                unsigned long long* acc8, const float* cq, int bpt, unsigned long long* strays,
                hipStream_t s, const EmBatch& b);
void em_dcu(aw_ctx& c, const aw_particles& p, const int32_t* off, const int32_t* idx,
            const float* const* F, unsigned long long* acc3, const float* c_dcu, int bpt,
            unsigned long long* strays, hipStream_t s, const EmBatch& b);
int em_push(aw_ctx& c, const aw_particles& p, const int32_t* off, const int32_t* idx,
            const float* const* F, unsigned long long* acc8, const float* cq, double qm,
            double dt, int bpt, unsigned long long* strays, double* ke_part, hipStream_t s,
            bool esk, const EmBatch& b);
void deposit_rho_j_amu(aw_ctx& c, const aw_particles& p, double qm, float* const out[8],
                       hipStream_t s);
void deposit_dcu(aw_ctx& c, const aw_particles& p, double qm, const float* const* F,
                 float* const out[3], hipStream_t s);
void push_boris_em(aw_ctx& c, const aw_particles& p, double qm, double dt, const float* const* F,
                   hipStream_t s);
void convert_multi(unsigned long long* acc, float* out, float* out2, int64_t ncell, int nq,
                   const double* inv, hipStream_t s, int runs, long long sa, long long so,
                   long long so2, int si, const long long* base0 = nullptr);
void form_e(const FormE& f, const aw_grid& g, const aw_run_params& rp, double t, hipStream_t s,
            int runs);
void dcu_scale(const float* const* F, int64_t ncell, float* part, double coef_abs,
               double coef_signed, double umax, const double* dev, float* c_dcu, double* inv_dcu,
               hipStream_t s, int runs, long long sE, long long sB, int si);

void check_grid(const aw_grid* g) {
    AW_REQUIRE(g != nullptr, "grid is NULL");
    AW_REQUIRE(g->nx > 0 && g->ny > 0, "grid nx, ny must be positive");
    AW_REQUIRE(g->dx > 0 && g->dy > 0, "grid dx, dy must be positive");
    AW_REQUIRE((int64_t)g->nx * g->ny < (1ll << 30), "grid too large");
}

void check_particles(const aw_particles* p) {
    AW_REQUIRE(p != nullptr, "particles is NULL");
    AW_REQUIRE(p->n >= 0 && p->n < (int64_t)INT32_MAX, "particle count out of range");
    if (p->n > 0)
        AW_REQUIRE(p->x && p->y && p->ux && p->uy && p->uz, "particle arrays must be non-NULL");
}

namespace {

thread_local std::string g_last_error;

// also writes the R2C input copy (c.rin) so the solve needs no extra copy
void fix_to_float(aw_ctx& c, float* out, bool assign, hipStream_t s,
                  unsigned long long* counter = nullptr) {
    c.rin.reserve((size_t)ncell(c));
    fix_to_float_add(c.acc.get(), out, ncell(c), c.fix_log2, true, s, assign, counter,
                     assign ? c.rin.get() : nullptr, 1, nullptr, assign ? c.rho_base : 0);
}

}  // namespace

void set_ctx_tiles(aw_ctx& c, int tile_x, int tile_y) {
    c.tx = std::min(tile_x > 0 ? tile_x : 16, c.g.nx);
    c.ty = std::min(tile_y > 0 ? tile_y : 16, c.g.ny);
    c.ntx = (c.g.nx + c.tx - 1) / c.tx;
    c.nty = (c.g.ny + c.ty - 1) / c.ty;
    c.ntiles = c.ntx * c.nty;
}

void sort_tiles(aw_ctx& c, const aw_particles& src, const aw_particles& dst, int32_t* perm,
                int32_t* off, hipStream_t s) {
    if (tile_sort_onepass(c, src, dst, off, nullptr, s)) return;
    c.sort.keyA.reserve((size_t)std::max<int64_t>(1, src.n));
    compute_tile_keys(c, src, c.sort.keyA.get(), false, s);
    stable_sort_keys(c, src.n, (uint32_t)c.ntiles, perm, off, s);
    permute_particles(src, dst, perm, s);
}

void set_last_error(const std::string& msg) { g_last_error = msg; }
const std::string& last_error() { return g_last_error; }

}  // namespace aw

using namespace aw;

aw_ctx::~aw_ctx() { destroy_fft(*this); }

// ============================================================================
// Simulation engine
// ============================================================================

struct aw_sim {
    std::unique_ptr<aw_ctx> ctx;
    aw_grid g{};
    aw_run_params rp{};
    std::vector<aw_species> species;
    aw_sim_config cfg{};
    hipStream_t stream = nullptr;
    int64_t n = 0;
    double w_uniform = 0.0;
    bool per_w = false;
    DevBuf<float> P[2][6];
    int cur = 0;
    DevBuf<int32_t> tile_off, perm;
    DevBuf<float> rho, E;
    DevBuf<unsigned long long> strays;
    DevBuf<double> ke_part, grid_part, diag_out;
    int diag_cap = 256;
    std::vector<int64_t> diag_pending;  // steps with a slot in diag_out (runs slots each)
    std::vector<std::vector<aw_diag_sample>> hist;  // per run

    // Darwin batch (aw_ensemble_create with the Darwin model): runs of one
    // shape (grid, particle count, weights, step parameters) stored run after
    // run in every particle and grid buffer; the kernels take the run from
    // blockIdx.y.  runs = 1 is a plain simulation.
    int runs = 1;
    std::vector<std::vector<aw_species>> rspecies;  // species of run r (runs > 1)
    std::vector<aw_run_params> rrp;                 // load parameters (seed) of run r
    std::vector<std::vector<double>> rmodes;        // mode amplitudes of run r
    size_t kcap = 0;                                // KE partial slots per run
    bool initialized = false, fields_ready = false, sort_requested = false;
    bool range_error = false;  // a Darwin moment left its fixed-point range
    int64_t steps_since_sort = 0, nsorts = 0;
    unsigned long long* stray_host = nullptr;
    hipEvent_t stray_ev = nullptr;
    bool stray_pending = false;
    int stray_age = 0;          // steps since the pending stray read-back was issued
    bool j_staggered = false;   // acc8's current is the Esirkepov (staggered) one
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_ev;
    double prof_ms = 0.0;
    int64_t prof_launches = 0;
    int bpt = 0;

    // ---- spectral Darwin (field_model == AW_MODEL_DARWIN) ----
    bool em = false;
    DevBuf<unsigned long long> acc8, acc3;  // [rho|J|amu] next state, [dcu]
    DevBuf<float> src;   // [rho|Jx|Jy|Jz] ++ [dcux|dcuy|dcuz|amu0..amu3]
    DevBuf<float> fftin; // same layout: the R2C inputs (may be overwritten by rocFFT)
    DevBuf<float> elb;   // [ELx|ELy|Bx|By|Bz] (B includes B0)
    DevBuf<float> et;    // [ETx|ETy|ETz]
    DevBuf<float> e3;    // E = E_L + E_T + pump (push)
    DevBuf<float> ed3;   // E_L + pump (dcu gather)
    DevBuf<double> inv;  // inverse scales: rho, J x3, amu x4, dcu x3
    DevBuf<float> cdcu, fmax_part;
    DevBuf<float> cqd;  // rho, J, amu coefficients (device)
    DevBuf<long long> rbase;  // per run: the mean charge density at the rho scale
    float cq[3] = {0.f, 0.f, 0.f};
    double dcu_abs = 0.0, dcu_signed = 0.0, umax = 1.0;
    bool moments_ready = false;

    float* Ex(int r = 0) const {
        return em ? e3.get() + (size_t)r * 3 * ncell(*ctx) : E.get();
    }
    float* Ey(int r = 0) const { return Ex(r) + ncell(*ctx); }
    float* rho_ptr(int r = 0) const {
        return em ? src.get() + (size_t)r * 11 * ncell(*ctx) : rho.get();
    }
    float* B3() const { return elb.get() + 2 * ncell(*ctx); }
    // the uniform batch description of the particle kernels (astride: the
    // per-run size of the accumulator they write)
    EmBatch batch(long long astride) const {
        const long long nc = ncell(*ctx);
        EmBatch b;
        b.runs = runs;
        b.pstride = n;
        b.estride = 3 * nc;
        b.bstride = 5 * nc;
        b.astride = astride;
        b.kstride = (long long)kcap;
        b.ostride = ctx->ntiles + 1;
        return b;
    }
    // E (3 grids) and B as the six component pointers of the particle kernels
    void fields6(const float* E3, const float* F[6]) const {
        const int64_t nc = ncell(*ctx);
        for (int q = 0; q < 3; ++q) {
            F[q] = E3 + q * nc;
            F[3 + q] = B3() + q * nc;
        }
    }

    void setup_em() {
        em = true;
        const int64_t nc = ncell(*ctx);
        enable_em_fft(*ctx);
        const size_t R = (size_t)runs;
        if (runs > 1) enable_em_batch_fft(*ctx, runs);
        acc8.resize(R * 8 * nc);
        acc3.resize(R * 3 * nc);
        src.resize(R * 11 * nc);
        fftin.resize(R * 11 * nc);  // [run][4] R2C inputs, then [run][7]
        elb.resize(R * 5 * nc);
        et.resize(R * 3 * nc);
        e3.resize(R * 3 * nc);
        ed3.resize(R * 3 * nc);
        inv.resize(R * 11);
        cdcu.resize(R);
        fmax_part.resize(R * (size_t)dcu_part_capacity());
        // static per-quantity fixed-point scales: the per-contribution bound maps
        // to 2^48, leaving 2^14 bound-sized contributions of headroom per cell
        double wmax = w_uniform;
        for (const auto& s : species)
            wmax = std::max(wmax, s.density * g.dx * g.dy / s.ppc);
        const double area = g.dx * g.dy;
        umax = 2.0 * std::max(1.0, rp.c);
        const double cabs = std::fabs(rp.qm) * wmax / area;
        auto expo = [](double bound) {
            return 48 - (int)std::ceil(std::log2(std::max(bound, 1e-300)));
        };
        const int Sr = expo(cabs), Sj = expo(cabs * umax), Sa = expo(cabs * umax * umax);
        const double mult = per_w ? 1.0 : w_uniform;  // per-particle w multiplies in-kernel
        cq[0] = (float)std::ldexp(rp.qm * mult / area, Sr);
        cq[1] = (float)std::ldexp(rp.qm * mult / area, Sj);
        cq[2] = (float)std::ldexp(rp.qm * mult / area, Sa);
        double h[11] = {std::ldexp(1.0, -Sr), std::ldexp(1.0, -Sj), std::ldexp(1.0, -Sj),
                        std::ldexp(1.0, -Sj), std::ldexp(1.0, -Sa), std::ldexp(1.0, -Sa),
                        std::ldexp(1.0, -Sa), std::ldexp(1.0, -Sa), 1.0, 1.0, 1.0};
        for (int r = 0; r < runs; ++r)
            AW_HIP(hipMemcpy(inv.get() + (size_t)r * 11, h, sizeof(h), hipMemcpyHostToDevice));
        // each run's mean charge density at the rho scale, taken off the E_L input
        std::vector<long long> hb((size_t)runs);
        for (int r = 0; r < runs; ++r) {
            const auto& sp = runs > 1 ? rspecies[(size_t)r] : species;
            double dens = cfg.legacy_load ? (rp.n0 > 0 ? rp.n0 : 1.0) : 0.0;
            for (const auto& q : sp) dens += q.density;
            hb[(size_t)r] = rho_baseline(rp.qm, dens, Sr);
        }
        rbase.resize(R);
        AW_HIP(hipMemcpy(rbase.get(), hb.data(), R * sizeof(long long), hipMemcpyHostToDevice));
        cqd.resize(3);
        AW_HIP(hipMemcpy(cqd.get(), cq, sizeof(cq), hipMemcpyHostToDevice));
        dcu_abs = rp.qm * rp.qm * wmax / area;
        dcu_signed = rp.qm * rp.qm * mult / area;
        acc8.zero(stream);
        acc3.zero(stream);
        et.zero(stream);
    }

    // The t = 0 moments and fields, as Simulation::init solves them
    // (include/pic/simulation.hpp:58-66): fields() holds E and B before the
    // first step.  The moments are consumed; step 0 deposits them again.
    void init_em() {
        acc8.zero(stream);
        acc3.zero(stream);
        moments_ready = false;
        solve_fields_em(0.0);
    }

    // darwin_fields (include/pic/simulation.hpp:103-119) at time t from the
    // moments of the current state (deposited here unless the last push did)
    void solve_fields_em(double t) {
        const hipStream_t s = stream;
        const int64_t nc = ncell(*ctx);
        const int R = runs;
        if (!moments_ready) {
            em_moments(*ctx, particles(), tile_off.get(), nullptr, acc8.get(), cqd.get(), bpt,
                       nullptr, s, batch(8 * nc));
            j_staggered = false;
        }
        moments_ready = false;
        // per run: rho, J -> src[0..3] and the R2C input block [run][4]; amu ->
        // src[7..10] and [run][7] entries 3..6 (fftin holds the copies rocFFT
        // may overwrite: [R x 4 grids][R x 7 grids])
        float* fin4 = fftin.get();
        float* fin7 = fftin.get() + (size_t)R * 4 * nc;
        convert_multi(acc8.get(), src.get(), fin4, nc, 4, inv.get(), s, R, 8 * nc, 11 * nc,
                      4 * nc, 11, rbase.get());
        convert_multi(acc8.get() + 4 * nc, src.get() + 7 * nc, fin7 + 3 * nc, nc, 4,
                      inv.get() + 4, s, R, 8 * nc, 11 * nc, 7 * nc, 11);
        solve_darwin_elb(*ctx, fin4, elb.get(), rp.eps0, rp.c, rp.B0, s, j_staggered, R);
        FormE fd{{ed3.get(), ed3.get() + nc, ed3.get() + 2 * nc},
                 {elb.get(), elb.get() + nc},
                 {nullptr, nullptr, nullptr}};
        fd.sE = 3 * nc;
        fd.sL = 5 * nc;
        form_e(fd, g, rp, t, s, R);
        if (sort_requested || (cfg.sort_every > 0 && steps_since_sort >= cfg.sort_every)) do_sort();
        const float* Fd[6];
        fields6(ed3.get(), Fd);
        dcu_scale(Fd, nc, fmax_part.get(), dcu_abs, dcu_signed, umax, nullptr, cdcu.get(),
                  inv.get() + 8, s, R, 3 * nc, 5 * nc, 11);
        em_dcu(*ctx, particles(), tile_off.get(), nullptr, Fd, acc3.get(), cdcu.get(), bpt,
               strays.get(), s, batch(3 * nc));
        convert_multi(acc3.get(), src.get() + 4 * nc, fin7, nc, 3, inv.get() + 8, s, R, 3 * nc,
                      11 * nc, 7 * nc, 11);
        solve_darwin_et(*ctx, fin7, et.get(), rp.eps0, rp.c, rp.n0, s, R);
        FormE fe{{e3.get(), e3.get() + nc, e3.get() + 2 * nc},
                 {elb.get(), elb.get() + nc},
                 {et.get(), et.get() + nc, et.get() + 2 * nc}};
        fe.sE = 3 * nc;
        fe.sL = 5 * nc;
        fe.sT = 3 * nc;
        form_e(fe, g, rp, t, s, R);
    }

    // darwin_fields + boris_em + migrate (include/pic/simulation.hpp:80-84)
    void step_em(int64_t nstep) {
        const hipStream_t s = stream;
        const int64_t nc = ncell(*ctx);
        solve_fields_em((double)nstep * rp.dt);
        const bool dump = rp.dump_every <= 0 || (nstep % rp.dump_every) == 0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (prof) {
            AW_HIP(hipEventCreate(&e0));
            AW_HIP(hipEventCreate(&e1));
            AW_HIP(hipEventRecord(e0, s));
        }
        const float* Fp[6];
        fields6(e3.get(), Fp);
        const bool esk = cfg.current_deposit == AW_CURRENT_ESIRKEPOV;
        const int nblk = em_push(*ctx, particles(), tile_off.get(), nullptr, Fp, acc8.get(),
                                 cqd.get(), rp.qm, rp.dt, bpt, strays.get(),
                                 dump ? ke_part.get() : nullptr, s, esk, batch(8 * nc));
        j_staggered = esk;
        if (prof) {
            AW_HIP(hipEventRecord(e1, s));
            prof_ev.emplace_back(e0, e1);
        }
        moments_ready = true;
        if (!stray_pending) {
            AW_HIP(hipMemcpyAsync(stray_host, strays.get(), sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s));
            AW_HIP(hipEventRecord(stray_ev, s));
            stray_pending = true;
            stray_age = 0;
        }
        AW_HIP(hipMemsetAsync(strays.get(), 0, sizeof(unsigned long long), s));
        if (dump) record_diag(nstep, nblk);
        ++steps_since_sort;
    }

    static constexpr size_t kGridPart = 3 * 512 + 3;  // grid partial slots per run

    void record_diag(int64_t nstep, int nblk) {
        if ((int)diag_pending.size() >= diag_cap) flush_diag();
        for (int r = 0; r < runs; ++r) {
            double* gp = grid_part.get() + (size_t)r * kGridPart;
            double* kp = ke_part.get() + (size_t)r * kcap;
            const int ng = grid_partials(Ex(r), Ey(r), rho_ptr(r), ncell(*ctx), gp, stream);
            finalize_diag(kp, nblk, per_w ? 1.0 : w_uniform, gp, ng,
                          diag_out.get() + 4 * (diag_pending.size() * runs + r), stream,
                          kp + nblk);
        }
        diag_pending.push_back(nstep);
    }

    // buffer set `which`, run r (n particles; run 0 also heads a batch)
    aw_particles particles(int which, int r = 0) const {
        const size_t o = (size_t)r * n;
        aw_particles p{};
        p.x = P[which][0].get() + o;
        p.y = P[which][1].get() + o;
        p.ux = P[which][2].get() + o;
        p.uy = P[which][3].get() + o;
        p.uz = P[which][4].get() + o;
        p.w = per_w ? P[which][5].get() + o : nullptr;
        p.n = n;
        p.w_uniform = w_uniform;
        return p;
    }
    aw_particles particles() const { return particles(cur); }

    ~aw_sim() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& e : prof_ev) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        if (stray_ev) (void)hipEventDestroy(stray_ev);
        if (stray_host) (void)hipHostFree(stray_host);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void do_sort() {
        for (int r = 0; r < runs; ++r)
            sort_tiles(*ctx, particles(cur, r), particles(cur ^ 1, r), perm.get(),
                       tile_off.get() + (size_t)r * (ctx->ntiles + 1), stream);
        cur ^= 1;
        steps_since_sort = 0;
        ++nsorts;
        sort_requested = false;
    }

    void flush_diag() {
        if (diag_pending.empty()) return;
        std::vector<double> h(diag_pending.size() * 4 * runs);
        AW_HIP(hipMemcpyAsync(h.data(), diag_out.get(), h.size() * sizeof(double),
                              hipMemcpyDeviceToHost, stream));
        AW_HIP(hipStreamSynchronize(stream));
        const double dxdy = g.dx * g.dy;
        for (size_t k = 0; k < diag_pending.size(); ++k)
            for (int r = 0; r < runs; ++r) {
                const double* v = h.data() + 4 * (k * runs + r);
                aw_diag_sample d{};
                d.step = diag_pending[k];
                d.time = (double)d.step * rp.dt;
                d.ke = 0.5 * v[0];
                d.ee = 0.5 * rp.eps0 * v[1] * dxdy;
                d.charge = v[3] * dxdy;
                d.max_e = std::sqrt(v[2]);
                d.total = d.ke + d.ee;
                hist[r].push_back(d);
            }
        diag_pending.clear();
    }

    // The stray count of a step is read back asynchronously.  Once the read-back
    // is kStrayLag steps old the host waits for it, so the re-sort decision
    // trails the device by at most that many steps however far ahead the host
    // enqueues (the queue then stays about that deep, enough to keep the device
    // busy).
    static constexpr int kStrayLag = 2;

    void poll_strays() {
        if (!stray_pending) return;
        if (++stray_age >= kStrayLag) AW_HIP(hipEventSynchronize(stray_ev));
        else if (hipEventQuery(stray_ev) != hipSuccess) return;
        stray_pending = false;
        const unsigned long long v = *stray_host;
        if (v >= kRangeFlag) range_error = true;
        const double frac = cfg.stray_frac > 0 ? cfg.stray_frac : 0.02;
        if ((double)(v % kRangeFlag) > frac * (double)n * runs) sort_requested = true;
    }

    // Tile geometry and workgroups per tile of the particle kernels.
    void apply_exec(int tile, int nbpt) {
        AW_HIP(hipStreamSynchronize(stream));
        set_ctx_tiles(*ctx, tile, tile);
        bpt = nbpt > 0 ? nbpt : tile_bpt(*ctx, n);
        size_buffers();
    }

    // tile offsets and KE partials (per run) for the current tiles and bpt
    void size_buffers() {
        tile_off.resize((size_t)runs * (ctx->ntiles + 1));
        kcap = ke_stage_capacity((size_t)ctx->ntiles * bpt);
        ke_part.resize((size_t)runs * kcap);
    }

    // Autotuner: for tile 8/16/32 the state is loaded and initialised, then the
    // particle kernels of one step are timed (HIP events, best of three, each on
    // a copy of the particles) for 1..16 workgroups per tile: the fused
    // push+deposit (ES), or the dcu deposit plus the EM push (Darwin, on zero
    // fields; a batch times its first run).  The fastest pair is kept and cached
    // per device, grid, particle count, weights and kernel variant for the
    // process.  Every deposit is fixed-point and order-free, so the choice
    // changes no result.
    void autotune() {
        const bool esk = cfg.current_deposit == AW_CURRENT_ESIRKEPOV;
        using Key = std::tuple<int, int, int, int64_t, bool, bool, bool, bool>;
        static std::mutex mu;
        static std::map<Key, std::pair<int, int>> cache;
        const Key key{ctx->device, g.nx, g.ny,  n,  per_w, uz_evolves(rp, cfg.has_b0 != 0),
                      em,          esk};
        {
            std::lock_guard<std::mutex> lk(mu);
            const auto it = cache.find(key);
            if (it != cache.end()) {
                apply_exec(it->second.first, it->second.second);
                return;
            }
        }
        struct Events {
            hipEvent_t a = nullptr, b = nullptr;
            ~Events() {
                if (a) (void)hipEventDestroy(a);
                if (b) (void)hipEventDestroy(b);
            }
        } ev;
        AW_HIP(hipEventCreate(&ev.a));
        AW_HIP(hipEventCreate(&ev.b));
        const hipStream_t s = stream;
        if (em) {  // the timed kernels read zero fields and a zero dcu coefficient
            e3.zero(s);
            elb.zero(s);
            cdcu.zero(s);
        }
        float best = 0.f;
        int best_tile = 0, best_bpt = 0;
        for (const int tile : {8, 16, 32}) {
            if (tile > g.nx || tile > g.ny) continue;
            if (em) {  // the EM push's LDS frame: 8 int64 moments + 6 field floats per node
                const size_t side = (size_t)tile + 2 * ctx->apron + 1;
                if (side * side * (8 * 8 + 24) > ctx->lds_per_block) continue;
            }
            apply_exec(tile, 1);
            init_state();
            const int64_t per_tile = n / ctx->ntiles;
            for (const int b : {1, 2, 4, 8, 16}) {
                if (b > 1 && per_tile / b < 256) break;
                float tmin = 0.f;
                for (int rep = 0; rep < 3; ++rep) {
                    const aw_particles src = particles(cur), dst = particles(cur ^ 1);
                    const size_t bytes = (size_t)n * sizeof(float);
                    AW_HIP(hipMemcpyAsync(dst.x, src.x, bytes, hipMemcpyDeviceToDevice, s));
                    AW_HIP(hipMemcpyAsync(dst.y, src.y, bytes, hipMemcpyDeviceToDevice, s));
                    AW_HIP(hipMemcpyAsync(dst.ux, src.ux, bytes, hipMemcpyDeviceToDevice, s));
                    AW_HIP(hipMemcpyAsync(dst.uy, src.uy, bytes, hipMemcpyDeviceToDevice, s));
                    AW_HIP(hipMemcpyAsync(dst.uz, src.uz, bytes, hipMemcpyDeviceToDevice, s));
                    AW_HIP(hipEventRecord(ev.a, s));
                    if (em) {
                        const float* F[6];
                        fields6(e3.get(), F);
                        em_dcu(*ctx, dst, tile_off.get(), nullptr, F, acc3.get(), cdcu.get(), b,
                               strays.get(), s, EmBatch{});
                        em_push(*ctx, dst, tile_off.get(), nullptr, F, acc8.get(), cqd.get(),
                                rp.qm, rp.dt, b, strays.get(), nullptr, s, esk, EmBatch{});
                    } else {
                        push_deposit_es(*ctx, dst, tile_off.get(), Ex(), Ey(), rp,
                                        cfg.has_b0 != 0, strays.get(), nullptr, b, s);
                    }
                    AW_HIP(hipEventRecord(ev.b, s));
                    AW_HIP(hipEventSynchronize(ev.b));
                    float ms = 0.f;
                    AW_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
                    if (rep == 0 || ms < tmin) tmin = ms;
                }
                if (best_tile == 0 || tmin < best) {
                    best = tmin;
                    best_tile = tile;
                    best_bpt = b;
                }
            }
        }
        if (best_tile == 0) return;  // grid smaller than every candidate tile
        apply_exec(best_tile, best_bpt);
        nsorts = 0;
        initialized = false;
        std::lock_guard<std::mutex> lk(mu);
        cache[key] = {best_tile, best_bpt};
    }

    void init() {
        if (cfg.autotune && n > 0) autotune();
        init_state();
    }

    void init_state() {
        const hipStream_t s = stream;
        const aw_particles p = particles(cur);
        if (cfg.legacy_load) {
            load_legacy(*ctx, rp, p, s);
        } else if (rspecies.empty()) {
            std::vector<double> wts(species.size());
            load_species(*ctx, species.data(), (int)species.size(), rp, cfg.mode_amp,
                         cfg.n_modes, p, wts.data(), s);
        } else {
            for (int r = 0; r < runs; ++r) {
                std::vector<double> wts(rspecies[r].size());
                load_species(*ctx, rspecies[r].data(), (int)rspecies[r].size(), rrp[r],
                             rmodes[r].empty() ? nullptr : rmodes[r].data(),
                             (int)rmodes[r].size(), particles(cur, r), wts.data(), s);
            }
        }
        do_sort();
        if (em) {  // Darwin: t=0 moments and fields; each step solves its own
            init_em();
            strays.zero(s);
            initialized = true;
            AW_HIP(hipStreamSynchronize(s));
            return;
        }
        ctx->acc.zero(s);
        deposit_rho_fixed(*ctx, particles(), rp.qm, AW_DEPOSIT_BINNED, tile_off.get(), s);
        fix_to_float(*ctx, rho.get(), true, s);
        solve_es_contig(*ctx, ctx->rin.get(), E.get(), rp.eps0, s);
        push_es(*ctx, particles(), Ex(), Ey(), rp, cfg.has_b0 != 0, true, s);
        strays.zero(s);  // afterwards cleared by each step's rho conversion
        fields_ready = true;
        initialized = true;
        AW_HIP(hipStreamSynchronize(s));
    }

    // The caller edited the particles between steps.  The reference's step
    // deposits its sources from the particles as they are
    // (include/pic/simulation.hpp:80-93), so such edits are part of its
    // contract; here the next step's sources come out of the previous push, so
    // they are rebuilt from the current state: re-sort, then rho (ES, into the
    // accumulator the next step solves from) or the moments (Darwin, deposited
    // by the next step).
    void particles_changed() {
        if (!initialized)
            fail(AW_ERR_STATE, "aw_sim_particles_changed before aw_sim_init", __FILE__, __LINE__);
        const hipStream_t s = stream;
        do_sort();
        if (em) {
            moments_ready = false;
            return;
        }
        ctx->acc.zero(s);
        deposit_rho_fixed(*ctx, particles(), rp.qm, AW_DEPOSIT_BINNED, tile_off.get(), s);
        fields_ready = false;
    }

    // The state a step reads, as checkpoint arrays: the particles, the
    // fixed-point sources the previous push deposited (ES: rho; Darwin: rho,
    // J, amu), and on the ES path the field and charge of the last solve.
    // flags = {model, fields_ready, moments_ready, j_staggered}.
    int32_t ck_flags[4] = {0, 0, 0, 0};
    std::vector<aw_checkpoint_array> ck_arrays() {
        if (!initialized)
            fail(AW_ERR_STATE, "checkpoint before aw_sim_init", __FILE__, __LINE__);
        AW_REQUIRE(runs == 1, "batched Darwin runs are not checkpointed");
        const aw_particles p = particles();
        const int64_t nc = ncell(*ctx);
        ck_flags[0] = em ? 1 : 0;
        ck_flags[1] = fields_ready ? 1 : 0;
        ck_flags[2] = moments_ready ? 1 : 0;
        ck_flags[3] = j_staggered ? 1 : 0;
        std::vector<aw_checkpoint_array> a;
        auto dev = [&](const char* nm, uint32_t t, uint64_t cnt, void* d) {
            a.push_back(aw_checkpoint_array{nm, t, cnt, d, 1});
        };
        dev("x", AW_CK_F32, (uint64_t)n, p.x);
        dev("y", AW_CK_F32, (uint64_t)n, p.y);
        dev("ux", AW_CK_F32, (uint64_t)n, p.ux);
        dev("uy", AW_CK_F32, (uint64_t)n, p.uy);
        dev("uz", AW_CK_F32, (uint64_t)n, p.uz);
        if (p.w) dev("w", AW_CK_F32, (uint64_t)n, p.w);
        if (em) {
            dev("moments_fixed", AW_CK_U64, (uint64_t)(8 * nc), acc8.get());
        } else {
            dev("rho_fixed", AW_CK_U64, (uint64_t)nc, ctx->acc.get());
            dev("E", AW_CK_F32, (uint64_t)(2 * nc), E.get());
            dev("rho", AW_CK_F32, (uint64_t)nc, rho.get());
        }
        a.push_back(aw_checkpoint_array{"state_flags", AW_CK_I32, 4, ck_flags, 0});
        return a;
    }

    void save(const char* path, const aw_checkpoint_header& h, const char* deck,
              const char* git) {
        const std::vector<aw_checkpoint_array> a = ck_arrays();
        AW_HIP(hipStreamSynchronize(stream));
        const aw_status st = aw_checkpoint_write(path, &h, deck, git, a.data(), (int32_t)a.size());
        if (st != AW_OK) fail(st, last_error(), __FILE__, __LINE__);
    }

    // Loads a checkpoint written by save() into a created simulation of the
    // same grid, species and model (no aw_sim_init), then re-sorts: results
    // continue bit for bit (the deposits do not depend on particle order).
    void restore(const char* path) {
        AW_REQUIRE(runs == 1, "batched Darwin runs are not checkpointed");
        aw_checkpoint* ck = nullptr;
        aw_status st = aw_checkpoint_open(path, &ck);
        if (st != AW_OK) fail(st, last_error(), __FILE__, __LINE__);
        std::unique_ptr<aw_checkpoint, aw_status (*)(aw_checkpoint*)> hold(ck, aw_checkpoint_close);
        int32_t na = 0;
        aw_checkpoint_info(ck, nullptr, &na, nullptr, nullptr);
        initialized = true;  // so ck_arrays() lays out the targets
        const std::vector<aw_checkpoint_array> want = ck_arrays();
        for (const auto& w : want) {
            int32_t found = -1;
            for (int32_t i = 0; i < na && found < 0; ++i) {
                char name[64];
                uint32_t t = 0;
                uint64_t cnt = 0;
                int32_t present = 0;
                aw_checkpoint_array_info(ck, i, name, sizeof name, &t, &cnt, &present);
                if (std::strcmp(name, w.name) == 0) {
                    if (t != w.dtype || cnt != w.count || !present) {
                        initialized = false;
                        fail(AW_ERR_INVALID,
                             std::string("checkpoint array '") + w.name +
                                 "' does not match this simulation (dtype, size or payload)",
                             __FILE__, __LINE__);
                    }
                    found = i;
                }
            }
            if (found < 0) {
                initialized = false;
                fail(AW_ERR_INVALID, std::string("checkpoint lacks array '") + w.name + "'",
                     __FILE__, __LINE__);
            }
            st = aw_checkpoint_read_array(ck, found, w.data, w.on_device);
            if (st != AW_OK) {
                initialized = false;
                fail(st, last_error(), __FILE__, __LINE__);
            }
        }
        if (ck_flags[0] != (em ? 1 : 0)) {
            initialized = false;
            fail(AW_ERR_INVALID, "checkpoint field model differs from this simulation", __FILE__,
                 __LINE__);
        }
        fields_ready = ck_flags[1] != 0;
        moments_ready = ck_flags[2] != 0;
        j_staggered = ck_flags[3] != 0;
        do_sort();
        strays.zero(stream);
        stray_pending = false;
        AW_HIP(hipStreamSynchronize(stream));
    }

    void step(int64_t nstep) {
        if (!initialized) fail(AW_ERR_STATE, "aw_sim_step before aw_sim_init", __FILE__, __LINE__);
        const hipStream_t s = stream;
        poll_strays();
        if (range_error)
            fail(AW_ERR_INVALID,
                 "Darwin moments left the fixed-point range: a particle exceeded the |u| <= "
                 "2 max(1, c) bound the moment scales assume (contributions saturated, no sum "
                 "wrapped)",
                 __FILE__, __LINE__);
        if (em) {
            step_em(nstep);
            return;
        }
        if (!fields_ready) {
            fix_to_float(*ctx, rho.get(), true, s, strays.get());
            solve_es_contig(*ctx, ctx->rin.get(), E.get(), rp.eps0, s);
        }
        fields_ready = false;
        if (sort_requested || (cfg.sort_every > 0 && steps_since_sort >= cfg.sort_every)) do_sort();
        const bool dump = rp.dump_every <= 0 || (nstep % rp.dump_every) == 0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (prof) {
            AW_HIP(hipEventCreate(&e0));
            AW_HIP(hipEventCreate(&e1));
            AW_HIP(hipEventRecord(e0, s));
        }
        const int nblk = push_deposit_es(*ctx, particles(), tile_off.get(), Ex(), Ey(), rp,
                                         cfg.has_b0 != 0, strays.get(),
                                         dump ? ke_part.get() : nullptr, bpt, s);
        if (prof) {
            AW_HIP(hipEventRecord(e1, s));
            prof_ev.emplace_back(e0, e1);
        }
        if (!stray_pending) {
            AW_HIP(hipMemcpyAsync(stray_host, strays.get(), sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s));
            AW_HIP(hipEventRecord(stray_ev, s));
            stray_pending = true;
            stray_age = 0;
        }
        if (dump) record_diag(nstep, nblk);
        ++steps_since_sort;
    }

    void collect_prof() {
        AW_HIP(hipStreamSynchronize(stream));
        for (auto& e : prof_ev) {
            float ms = 0.f;
            AW_HIP(hipEventElapsedTime(&ms, e.first, e.second));
            prof_ms += ms;
            ++prof_launches;
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        prof_ev.clear();
    }
};

// ============================================================================
// C ABI
// ============================================================================

namespace aw {

aw_ctx* make_ctx(const aw_grid* g, int tile_x, int tile_y) {
    check_grid(g);
    auto c = std::make_unique<aw_ctx>();
    c->g = *g;
    if (c->g.Lx <= 0) c->g.Lx = g->nx * g->dx;
    if (c->g.Ly <= 0) c->g.Ly = g->ny * g->dy;
    set_ctx_tiles(*c, tile_x, tile_y);
    AW_HIP(hipGetDevice(&c->device));
    AW_HIP(hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, c->device));
    int lds = 0;
    AW_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    c->lds_per_block = (size_t)lds;
    init_fft(*c);
    c->acc.resize((size_t)ncell(*c));
    AW_HIP(hipMemset(c->acc.get(), 0, (size_t)ncell(*c) * sizeof(unsigned long long)));
    c->red.resize((size_t)reduce_partial_capacity());
    return c.release();
}

// Lowers S until no int64 cell sum of rho can wrap: even every particle's
// largest contribution, |qm| wmax / (dx dy) 2^S, summed in one cell stays
// below 2^62.  A configuration that needs S < 8 is an error, not a wrap.
int rho_scale_guarded(const aw_grid& g, double qm, double wmax, int64_t n, int S) {
    const double cmax = std::fabs(qm) * wmax / (g.dx * g.dy);
    AW_REQUIRE(std::isfinite(cmax), "particle weight or charge is not finite");
    const double total = (double)std::max<int64_t>(1, n) * cmax;
    while (S > 8 && total * std::ldexp(1.0, S) >= std::ldexp(1.0, 62)) --S;
    if (total * std::ldexp(1.0, S) >= std::ldexp(1.0, 62))
        fail(AW_ERR_INVALID,
             "rho fixed point: n |qm| w / (dx dy) = " + std::to_string(total) +
                 " leaves no scale >= 2^8 whose int64 cell sums cannot wrap",
             __FILE__, __LINE__);
    return S;
}

int fixed_scale_for(const std::vector<aw_species>& sp, const aw_run_params& rp, bool legacy) {
    // rho magnitude ~ |qm| * total density; keep 2^23 headroom over it
    double dens = 0.0;
    if (legacy) dens = rp.n0 > 0 ? rp.n0 : 1.0;
    for (const auto& s : sp) dens += std::fabs(s.density);
    const double rmax = std::max(1e-30, std::fabs(rp.qm) * std::max(dens, 1e-30));
    const int e = (int)std::ceil(std::log2(rmax));
    return std::clamp(40 - e, 8, 60);
}

}  // namespace aw

extern "C" {

void aw_last_error(char* buf, size_t len) {
    if (!buf || !len) return;
    std::strncpy(buf, last_error().c_str(), len - 1);
    buf[len - 1] = 0;
}

const char* aw_version(void) { return "arcwarden-amd 0.1 gfx950"; }

aw_status aw_rho_fixed_scale(const aw_grid* g, double qm, double density, double wmax,
                             int64_t n, int32_t* scale_log2) {
    return guard("aw_rho_fixed_scale", [&] {
        check_grid(g);
        AW_REQUIRE(scale_log2, "scale_log2 is NULL");
        AW_REQUIRE(n >= 0 && wmax >= 0.0 && density >= 0.0, "n, wmax and density must be >= 0");
        aw_species sp{};
        sp.density = density;
        sp.ppc = 1;
        aw_run_params rp{};
        rp.qm = qm;
        const int S = fixed_scale_for(std::vector<aw_species>{sp}, rp, false);
        *scale_log2 = rho_scale_guarded(*g, qm, wmax, n, S);
    });
}

aw_status aw_device_alloc(size_t bytes, void** out) {
    return guard("aw_device_alloc", [&] {
        AW_REQUIRE(out, "out is NULL");
        *out = nullptr;
        if (bytes) AW_HIP(hipMalloc(out, bytes));
    });
}

aw_status aw_device_free(void* p) {
    return guard("aw_device_free", [&] {
        if (p) AW_HIP(hipFree(p));
    });
}

aw_status aw_copy(void* dst, const void* src, size_t bytes) {
    return guard("aw_copy", [&] {
        if (!bytes) return;
        AW_REQUIRE(dst && src, "NULL pointer");
        AW_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    });
}

aw_status aw_memset_zero(void* p, size_t bytes, aw_stream s) {
    return guard("aw_memset_zero", [&] {
        if (!bytes) return;
        AW_REQUIRE(p, "NULL pointer");
        AW_HIP(hipMemsetAsync(p, 0, bytes, S(s)));
    });
}

aw_status aw_stream_synchronize(aw_stream s) {
    return guard("aw_stream_synchronize", [&] { AW_HIP(hipStreamSynchronize(S(s))); });
}

aw_status aw_device_synchronize(void) {
    return guard("aw_device_synchronize", [&] { AW_HIP(hipDeviceSynchronize()); });
}

void aw_run_params_defaults(aw_run_params* rp) {
    if (!rp) return;
    std::memset(rp, 0, sizeof(*rp));
    rp->n0 = 1.0;
    rp->wpe = 1.0;
    rp->eps0 = 1.0;
    rp->qm = -1.0;
    rp->c = 1.0;
    rp->ndc = 1;
    rp->bnd_nd = 64;
    rp->bnd_numax = 1.0;
    rp->df_tpar = 1.0;
    rp->df_tperp = 1.0;
    rp->b0_sref = 1.0;
    rp->b0_c[0] = 1.0;
    rp->df_rho = 1.0;
    rp->df_kappa = 0.3;
    rp->ant_sigma = 2.0;
    rp->weight = 1.0;
    rp->perturb_kx = 1;
}

void aw_species_defaults(aw_species* sp) {
    if (!sp) return;
    std::memset(sp, 0, sizeof(*sp));
    std::strncpy(sp->name, "electrons", sizeof(sp->name) - 1);
    sp->density = 1.0;
    sp->lc_rho = 1.0;
    sp->lc_kappa = 0.3;
}

aw_status aw_ctx_create(const aw_grid* g, int32_t tile_x, int32_t tile_y, aw_ctx** out) {
    return guard("aw_ctx_create", [&] {
        AW_REQUIRE(out != nullptr, "out is NULL");
        *out = make_ctx(g, tile_x, tile_y);
    });
}

aw_status aw_ctx_destroy(aw_ctx* ctx) {
    return guard("aw_ctx_destroy", [&] {
        if (ctx) (void)hipDeviceSynchronize();
        delete ctx;
    });
}

aw_status aw_ctx_set_fixed_scale(aw_ctx* ctx, int32_t scale_log2) {
    return guard("aw_ctx_set_fixed_scale", [&] {
        AW_REQUIRE(ctx, "ctx is NULL");
        AW_REQUIRE(scale_log2 >= 0 && scale_log2 <= 62, "scale_log2 must be in [0, 62]");
        ctx->fix_log2 = scale_log2;
    });
}

aw_status aw_ctx_tiles(const aw_ctx* ctx, int32_t* ntiles, int32_t* ntx, int32_t* tx,
                       int32_t* ty) {
    return guard("aw_ctx_tiles", [&] {
        AW_REQUIRE(ctx, "ctx is NULL");
        if (ntiles) *ntiles = ctx->ntiles;
        if (ntx) *ntx = ctx->ntx;
        if (tx) *tx = ctx->tx;
        if (ty) *ty = ctx->ty;
    });
}

aw_status aw_cell_index(aw_ctx* ctx, const aw_particles* p, int32_t* cell, aw_stream s) {
    return guard("aw_cell_index", [&] {
        AW_REQUIRE(ctx && cell, "ctx/cell is NULL");
        check_particles(p);
        cell_index(*ctx, *p, cell, S(s));
    });
}

aw_status aw_bin_tiles(aw_ctx* ctx, const aw_particles* p, int32_t* off, int32_t* idx,
                       aw_stream s) {
    return guard("aw_bin_tiles", [&] {
        AW_REQUIRE(ctx && off && idx, "ctx/off/idx is NULL");
        check_particles(p);
        bin_tiles(*ctx, *p, idx, off, S(s));
    });
}

aw_status aw_sort_cells(aw_ctx* ctx, const aw_particles* p, int32_t* perm, int32_t* off,
                        aw_stream s) {
    return guard("aw_sort_cells", [&] {
        AW_REQUIRE(ctx && perm, "ctx/perm is NULL");
        check_particles(p);
        ctx->sort.keyA.reserve((size_t)std::max<int64_t>(1, p->n));
        compute_tile_keys(*ctx, *p, ctx->sort.keyA.get(), true, S(s));
        stable_sort_keys(*ctx, p->n, (uint32_t)(ctx->ntiles * ctx->tx * ctx->ty), perm, off, S(s));
    });
}

aw_status aw_sort_by_tile(aw_ctx* ctx, const aw_particles* src, aw_particles* dst,
                          int32_t* off, aw_stream s) {
    return guard("aw_sort_by_tile", [&] {
        AW_REQUIRE(ctx && dst && off, "ctx/dst/off is NULL");
        check_particles(src);
        check_particles(dst);
        AW_REQUIRE(src->n == dst->n, "src and dst particle counts differ");
        AW_REQUIRE((src->w == nullptr) == (dst->w == nullptr), "w present in only one of src/dst");
        AW_REQUIRE(src->n == 0 || src->x != dst->x,
                   "aw_sort_by_tile needs distinct src and dst buffers");
        ctx->sort.bin_perm.reserve((size_t)std::max<int64_t>(1, src->n));
        sort_tiles(*ctx, *src, *dst, ctx->sort.bin_perm.get(), off, S(s));
    });
}

aw_status aw_migrate(aw_ctx* ctx, aw_particles* p, int32_t* cell, aw_stream s) {
    return guard("aw_migrate", [&] {
        AW_REQUIRE(ctx, "ctx is NULL");
        check_particles(p);
        migrate(*ctx, *p, cell, S(s));
    });
}

aw_status aw_load_species(aw_ctx* ctx, const aw_species* sp, int32_t nspecies,
                          const aw_run_params* rp, const double* mode_amp, int32_t n_modes,
                          aw_particles* p, double* weights_out, aw_stream s) {
    return guard("aw_load_species", [&] {
        AW_REQUIRE(ctx && sp && rp && nspecies > 0, "ctx/species/rp invalid");
        AW_REQUIRE(n_modes == 0 || mode_amp, "mode_amp is NULL");
        check_particles(p);
        load_species(*ctx, sp, nspecies, *rp, mode_amp, n_modes, *p, weights_out, S(s));
    });
}

aw_status aw_load_legacy(aw_ctx* ctx, const aw_run_params* rp, aw_particles* p, aw_stream s) {
    return guard("aw_load_legacy", [&] {
        AW_REQUIRE(ctx && rp, "ctx/rp is NULL");
        check_particles(p);
        load_legacy(*ctx, *rp, *p, S(s));
    });
}

aw_status aw_deposit_rho(aw_ctx* ctx, const aw_particles* p, double qm, int32_t policy,
                         float* rho, aw_stream s) {
    return guard("aw_deposit_rho", [&] {
        AW_REQUIRE(ctx && rho, "ctx/rho is NULL");
        check_particles(p);
        AW_REQUIRE(policy >= AW_DEPOSIT_GLOBAL && policy <= AW_DEPOSIT_AUTO, "bad policy");
        ctx->acc.zero(S(s));
        deposit_rho_fixed(*ctx, *p, qm, policy, nullptr, S(s));
        fix_to_float(*ctx, rho, false, S(s));
    });
}

aw_status aw_deposit_esirkepov(aw_ctx* ctx, const float* x0, const float* y0,
                               const aw_particles* p, double qm, double dt, int32_t policy,
                               float* jx, float* jy, float* jz, aw_stream s) {
    return guard("aw_deposit_esirkepov", [&] {
        AW_REQUIRE(ctx && jx && jy && jz, "ctx/jx/jy/jz is NULL");
        check_particles(p);
        AW_REQUIRE(p->n == 0 || (x0 && y0), "x0/y0 is NULL");
        AW_REQUIRE(policy >= AW_DEPOSIT_GLOBAL && policy <= AW_DEPOSIT_AUTO, "bad policy");
        deposit_esirkepov(*ctx, x0, y0, *p, qm, dt, policy, jx, jy, jz, S(s));
    });
}

aw_status aw_push_es(aw_ctx* ctx, aw_particles* p, const float* Ex, const float* Ey,
                     const aw_run_params* rp, int32_t has_b0, aw_stream s) {
    return guard("aw_push_es", [&] {
        AW_REQUIRE(ctx && Ex && Ey && rp, "ctx/E/rp is NULL");
        check_particles(p);
        push_es(*ctx, *p, Ex, Ey, *rp, has_b0 != 0, false, S(s));
    });
}

aw_status aw_half_step_back(aw_ctx* ctx, aw_particles* p, const float* Ex, const float* Ey,
                            const aw_run_params* rp, int32_t has_b0, aw_stream s) {
    return guard("aw_half_step_back", [&] {
        AW_REQUIRE(ctx && Ex && Ey && rp, "ctx/E/rp is NULL");
        check_particles(p);
        push_es(*ctx, *p, Ex, Ey, *rp, has_b0 != 0, true, S(s));
    });
}

aw_status aw_push_deposit_es(aw_ctx* ctx, aw_particles* p, const int32_t* off,
                             const float* Ex, const float* Ey, const aw_run_params* rp,
                             int32_t has_b0, float* rho_next, int64_t* strays_out,
                             aw_stream s) {
    return guard("aw_push_deposit_es", [&] {
        AW_REQUIRE(ctx && off && Ex && Ey && rp && rho_next, "NULL argument");
        check_particles(p);
        DevBuf<unsigned long long> cnt(1);
        cnt.zero(S(s));
        ctx->acc.zero(S(s));
        push_deposit_es(*ctx, *p, off, Ex, Ey, *rp, has_b0 != 0, cnt.get(), nullptr, 0, S(s));
        fix_to_float(*ctx, rho_next, false, S(s));
        unsigned long long h = 0;
        AW_HIP(hipMemcpyAsync(&h, cnt.get(), sizeof(h), hipMemcpyDeviceToHost, S(s)));
        AW_HIP(hipStreamSynchronize(S(s)));
        if (strays_out) *strays_out = (int64_t)h;
    });
}

aw_status aw_fft_r2c(aw_ctx* ctx, const float* in, void* out_k, aw_stream s) {
    return guard("aw_fft_r2c", [&] {
        AW_REQUIRE(ctx && in && out_k, "NULL argument");
        // rocFFT may overwrite an out-of-place input; keep the caller's array intact
        ctx->rin.reserve((size_t)ncell(*ctx));
        AW_HIP(hipMemcpyAsync(ctx->rin.get(), in, (size_t)ncell(*ctx) * sizeof(float),
                              hipMemcpyDeviceToDevice, S(s)));
        fft_exec(*ctx, ctx->r2c, ctx->rin.get(), out_k, S(s));
    });
}

aw_status aw_fft_c2r(aw_ctx* ctx, void* in_k, float* out, aw_stream s) {
    return guard("aw_fft_c2r", [&] {
        AW_REQUIRE(ctx && in_k && out, "NULL argument");
        fft_exec(*ctx, ctx->c2r, in_k, out, S(s));
    });
}

aw_status aw_solve_es(aw_ctx* ctx, const float* rho, float* Ex, float* Ey, double eps0,
                      aw_stream s) {
    return guard("aw_solve_es", [&] {
        AW_REQUIRE(ctx && rho && Ex && Ey, "NULL argument");
        AW_REQUIRE(eps0 > 0, "eps0 must be positive");
        solve_es(*ctx, rho, Ex, Ey, eps0, S(s));
    });
}

aw_status aw_solve_darwin_el_b(aw_ctx* ctx, const float* src4, float* out5, double eps0,
                               double c, const float B0[3], aw_stream s) {
    return guard("aw_solve_darwin_el_b", [&] {
        AW_REQUIRE(ctx && src4 && out5, "NULL argument");
        AW_REQUIRE(eps0 > 0 && c > 0, "eps0 and c must be positive");
        const float zero[3] = {0.f, 0.f, 0.f};
        enable_em_fft(*ctx);
        solve_darwin_elb(*ctx, src4, out5, eps0, c, B0 ? B0 : zero, S(s), false, 1);
    });
}

aw_status aw_solve_darwin_el_b_ex(aw_ctx* ctx, const float* src4, float* out5, double eps0,
                                  double c, const float B0[3], int32_t flags, aw_stream s) {
    return guard("aw_solve_darwin_el_b_ex", [&] {
        AW_REQUIRE(ctx && src4 && out5, "NULL argument");
        AW_REQUIRE(eps0 > 0 && c > 0, "eps0 and c must be positive");
        AW_REQUIRE((flags & ~AW_J_STAGGERED) == 0, "unknown flags");
        const float zero[3] = {0.f, 0.f, 0.f};
        enable_em_fft(*ctx);
        solve_darwin_elb(*ctx, src4, out5, eps0, c, B0 ? B0 : zero, S(s),
                         (flags & AW_J_STAGGERED) != 0, 1);
    });
}

aw_status aw_solve_darwin_et(aw_ctx* ctx, const float* src7, float* out3, double eps0,
                             double c, double n0, aw_stream s) {
    return guard("aw_solve_darwin_et", [&] {
        AW_REQUIRE(ctx && src7 && out3, "NULL argument");
        AW_REQUIRE(eps0 > 0 && c > 0, "eps0 and c must be positive");
        enable_em_fft(*ctx);
        solve_darwin_et(*ctx, src7, out3, eps0, c, n0, S(s), 1);
    });
}

aw_status aw_form_e(aw_ctx* ctx, aw_fields* f, const aw_run_params* rp, double t,
                    aw_stream s) {
    return guard("aw_form_e", [&] {
        AW_REQUIRE(ctx && f && rp, "NULL argument");
        AW_REQUIRE(f->Ex && f->Ey && f->Ez && f->ELx && f->ELy, "E / E_L grids must be non-NULL");
        AW_REQUIRE(!f->ETx || (f->ETy && f->ETz), "E_T must be all NULL or all non-NULL");
        form_e(FormE{{f->Ex, f->Ey, f->Ez}, {f->ELx, f->ELy}, {f->ETx, f->ETy, f->ETz}},
               ctx->g, *rp, t, S(s), 1);
    });
}

namespace {
void fields_in(const aw_fields* f, const float* F[6]) {
    AW_REQUIRE(f && f->Ex && f->Ey && f->Ez && f->Bx && f->By && f->Bz,
               "fields: Ex, Ey, Ez, Bx, By, Bz must be non-NULL");
    F[0] = f->Ex;
    F[1] = f->Ey;
    F[2] = f->Ez;
    F[3] = f->Bx;
    F[4] = f->By;
    F[5] = f->Bz;
}
}  // namespace

aw_status aw_deposit_rho_j_amu(aw_ctx* ctx, const aw_particles* p, double qm, aw_sources* out,
                               aw_stream s) {
    return guard("aw_deposit_rho_j_amu", [&] {
        AW_REQUIRE(ctx && out, "NULL argument");
        check_particles(p);
        float* const o[8] = {out->rho, out->Jx, out->Jy, out->Jz,
                             out->amu0, out->amu1, out->amu2, out->amu3};
        bool any = false;
        for (float* q : o) any = any || q;
        AW_REQUIRE(any, "sources: no rho, J or amu grid given");
        deposit_rho_j_amu(*ctx, *p, qm, o, S(s));
    });
}

aw_status aw_deposit_dcu(aw_ctx* ctx, const aw_particles* p, double qm, const aw_fields* f,
                         aw_sources* out, aw_stream s) {
    return guard("aw_deposit_dcu", [&] {
        AW_REQUIRE(ctx && out, "NULL argument");
        check_particles(p);
        const float* F[6];
        fields_in(f, F);
        float* const o[3] = {out->dcux, out->dcuy, out->dcuz};
        for (float* q : o) AW_REQUIRE(q, "sources: dcu grids must be non-NULL");
        deposit_dcu(*ctx, *p, qm, F, o, S(s));
    });
}

aw_status aw_push_boris_em(aw_ctx* ctx, aw_particles* p, const aw_fields* f,
                           const aw_run_params* rp, aw_stream s) {
    return guard("aw_push_boris_em", [&] {
        AW_REQUIRE(ctx && rp, "NULL argument");
        check_particles(p);
        const float* F[6];
        fields_in(f, F);
        push_boris_em(*ctx, *p, rp->qm, rp->dt, F, S(s));
    });
}

aw_status aw_reduce_energies(aw_ctx* ctx, const aw_particles* p, const float* Ex,
                             const float* Ey, const float* rho, double eps0,
                             aw_diag_sample* out, aw_stream s) {
    return guard("aw_reduce_energies", [&] {
        AW_REQUIRE(ctx && Ex && Ey && out, "NULL argument");
        check_particles(p);
        double* part = ctx->red.get();
        const int nke = p->n ? ke_partials(*p, part, S(s)) : 0;
        double* gp = part + reduce_grid_offset();
        double* o = part + reduce_out_offset();
        const int ng = grid_partials(Ex, Ey, rho, ncell(*ctx), gp, S(s));
        finalize_diag(part, nke, p->w ? 1.0 : p->w_uniform, gp, ng, o, S(s), nullptr);
        double h[4];
        AW_HIP(hipMemcpyAsync(h, o, sizeof(h), hipMemcpyDeviceToHost, S(s)));
        AW_HIP(hipStreamSynchronize(S(s)));
        const double dxdy = ctx->g.dx * ctx->g.dy;
        out->step = 0;
        out->time = 0.0;
        out->ke = 0.5 * h[0];
        out->ee = 0.5 * eps0 * h[1] * dxdy;
        out->max_e = std::sqrt(h[2]);
        out->charge = h[3] * dxdy;
        out->total = out->ke + out->ee;
    });
}

// ---- simulation ------------------------------------------------------------------

// A validated simulation with its context and stream (buffers: finish_sim).
static std::unique_ptr<aw_sim> make_sim(const aw_grid* g, const aw_run_params* rp,
                                        const aw_species* sp, int32_t nspecies,
                                        const aw_sim_config* cfg) {
    AW_REQUIRE(rp && cfg, "rp/cfg is NULL");
    check_grid(g);
    AW_REQUIRE(cfg->field_model == AW_MODEL_ES || cfg->field_model == AW_MODEL_DARWIN,
               "field_model must be AW_MODEL_ES or AW_MODEL_DARWIN");
    if (cfg->field_model == AW_MODEL_DARWIN)
        AW_REQUIRE(rp->c > 0 && rp->n0 >= 0, "Darwin needs rp.c > 0 and rp.n0 >= 0");
    AW_REQUIRE(rp->dt > 0, "rp.dt must be positive");
    AW_REQUIRE(rp->eps0 > 0, "rp.eps0 must be positive");
    AW_REQUIRE(cfg->current_deposit == AW_CURRENT_CIC ||
                   (cfg->current_deposit == AW_CURRENT_ESIRKEPOV &&
                    cfg->field_model == AW_MODEL_DARWIN),
               "current_deposit must be AW_CURRENT_CIC, or AW_CURRENT_ESIRKEPOV with "
               "the Darwin model");
    auto sim = std::make_unique<aw_sim>();
    sim->g = *g;
    if (sim->g.Lx <= 0) sim->g.Lx = g->nx * g->dx;
    if (sim->g.Ly <= 0) sim->g.Ly = g->ny * g->dy;
    sim->rp = *rp;
    sim->cfg = *cfg;
    AW_REQUIRE(cfg->n_modes >= 0 && cfg->n_modes <= 32, "n_modes must be in [0, 32]");
    const int64_t nc = (int64_t)g->nx * g->ny;
    if (cfg->legacy_load) {
        AW_REQUIRE(rp->ppc > 0, "legacy load needs rp.ppc > 0");
        sim->n = (int64_t)rp->ppc * nc;
        sim->w_uniform = rp->weight > 0 ? rp->weight : rp->n0 * g->dx * g->dy / rp->ppc;
        sim->rp.weight = sim->w_uniform;
    } else {
        AW_REQUIRE(sp && nspecies > 0, "at least one species is required");
        sim->species.assign(sp, sp + nspecies);
        int64_t total = 0;
        double w0 = -1.0;
        bool uniform = true;
        for (int k = 0; k < nspecies; ++k) {
            AW_REQUIRE(sp[k].ppc > 0, "species ppc must be positive");
            total += (int64_t)sp[k].ppc * nc;
            const double wk = sp[k].density * g->dx * g->dy / sp[k].ppc;
            if (w0 < 0) w0 = wk;
            else if (wk != w0) uniform = false;
        }
        sim->n = total;
        sim->w_uniform = w0;
        sim->per_w = !uniform;
    }
    AW_REQUIRE(sim->n < (int64_t)INT32_MAX, "particle count must fit int32");
    AW_HIP(hipStreamCreateWithFlags(&sim->stream, hipStreamNonBlocking));
    // default ES tile: 8x8 at >= 64 particles per cell (more, smaller
    // workgroups; the apron overhead stays small), else 16x16
    int tx = cfg->tile_x, ty = cfg->tile_y;
    if (tx <= 0 && ty <= 0 && cfg->field_model != AW_MODEL_DARWIN)
        tx = ty = sim->n >= 64 * nc ? 8 : 16;
    sim->ctx.reset(make_ctx(&sim->g, tx, ty));
    sim->ctx->fix_log2 = cfg->fixed_scale > 0
                             ? cfg->fixed_scale
                             : fixed_scale_for(sim->species, *rp, cfg->legacy_load != 0);
    if (cfg->fixed_scale <= 0) {
        double wmax = cfg->legacy_load ? std::fabs(rp->weight) : 0.0;
        for (const auto& s : sim->species)
            wmax = std::max(wmax, std::fabs(s.density) * g->dx * g->dy / s.ppc);
        sim->ctx->fix_log2 = rho_scale_guarded(*g, rp->qm, wmax, sim->n, sim->ctx->fix_log2);
    }
    double dens = cfg->legacy_load ? (rp->n0 > 0 ? rp->n0 : 1.0) : 0.0;
    for (const auto& s : sim->species) dens += s.density;
    sim->ctx->rho_base = rho_baseline(rp->qm, dens, sim->ctx->fix_log2);
    return sim;
}

// Buffers of a validated simulation for `runs` runs of its shape.
static std::unique_ptr<aw_sim> finish_sim(std::unique_ptr<aw_sim> sim, int runs) {
    const int64_t nc = ncell(*sim->ctx);
    sim->runs = runs;
    const size_t n = (size_t)std::max<int64_t>(1, sim->n);
    const size_t R = (size_t)runs;
    for (int b = 0; b < 2; ++b)
        for (int a = 0; a < (sim->per_w ? 6 : 5); ++a) sim->P[b][a].resize(R * n);
    sim->perm.resize(n);
    sim->rho.resize((size_t)nc);
    sim->E.resize((size_t)2 * nc);
    sim->strays.resize(1);
    sim->bpt = sim->cfg.blocks_per_tile > 0 ? sim->cfg.blocks_per_tile
                                            : tile_bpt(*sim->ctx, sim->n);
    sim->size_buffers();
    sim->grid_part.resize(R * aw_sim::kGridPart);
    sim->diag_out.resize(R * 4 * sim->diag_cap);
    sim->hist.assign(R, {});
    if (sim->cfg.field_model == AW_MODEL_DARWIN) sim->setup_em();
    AW_HIP(hipHostMalloc((void**)&sim->stray_host, sizeof(unsigned long long),
                         hipHostMallocDefault));
    AW_HIP(hipEventCreateWithFlags(&sim->stray_ev, hipEventDisableTiming));
    return sim;
}

}  // extern "C"

namespace aw {

// The autotuner's tile side and workgroups per tile for one electrostatic run
// (an ensemble member): a temporary simulation of the run is tuned as
// aw_sim_init tunes one -- loaded once per candidate tile, its fused kernel
// timed -- or its shape is found in the per-process cache.
void autotune_run(const aw_grid& g, const aw_run_params& rp, const aw_species* sp, int nsp,
                  bool has_b0, int& tile, int& bpt) {
    aw_sim_config c{};
    c.has_b0 = has_b0 ? 1 : 0;
    c.field_model = AW_MODEL_ES;
    c.autotune = 1;
    std::unique_ptr<aw_sim> sim = finish_sim(make_sim(&g, &rp, sp, nsp, &c), 1);
    sim->autotune();
    tile = sim->ctx->tx;
    bpt = sim->bpt;
}

}  // namespace aw

extern "C" {

aw_status aw_sim_create(const aw_grid* g, const aw_run_params* rp, const aw_species* sp,
                        int32_t nspecies, const aw_sim_config* cfg, aw_sim** out) {
    return guard("aw_sim_create", [&] {
        AW_REQUIRE(out, "out is NULL");
        *out = finish_sim(make_sim(g, rp, sp, nspecies, cfg), 1).release();
    });
}

aw_status aw_sim_destroy(aw_sim* sim) {
    return guard("aw_sim_destroy", [&] { delete sim; });
}

aw_status aw_sim_init(aw_sim* sim) {
    return guard("aw_sim_init", [&] {
        AW_REQUIRE(sim, "sim is NULL");
        sim->init();
    });
}

aw_status aw_sim_step(aw_sim* sim, int64_t n) {
    return guard("aw_sim_step", [&] {
        AW_REQUIRE(sim, "sim is NULL");
        sim->step(n);
    });
}

aw_status aw_sim_run(aw_sim* sim, int64_t nsteps) {
    return guard("aw_sim_run", [&] {
        AW_REQUIRE(sim, "sim is NULL");
        for (int64_t k = 0; k < nsteps; ++k) sim->step(k);
        AW_HIP(hipStreamSynchronize(sim->stream));
    });
}

aw_status aw_sim_synchronize(aw_sim* sim) {
    return guard("aw_sim_synchronize", [&] {
        AW_REQUIRE(sim, "sim is NULL");
        AW_HIP(hipStreamSynchronize(sim->stream));
    });
}

aw_status aw_sim_stream(aw_sim* sim, aw_stream* out) {
    return guard("aw_sim_stream", [&] {
        AW_REQUIRE(sim && out, "NULL argument");
        *out = (aw_stream)sim->stream;
    });
}

aw_status aw_sim_grid(aw_sim* sim, aw_grid* out) {
    return guard("aw_sim_grid", [&] {
        AW_REQUIRE(sim && out, "NULL argument");
        *out = sim->g;
    });
}

aw_status aw_sim_particles(aw_sim* sim, aw_particles* out) {
    return guard("aw_sim_particles", [&] {
        AW_REQUIRE(sim && out, "NULL argument");
        *out = sim->particles();
    });
}

aw_status aw_sim_fields(aw_sim* sim, aw_fields* out) {
    return guard("aw_sim_fields", [&] {
        AW_REQUIRE(sim && out, "NULL argument");
        std::memset(out, 0, sizeof(*out));
        out->Ex = sim->Ex();
        out->Ey = sim->Ey();
        if (sim->em) {
            const int64_t nc = ncell(*sim->ctx);
            out->Ez = sim->e3.get() + 2 * nc;
            out->Bx = sim->B3();
            out->By = sim->B3() + nc;
            out->Bz = sim->B3() + 2 * nc;
            out->ELx = sim->elb.get();
            out->ELy = sim->elb.get() + nc;
            out->ETx = sim->et.get();
            out->ETy = sim->et.get() + nc;
            out->ETz = sim->et.get() + 2 * nc;
        }
    });
}

aw_status aw_sim_sources(aw_sim* sim, aw_sources* out) {
    return guard("aw_sim_sources", [&] {
        AW_REQUIRE(sim && out, "NULL argument");
        std::memset(out, 0, sizeof(*out));
        out->rho = sim->rho_ptr();
        if (sim->em) {
            const int64_t nc = ncell(*sim->ctx);
            float* s = sim->src.get();
            out->Jx = s + nc;
            out->Jy = s + 2 * nc;
            out->Jz = s + 3 * nc;
            out->dcux = s + 4 * nc;
            out->dcuy = s + 5 * nc;
            out->dcuz = s + 6 * nc;
            out->amu0 = s + 7 * nc;
            out->amu1 = s + 8 * nc;
            out->amu2 = s + 9 * nc;
            out->amu3 = s + 10 * nc;
        }
    });
}

aw_status aw_sim_history(aw_sim* sim, aw_diag_sample* out, int64_t cap, int64_t* count) {
    return guard("aw_sim_history", [&] {
        AW_REQUIRE(sim && count, "NULL argument");
        sim->flush_diag();
        const int64_t nh = (int64_t)sim->hist[0].size();
        if (out)
            for (int64_t k = 0; k < std::min(cap, nh); ++k) out[k] = sim->hist[0][k];
        *count = nh;
    });
}

aw_status aw_sim_diag_now(aw_sim* sim, int64_t step, aw_diag_sample* out) {
    return guard("aw_sim_diag_now", [&] {
        AW_REQUIRE(sim && out, "NULL argument");
        const aw_particles p = sim->particles();
        aw_ctx& c = *sim->ctx;
        double* part = c.red.get();
        const int nke = p.n ? ke_partials(p, part, sim->stream) : 0;
        double* gp = part + reduce_grid_offset();
        double* o = part + reduce_out_offset();
        const int ng = grid_partials(sim->Ex(), sim->Ey(), sim->rho_ptr(), ncell(c), gp,
                                     sim->stream);
        finalize_diag(part, nke, p.w ? 1.0 : p.w_uniform, gp, ng, o, sim->stream, nullptr);
        double h[4];
        AW_HIP(hipMemcpyAsync(h, o, sizeof(h), hipMemcpyDeviceToHost, sim->stream));
        AW_HIP(hipStreamSynchronize(sim->stream));
        const double dxdy = sim->g.dx * sim->g.dy;
        out->step = step;
        out->time = step * sim->rp.dt;
        out->ke = 0.5 * h[0];
        out->ee = 0.5 * sim->rp.eps0 * h[1] * dxdy;
        out->max_e = std::sqrt(h[2]);
        out->charge = h[3] * dxdy;
        out->total = out->ke + out->ee;
    });
}

aw_status aw_sim_checkpoint_save(aw_sim* sim, const char* path, int64_t step, double time,
                                 const char* deck, const char* git_hash) {
    return guard("aw_sim_checkpoint_save", [&] {
        AW_REQUIRE(sim && path, "NULL argument");
        aw_checkpoint_header h{};
        h.version = 1;
        h.rng_seed = sim->rp.rng_seed;
        h.step = step;
        h.time = time;
        h.sd = 1.0;
        h.eps_l = 0.0;
        sim->save(path, h, deck, git_hash);
    });
}

aw_status aw_sim_checkpoint_restore(aw_sim* sim, const char* path) {
    return guard("aw_sim_checkpoint_restore", [&] {
        AW_REQUIRE(sim && path, "NULL argument");
        sim->restore(path);
    });
}

aw_status aw_sim_particles_changed(aw_sim* sim) {
    return guard("aw_sim_particles_changed", [&] {
        AW_REQUIRE(sim, "sim is NULL");
        sim->particles_changed();
    });
}

aw_status aw_sim_request_sort(aw_sim* sim) {
    return guard("aw_sim_request_sort", [&] {
        AW_REQUIRE(sim, "sim is NULL");
        sim->sort_requested = true;
    });
}

aw_status aw_sim_set_profiling(aw_sim* sim, int32_t on) {
    return guard("aw_sim_set_profiling", [&] {
        AW_REQUIRE(sim, "sim is NULL");
        sim->collect_prof();
        sim->prof = on != 0;
        sim->prof_ms = 0.0;
        sim->prof_launches = 0;
    });
}

aw_status aw_sim_kernel_time(aw_sim* sim, double* total_ms, int64_t* launches, int64_t* sorts) {
    return guard("aw_sim_kernel_time", [&] {
        AW_REQUIRE(sim, "sim is NULL");
        sim->collect_prof();
        if (total_ms) *total_ms = sim->prof_ms;
        if (launches) *launches = sim->prof_launches;
        if (sorts) *sorts = sim->nsorts;
    });
}

aw_status aw_sim_exec_config(aw_sim* sim, int32_t* tile_x, int32_t* tile_y,
                             int32_t* blocks_per_tile) {
    return guard("aw_sim_exec_config", [&] {
        AW_REQUIRE(sim, "sim is NULL");
        if (tile_x) *tile_x = sim->ctx->tx;
        if (tile_y) *tile_y = sim->ctx->ty;
        if (blocks_per_tile) *blocks_per_tile = sim->bpt;
    });
}

}  // extern "C"

// ============================================================================
// Darwin ensembles: one aw_sim stepping a uniform batch of runs
// ============================================================================

namespace aw {

// Runs of one Darwin batch share the grid, the species layout (count, ppc,
// density: equal weights) and every step parameter; they differ in their
// loads (thermal and drift velocities, seed, mode amplitudes).
aw_sim* darwin_batch_create(const aw_run_desc* runs, int nruns, const aw_sim_config* cfg) {
    AW_REQUIRE(runs && cfg, "NULL argument");
    AW_REQUIRE(nruns >= 1 && nruns <= 64, "a Darwin ensemble holds 1..64 runs");
    AW_REQUIRE(cfg->field_model == AW_MODEL_DARWIN, "not a Darwin configuration");
    AW_REQUIRE(!cfg->legacy_load, "Darwin ensembles load species (legacy_load = 0)");
    const aw_run_desc& r0 = runs[0];
    for (int k = 0; k < nruns; ++k) {
        const aw_run_desc& r = runs[k];
        AW_REQUIRE(r.species && r.nspecies > 0, "every run needs species");
        AW_REQUIRE(r.n_modes >= 0 && r.n_modes <= 32, "n_modes must be in [0, 32]");
        AW_REQUIRE(r.grid.nx == r0.grid.nx && r.grid.ny == r0.grid.ny &&
                       r.grid.dx == r0.grid.dx && r.grid.dy == r0.grid.dy,
                   "Darwin ensemble runs must share the grid");
        AW_REQUIRE(r.nspecies == r0.nspecies && r.has_b0 == r0.has_b0,
                   "Darwin ensemble runs must share the species layout and has_b0");
        for (int q = 0; q < r.nspecies; ++q)
            AW_REQUIRE(r.species[q].ppc == r0.species[q].ppc &&
                           r.species[q].density == r0.species[q].density,
                       "Darwin ensemble runs must share species ppc and density");
        const aw_run_params &a = r.rp, &b = r0.rp;
        AW_REQUIRE(a.dt == b.dt && a.qm == b.qm && a.eps0 == b.eps0 && a.c == b.c &&
                       a.n0 == b.n0 && a.B0[0] == b.B0[0] && a.B0[1] == b.B0[1] &&
                       a.B0[2] == b.B0[2] && a.pump == b.pump && a.pump_ex == b.pump_ex &&
                       a.pump_ey == b.pump_ey && a.pump_ez == b.pump_ez &&
                       a.pump_k0 == b.pump_k0 && a.pump_w0 == b.pump_w0 &&
                       a.pump_trmp == b.pump_trmp && a.pump_toff == b.pump_toff &&
                       a.dump_every == b.dump_every,
                   "Darwin ensemble runs must share dt, qm, eps0, c, n0, B0, the pump and "
                   "dump_every");
    }
    aw_sim_config c = *cfg;
    c.has_b0 = r0.has_b0;
    c.n_modes = 0;
    auto sim = make_sim(&r0.grid, &r0.rp, r0.species, r0.nspecies, &c);
    AW_REQUIRE((int64_t)sim->n * nruns < (int64_t)INT32_MAX,
               "ensemble particle count must fit int32");
    sim->rspecies.resize(nruns);
    sim->rrp.resize(nruns);
    sim->rmodes.resize(nruns);
    for (int k = 0; k < nruns; ++k) {
        sim->rspecies[k].assign(runs[k].species, runs[k].species + runs[k].nspecies);
        sim->rrp[k] = runs[k].rp;
        sim->rmodes[k].assign(runs[k].mode_amp, runs[k].mode_amp + runs[k].n_modes);
    }
    return finish_sim(std::move(sim), nruns).release();
}

void darwin_batch_destroy(aw_sim* s) noexcept { delete s; }
void darwin_batch_init(aw_sim* s) { s->init(); }
void darwin_batch_step(aw_sim* s, int64_t n) { s->step(n); }
void darwin_batch_sync(aw_sim* s) { AW_HIP(hipStreamSynchronize(s->stream)); }

void darwin_batch_info(aw_sim* s, aw_ensemble_stats* out) {
    std::memset(out, 0, sizeof(*out));
    out->nruns = s->runs;
    out->ngroups = 1;
    out->particles = s->n * s->runs;
    out->cells = ncell(*s->ctx) * s->runs;
    out->workgroups = s->ctx->ntiles * s->bpt * s->runs;
    out->sorts = s->nsorts;
    out->run_sorts = s->nsorts * s->runs;  // a uniform batch re-sorts every run
}

static int check_run(const aw_sim* s, int run) {
    AW_REQUIRE(run >= 0 && run < s->runs, "run index out of range");
    return run;
}

aw_particles darwin_batch_particles(const aw_sim* s, int run) {
    return s->particles(s->cur, check_run(s, run));
}

aw_fields darwin_batch_fields(const aw_sim* s, int run) {
    const size_t nc = (size_t)ncell(*s->ctx), r = (size_t)check_run(s, run);
    aw_fields f{};
    f.Ex = s->e3.get() + r * 3 * nc;
    f.Ey = f.Ex + nc;
    f.Ez = f.Ex + 2 * nc;
    f.ELx = s->elb.get() + r * 5 * nc;
    f.ELy = f.ELx + nc;
    f.Bx = f.ELx + 2 * nc;
    f.By = f.ELx + 3 * nc;
    f.Bz = f.ELx + 4 * nc;
    f.ETx = s->et.get() + r * 3 * nc;
    f.ETy = f.ETx + nc;
    f.ETz = f.ETx + 2 * nc;
    return f;
}

float* darwin_batch_rho(const aw_sim* s, int run) { return s->rho_ptr(check_run(s, run)); }

const std::vector<aw_diag_sample>& darwin_batch_history(aw_sim* s, int run) {
    s->flush_diag();
    return s->hist[check_run(s, run)];
}

}  // namespace aw
