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
#include <memory>

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
                   double* out4, hipStream_t s);
int reduce_partial_capacity();
int reduce_grid_offset();
int reduce_out_offset();
int tile_bpt(const aw_ctx& c, int64_t n);

namespace {

thread_local std::string g_last_error;

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

void fix_to_float(aw_ctx& c, float* out, bool assign, hipStream_t s,
                  unsigned long long* counter = nullptr) {
    fix_to_float_add(c.acc.get(), out, ncell(c), c.fix_log2, true, s, assign, counter);
}

void sort_tiles(aw_ctx& c, const aw_particles& src, const aw_particles& dst, int32_t* perm,
                int32_t* off, hipStream_t s) {
    if (tile_sort_onepass(c, src, dst, off, nullptr, s)) return;
    c.sort.keyA.reserve((size_t)std::max<int64_t>(1, src.n));
    compute_tile_keys(c, src, c.sort.keyA.get(), false, s);
    stable_sort_keys(c, src.n, (uint32_t)c.ntiles, perm, off, s);
    permute_particles(src, dst, perm, s);
}

}  // namespace

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
    std::vector<int64_t> diag_pending;  // steps with a slot in diag_out
    std::vector<aw_diag_sample> history;
    bool initialized = false, fields_ready = false, sort_requested = false;
    int64_t steps_since_sort = 0, nsorts = 0;
    unsigned long long* stray_host = nullptr;
    hipEvent_t stray_ev = nullptr;
    bool stray_pending = false;
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_ev;
    double prof_ms = 0.0;
    int64_t prof_launches = 0;
    int bpt = 0;

    aw_particles particles(int which) const {
        aw_particles p{};
        p.x = P[which][0].get();
        p.y = P[which][1].get();
        p.ux = P[which][2].get();
        p.uy = P[which][3].get();
        p.uz = P[which][4].get();
        p.w = per_w ? P[which][5].get() : nullptr;
        p.n = n;
        p.w_uniform = w_uniform;
        return p;
    }
    aw_particles particles() const { return particles(cur); }
    float* Ex() const { return E.get(); }
    float* Ey() const { return E.get() + ncell(*ctx); }

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
        const aw_particles src = particles(cur), dst = particles(cur ^ 1);
        sort_tiles(*ctx, src, dst, perm.get(), tile_off.get(), stream);
        cur ^= 1;
        steps_since_sort = 0;
        ++nsorts;
        sort_requested = false;
    }

    void flush_diag() {
        if (diag_pending.empty()) return;
        std::vector<double> h(diag_pending.size() * 4);
        AW_HIP(hipMemcpyAsync(h.data(), diag_out.get(), h.size() * sizeof(double),
                              hipMemcpyDeviceToHost, stream));
        AW_HIP(hipStreamSynchronize(stream));
        const double dxdy = g.dx * g.dy;
        for (size_t k = 0; k < diag_pending.size(); ++k) {
            aw_diag_sample d{};
            d.step = diag_pending[k];
            d.time = (double)d.step * rp.dt;
            d.ke = 0.5 * h[4 * k];
            d.ee = 0.5 * rp.eps0 * h[4 * k + 1] * dxdy;
            d.charge = h[4 * k + 3] * dxdy;
            d.max_e = std::sqrt(h[4 * k + 2]);
            d.total = d.ke + d.ee;
            history.push_back(d);
        }
        diag_pending.clear();
    }

    void poll_strays() {
        if (!stray_pending) return;
        if (hipEventQuery(stray_ev) != hipSuccess) return;
        stray_pending = false;
        const double frac = cfg.stray_frac > 0 ? cfg.stray_frac : 0.02;
        if ((double)*stray_host > frac * (double)n) sort_requested = true;
    }

    void init() {
        const hipStream_t s = stream;
        const aw_particles p = particles(cur);
        if (cfg.legacy_load) {
            load_legacy(*ctx, rp, p, s);
        } else {
            std::vector<double> wts(species.size());
            load_species(*ctx, species.data(), (int)species.size(), rp, cfg.mode_amp,
                         cfg.n_modes, p, wts.data(), s);
        }
        do_sort();
        ctx->acc.zero(s);
        deposit_rho_fixed(*ctx, particles(), rp.qm, AW_DEPOSIT_BINNED, tile_off.get(), s);
        fix_to_float(*ctx, rho.get(), true, s);
        solve_es_contig(*ctx, rho.get(), E.get(), rp.eps0, s);
        push_es(*ctx, particles(), Ex(), Ey(), rp, cfg.has_b0 != 0, true, s);
        strays.zero(s);  // afterwards cleared by each step's rho conversion
        fields_ready = true;
        initialized = true;
        AW_HIP(hipStreamSynchronize(s));
    }

    void step(int64_t nstep) {
        if (!initialized) fail(AW_ERR_STATE, "aw_sim_step before aw_sim_init", __FILE__, __LINE__);
        const hipStream_t s = stream;
        poll_strays();
        if (!fields_ready) {
            fix_to_float(*ctx, rho.get(), true, s, strays.get());
            solve_es_contig(*ctx, rho.get(), E.get(), rp.eps0, s);
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
        }
        if (dump) {
            if ((int)diag_pending.size() >= diag_cap) flush_diag();
            const int ng = grid_partials(Ex(), Ey(), rho.get(), ncell(*ctx), grid_part.get(), s);
            finalize_diag(ke_part.get(), nblk, per_w ? 1.0 : w_uniform, grid_part.get(), ng,
                          diag_out.get() + 4 * diag_pending.size(), s);
            diag_pending.push_back(nstep);
        }
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

namespace {

aw_ctx* make_ctx(const aw_grid* g, int tile_x, int tile_y) {
    check_grid(g);
    auto c = std::make_unique<aw_ctx>();
    c->g = *g;
    if (c->g.Lx <= 0) c->g.Lx = g->nx * g->dx;
    if (c->g.Ly <= 0) c->g.Ly = g->ny * g->dy;
    c->tx = std::min(tile_x > 0 ? tile_x : 16, g->nx);
    c->ty = std::min(tile_y > 0 ? tile_y : 16, g->ny);
    c->ntx = (g->nx + c->tx - 1) / c->tx;
    c->nty = (g->ny + c->ty - 1) / c->ty;
    c->ntiles = c->ntx * c->nty;
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

int fixed_scale_for(const std::vector<aw_species>& sp, const aw_run_params& rp, bool legacy) {
    // rho magnitude ~ |qm| * total density; keep 2^23 headroom over it
    double dens = 0.0;
    if (legacy) dens = rp.n0 > 0 ? rp.n0 : 1.0;
    for (const auto& s : sp) dens += std::fabs(s.density);
    const double rmax = std::max(1e-30, std::fabs(rp.qm) * std::max(dens, 1e-30));
    const int e = (int)std::ceil(std::log2(rmax));
    return std::clamp(40 - e, 8, 60);
}

}  // namespace

extern "C" {

void aw_last_error(char* buf, size_t len) {
    if (!buf || !len) return;
    std::strncpy(buf, last_error().c_str(), len - 1);
    buf[len - 1] = 0;
}

const char* aw_version(void) { return "arcwarden-amd 0.1 gfx950"; }

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
        ctx->sort.keyA.reserve((size_t)std::max<int64_t>(1, p->n));
        compute_tile_keys(*ctx, *p, ctx->sort.keyA.get(), false, S(s));
        stable_sort_keys(*ctx, p->n, (uint32_t)ctx->ntiles, idx, off, S(s));
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
        fft_exec(*ctx, ctx->r2c, const_cast<float*>(in), out_k, S(s));
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
        finalize_diag(part, nke, p->w ? 1.0 : p->w_uniform, gp, ng, o, S(s));
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

aw_status aw_sim_create(const aw_grid* g, const aw_run_params* rp, const aw_species* sp,
                        int32_t nspecies, const aw_sim_config* cfg, aw_sim** out) {
    return guard("aw_sim_create", [&] {
        AW_REQUIRE(out && rp && cfg, "out/rp/cfg is NULL");
        check_grid(g);
        AW_REQUIRE(cfg->field_model == AW_MODEL_ES,
                   "only the electrostatic field model is available in this build");
        AW_REQUIRE(rp->dt > 0, "rp.dt must be positive");
        AW_REQUIRE(rp->eps0 > 0, "rp.eps0 must be positive");
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
        sim->ctx.reset(make_ctx(&sim->g, cfg->tile_x, cfg->tile_y));
        sim->ctx->fix_log2 = cfg->fixed_scale > 0
                                 ? cfg->fixed_scale
                                 : fixed_scale_for(sim->species, *rp, cfg->legacy_load != 0);
        const size_t n = (size_t)std::max<int64_t>(1, sim->n);
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < (sim->per_w ? 6 : 5); ++a) sim->P[b][a].resize(n);
        sim->tile_off.resize((size_t)sim->ctx->ntiles + 1);
        sim->perm.resize(n);
        sim->rho.resize((size_t)nc);
        sim->E.resize((size_t)2 * nc);
        sim->strays.resize(1);
        sim->bpt = cfg->blocks_per_tile > 0 ? cfg->blocks_per_tile : tile_bpt(*sim->ctx, sim->n);
        sim->ke_part.resize((size_t)sim->ctx->ntiles * sim->bpt + 1);
        sim->grid_part.resize((size_t)3 * 512 + 3);
        sim->diag_out.resize((size_t)4 * sim->diag_cap);
        AW_HIP(hipHostMalloc((void**)&sim->stray_host, sizeof(unsigned long long),
                             hipHostMallocDefault));
        AW_HIP(hipEventCreateWithFlags(&sim->stray_ev, hipEventDisableTiming));
        *out = sim.release();
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
    });
}

aw_status aw_sim_sources(aw_sim* sim, aw_sources* out) {
    return guard("aw_sim_sources", [&] {
        AW_REQUIRE(sim && out, "NULL argument");
        std::memset(out, 0, sizeof(*out));
        out->rho = sim->rho.get();
    });
}

aw_status aw_sim_history(aw_sim* sim, aw_diag_sample* out, int64_t cap, int64_t* count) {
    return guard("aw_sim_history", [&] {
        AW_REQUIRE(sim && count, "NULL argument");
        sim->flush_diag();
        const int64_t nh = (int64_t)sim->history.size();
        if (out)
            for (int64_t k = 0; k < std::min(cap, nh); ++k) out[k] = sim->history[k];
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
        const int ng = grid_partials(sim->Ex(), sim->Ey(), sim->rho.get(), ncell(c), gp,
                                     sim->stream);
        finalize_diag(part, nke, p.w ? 1.0 : p.w_uniform, gp, ng, o, sim->stream);
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

}  // extern "C"

