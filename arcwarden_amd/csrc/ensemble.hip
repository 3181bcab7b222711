// ensemble.hip -- batched electrostatic ensembles: many independent runs
// stepped together.
//
// No reference counterpart: the reference runs one Simulation per process
// (include/pic/simulation.hpp:38-140) and caches device attributes in
// statics keyed to the first grid (include/pic/depositor.hpp:450-484).  Here
// every run keeps its own grid, RunParams, species and tile geometry, and a
// step of the whole ensemble is
//   1 launch   fixed-point rho^n -> float for every run (concatenated grids),
//   3 launches per grid size: R2C over the size's runs, k-space, C2R (Ex|Ey);
//              the sizes are independent and run concurrently, on the main
//              stream and up to kSideStreams side streams (fork/join events),
//   1 launch   fused gather + Boris + move + wrap + deposit over every run
//              (each workgroup finds its run in a device-resident run table),
// with the stable tile re-sort of every run on a cadence, and of each run
// whose own particles strayed from their tile apron beyond a fraction of its
// particle count (per-run stray counters; the other runs keep their order
// and buffers).  The step can
// be captured once into a HIP graph and replayed (two graphs: with and
// without the diagnostics reduction); the run table lives in device memory,
// so the re-sort's buffer swap only rewrites the table, not the graphs.
#include "aw_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

namespace aw {

namespace {

constexpr int kThreads = 256;
// streams beside the main one for the per-size solves (the process has four
// hardware queues)
#ifndef AW_ENSEMBLE_SIDE_STREAMS
#define AW_ENSEMBLE_SIDE_STREAMS 2
#endif
constexpr int kSideStreams = AW_ENSEMBLE_SIDE_STREAMS;

inline int blocks_for(int64_t n) { return (int)((n + kThreads - 1) / kThreads); }

struct DiagRun {
    int wg_begin, nwg;
    long long cbase;
    int ncell;
    double w_uniform;
};

// per run: KE partial sum, sum |E|^2, max |E|^2, sum rho -> out[r][4]
__global__ void __launch_bounds__(kThreads)
ens_diag_kernel(const DiagRun* __restrict__ runs, const double* __restrict__ ke_part,
                const float* __restrict__ E, const float* __restrict__ rho,
                double* __restrict__ out) {
    const DiagRun d = runs[blockIdx.x];
    double ke = 0.0, e2 = 0.0, q = 0.0, m2 = 0.0;
    for (int w = threadIdx.x; w < d.nwg; w += kThreads) ke += ke_part[d.wg_begin + w];
    const float* ex = E + 2 * d.cbase;
    const float* ey = ex + d.ncell;
    for (int c = threadIdx.x; c < d.ncell; c += kThreads) {
        const double v = (double)ex[c] * ex[c] + (double)ey[c] * ey[c];
        e2 += v;
        m2 = fmax(m2, v);
        q += (double)rho[d.cbase + c];
    }
    __shared__ double sh[4][kThreads];
    sh[0][threadIdx.x] = ke;
    sh[1][threadIdx.x] = e2;
    sh[2][threadIdx.x] = m2;
    sh[3][threadIdx.x] = q;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
            sh[2][threadIdx.x] = fmax(sh[2][threadIdx.x], sh[2][threadIdx.x + o]);
            sh[3][threadIdx.x] += sh[3][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[4 * blockIdx.x + 0] = sh[0][0] * d.w_uniform;
        out[4 * blockIdx.x + 1] = sh[1][0];
        out[4 * blockIdx.x + 2] = sh[2][0];
        out[4 * blockIdx.x + 3] = sh[3][0];
    }
}

}  // namespace

}  // namespace aw

using namespace aw;

namespace aw {
// Darwin ensembles (capi.hip): a batched aw_sim
aw_sim* darwin_batch_create(const aw_run_desc* runs, int nruns, const aw_sim_config* cfg);
void darwin_batch_destroy(aw_sim* s) noexcept;
void darwin_batch_init(aw_sim* s);
void darwin_batch_step(aw_sim* s, int64_t n);
void darwin_batch_sync(aw_sim* s);
void darwin_batch_info(aw_sim* s, aw_ensemble_stats* out);
aw_particles darwin_batch_particles(const aw_sim* s, int run);
aw_fields darwin_batch_fields(const aw_sim* s, int run);
float* darwin_batch_rho(const aw_sim* s, int run);
const std::vector<aw_diag_sample>& darwin_batch_history(aw_sim* s, int run);
}  // namespace aw

struct aw_ensemble {
    // the Darwin model: a uniform batch stepped by one aw_sim (every call
    // below forwards to it); the members after it serve the ES model
    aw_sim* dar = nullptr;
    bool dar_initialized = false;

    struct Run {
        int user = 0;  // index in the caller's order
        aw_grid g{};
        aw_run_params rp{};
        std::vector<aw_species> species;
        std::vector<double> modes;
        std::unique_ptr<aw_ctx> ctx;  // tile geometry, sort scratch, init solve
        int64_t n = 0, pbase = 0, cbase = 0, obase = 0;
        double w_uniform = 0.0;
        bool has_b0 = false;
        int bpt = 0, wg_begin = 0, nwg = 0;
        int cur = 0;            // particle buffer (P[cur]) holding the run
        bool sort_req = false;  // strays beyond the threshold: re-sort
        std::vector<aw_diag_sample> history;
    };
    struct Group {
        int first = 0, count = 0;  // runs [first, first + count) share nx, ny
        int64_t cbase = 0;
    };

    std::vector<Run> runs;  // grouped by grid size
    std::vector<Group> groups;
    aw_sim_config cfg{};
    hipStream_t stream = nullptr;
    int64_t ntot = 0, nctot = 0;
    int total_wg = 0, fix_log2 = 40;
    double eps0 = 1.0;
    bool uz = false;
    size_t lds = 0;
    DevBuf<float> P[2][5];
    std::vector<hipStream_t> side;         // concurrent per-size solves
    std::vector<hipEvent_t> fork_join;     // [0] fork, [1 + i] join of side[i]
    DevBuf<int32_t> tile_off, perm;
    DevBuf<int32_t> wg_run;  // workgroup -> run (slot) of the multi-run launch
    DevBuf<unsigned long long> acc, strays;
    DevBuf<float> rho, rin, E;
    DevBuf<long long> rbase;  // per cell: its run's mean charge density (fixed point)
    DevBuf<double> ke_part, diag_stage, diag_ring;
    DevBuf<unsigned char> slots[2], active;
    DevBuf<DiagRun> diag_runs;
    int diag_cap = 64;
    std::vector<int64_t> diag_pending;
    bool initialized = false, fields_ready = false;
    int64_t steps_since_sort = 0, nsorts = 0, run_sorts = 0;
    bool use_graph = false;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // [dump]
    int64_t graph_launches = 0;
    // per-run stray counts, read back asynchronously, at most kStrayLag
    // steps late
    unsigned long long* stray_host = nullptr;
    hipEvent_t stray_ev = nullptr;
    bool stray_pending = false, sort_requested = false;
    int stray_age = 0;  // steps since the pending stray read-back was issued

    ~aw_ensemble() {
        if (dar) darwin_batch_destroy(dar);
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& gx : graph)
            if (gx) (void)hipGraphExecDestroy(gx);
        if (stray_ev) (void)hipEventDestroy(stray_ev);
        if (stray_host) (void)hipHostFree(stray_host);
        for (hipEvent_t ev : fork_join) (void)hipEventDestroy(ev);
        for (hipStream_t ss : side) {
            (void)hipStreamSynchronize(ss);
            (void)hipStreamDestroy(ss);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Asynchronous stray read-back, waited for once it is kStrayLag steps old so
    // that the re-sort decision never trails the device by more than that.
    static constexpr int kStrayLag = 2;

    void poll_strays() {
        if (!stray_pending) return;
        if (++stray_age >= kStrayLag) AW_HIP(hipEventSynchronize(stray_ev));
        else if (hipEventQuery(stray_ev) != hipSuccess) return;
        stray_pending = false;
        const double frac = cfg.stray_frac > 0 ? cfg.stray_frac : 0.02;
        for (size_t k = 0; k < runs.size(); ++k)
            if ((double)stray_host[k] > frac * (double)runs[k].n) {
                runs[k].sort_req = true;
                sort_requested = true;
            }
    }

    aw_particles run_particles(const Run& r, int which) const {
        aw_particles p{};
        p.x = P[which][0].get() + r.pbase;
        p.y = P[which][1].get() + r.pbase;
        p.ux = P[which][2].get() + r.pbase;
        p.uy = P[which][3].get() + r.pbase;
        p.uz = P[which][4].get() + r.pbase;
        p.w = nullptr;
        p.n = r.n;
        p.w_uniform = r.w_uniform;
        return p;
    }

    void build_slots() {
        const size_t sb = run_slot_bytes();
        for (int which = 0; which < 2; ++which) {
            std::vector<unsigned char> h(sb * runs.size());
            for (size_t k = 0; k < runs.size(); ++k) {
                Run& r = runs[k];
                fill_run_slot(h.data() + sb * k, *r.ctx, run_particles(r, which),
                              tile_off.get() + r.obase, E.get() + 2 * r.cbase,
                              E.get() + 2 * r.cbase + ncell(*r.ctx), r.rp, r.has_b0,
                              acc.get() + r.cbase, strays.get() + k, ke_part.get() + r.wg_begin,
                              r.bpt, r.wg_begin);
            }
            slots[which].resize(h.size());
            AW_HIP(hipMemcpy(slots[which].get(), h.data(), h.size(), hipMemcpyHostToDevice));
        }
        active.resize(sb * runs.size());
        for (size_t k = 0; k < runs.size(); ++k) activate(k);
    }

    // the run table entry of run k <- its slot for the buffer that holds it
    void activate(size_t k) {
        const size_t sb = run_slot_bytes();
        AW_HIP(hipMemcpyAsync(active.get() + sb * k, slots[runs[k].cur].get() + sb * k, sb,
                              hipMemcpyDeviceToDevice, stream));
    }

    // Stable tile re-sort of every run (all) or of the runs whose strays
    // crossed the threshold; a sorted run moves to its other buffer.
    void sort_runs(bool all) {
        bool any = false;
        for (size_t k = 0; k < runs.size(); ++k) {
            Run& r = runs[k];
            if (!all && !r.sort_req) continue;
            sort_tiles(*r.ctx, run_particles(r, r.cur), run_particles(r, r.cur ^ 1), perm.get(),
                       tile_off.get() + r.obase, stream);
            r.cur ^= 1;
            r.sort_req = false;
            activate(k);
            ++run_sorts;
            any = true;
        }
        sort_requested = false;
        if (!any) return;
        steps_since_sort = 0;
        ++nsorts;
    }

    void convert_and_solve() {
        // rho^n for every run; rin receives the R2C input copy, strays reset
        fix_to_float_add(acc.get(), rho.get(), nctot, fix_log2, true, stream, true, strays.get(),
                         rin.get(), (int)runs.size(), rbase.get());
        // one batched solve per grid size; the sizes touch disjoint buffers
        // (each its own plans, work area and k-space scratch), so they fork
        // onto the side streams and join before the push
        const size_t nside = std::min(side.size(), groups.size() - 1);
        if (nside) {
            AW_HIP(hipEventRecord(fork_join[0], stream));
            for (size_t i = 0; i < nside; ++i)
                AW_HIP(hipStreamWaitEvent(side[i], fork_join[0], 0));
        }
        for (size_t g = 0; g < groups.size(); ++g) {
            const Group& gr = groups[g];
            const size_t lane = nside ? g % (nside + 1) : 0;
            solve_es_batch(*runs[gr.first].ctx, gr.count, rin.get() + gr.cbase,
                           E.get() + 2 * gr.cbase, eps0, lane ? side[lane - 1] : stream);
        }
        for (size_t i = 0; i < nside; ++i) {
            AW_HIP(hipEventRecord(fork_join[1 + i], side[i]));
            AW_HIP(hipStreamWaitEvent(stream, fork_join[1 + i], 0));
        }
    }

    void enqueue_step(bool dump) {
        convert_and_solve();
        push_deposit_multi(active.get(), wg_run.get(), total_wg, lds, dump, uz, stream);
        if (dump) {
            ens_diag_kernel<<<(int)runs.size(), kThreads, 0, stream>>>(
                diag_runs.get(), ke_part.get(), E.get(), rho.get(), diag_stage.get());
            AW_LAUNCH_CHECK();
        }
    }

    void capture_graphs() {
        for (int d = 0; d < 2; ++d) {
            hipGraph_t g = nullptr;
            AW_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_step(d == 1);
            } catch (...) {
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            AW_HIP(hipStreamEndCapture(stream, &g));
            hipGraphExec_t ex = nullptr;
            const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            AW_HIP(e);
            graph[d] = ex;
        }
    }

    void init() {
        const hipStream_t s = stream;
        for (Run& r : runs) {
            const aw_particles p = run_particles(r, r.cur);
            std::vector<double> wts(r.species.size());
            load_species(*r.ctx, r.species.data(), (int)r.species.size(), r.rp,
                         r.modes.empty() ? nullptr : r.modes.data(), (int)r.modes.size(), p,
                         wts.data(), s);
        }
        sort_runs(true);
        nsorts = run_sorts = 0;
        for (Run& r : runs) {
            deposit_rho_fixed(*r.ctx, run_particles(r, r.cur), r.rp.qm, AW_DEPOSIT_BINNED,
                              tile_off.get() + r.obase, s, acc.get() + r.cbase);
        }
        convert_and_solve();
        for (Run& r : runs) {
            const int64_t nc = ncell(*r.ctx);
            push_es(*r.ctx, run_particles(r, r.cur), E.get() + 2 * r.cbase,
                    E.get() + 2 * r.cbase + nc, r.rp, r.has_b0, true, s);
        }
        fields_ready = true;
        initialized = true;
        AW_HIP(hipStreamSynchronize(s));
        if (use_graph) capture_graphs();
    }

    void step(int64_t nstep) {
        if (!initialized) fail(AW_ERR_STATE, "aw_ensemble_step before aw_ensemble_init",
                               __FILE__, __LINE__);
        const hipStream_t s = stream;
        poll_strays();
        if (cfg.sort_every > 0 && steps_since_sort >= cfg.sort_every) sort_runs(true);
        else if (sort_requested) sort_runs(false);
        const bool dump = cfg_dump(nstep);
        if (fields_ready) {
            // the init solve already holds E^0: push only (first step)
            push_deposit_multi(active.get(), wg_run.get(), total_wg, lds, dump, uz, s);
            if (dump) {
                ens_diag_kernel<<<(int)runs.size(), kThreads, 0, s>>>(
                    diag_runs.get(), ke_part.get(), E.get(), rho.get(), diag_stage.get());
                AW_LAUNCH_CHECK();
            }
            fields_ready = false;
        } else if (use_graph) {
            AW_HIP(hipGraphLaunch(graph[dump ? 1 : 0], s));
            ++graph_launches;
        } else {
            enqueue_step(dump);
        }
        if (!stray_pending) {  // this step's stray counts (reset by the next conversion)
            AW_HIP(hipMemcpyAsync(stray_host, strays.get(),
                                  runs.size() * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s));
            AW_HIP(hipEventRecord(stray_ev, s));
            stray_pending = true;
            stray_age = 0;
        }
        if (dump) record_diag(nstep);
        ++steps_since_sort;
    }

    int64_t dump_every = 0;
    bool cfg_dump(int64_t nstep) const {
        return dump_every > 0 && (nstep % dump_every) == 0;
    }

    void record_diag(int64_t nstep) {
        if ((int)diag_pending.size() >= diag_cap) flush_diag();
        AW_HIP(hipMemcpyAsync(diag_ring.get() + diag_pending.size() * 4 * runs.size(),
                              diag_stage.get(), 4 * runs.size() * sizeof(double),
                              hipMemcpyDeviceToDevice, stream));
        diag_pending.push_back(nstep);
    }

    void flush_diag() {
        if (diag_pending.empty()) return;
        const size_t R = runs.size();
        std::vector<double> h(diag_pending.size() * 4 * R);
        AW_HIP(hipMemcpyAsync(h.data(), diag_ring.get(), h.size() * sizeof(double),
                              hipMemcpyDeviceToHost, stream));
        AW_HIP(hipStreamSynchronize(stream));
        for (size_t k = 0; k < diag_pending.size(); ++k)
            for (size_t i = 0; i < R; ++i) {
                Run& r = runs[i];
                const double* v = h.data() + (k * R + i) * 4;
                aw_diag_sample d{};
                d.step = diag_pending[k];
                d.time = (double)d.step * r.rp.dt;
                const double dxdy = r.g.dx * r.g.dy;
                d.ke = 0.5 * v[0];
                d.ee = 0.5 * r.rp.eps0 * v[1] * dxdy;
                d.max_e = std::sqrt(v[2]);
                d.charge = v[3] * dxdy;
                d.total = d.ke + d.ee;
                r.history.push_back(d);
            }
        diag_pending.clear();
    }

    Run& by_user(int i) {
        for (Run& r : runs)
            if (r.user == i) return r;
        fail(AW_ERR_INVALID, "run index out of range", __FILE__, __LINE__);
    }
};

extern "C" {

aw_status aw_ensemble_create(const aw_run_desc* descs, int32_t nruns, const aw_sim_config* cfg,
                             aw_ensemble** out) {
    return guard("aw_ensemble_create", [&] {
        AW_REQUIRE(descs && nruns > 0 && out, "descs / nruns / out");
        *out = nullptr;
        auto e = std::make_unique<aw_ensemble>();
        if (cfg) e->cfg = *cfg;
        if (e->cfg.field_model == AW_MODEL_DARWIN) {
            e->dar = darwin_batch_create(descs, nruns, &e->cfg);
            *out = e.release();
            return;
        }
        AW_REQUIRE(e->cfg.field_model == AW_MODEL_ES,
                   "field_model must be AW_MODEL_ES or AW_MODEL_DARWIN");
        // order runs by grid size so same-size runs form contiguous FFT batches
        std::vector<int> order(nruns);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            const aw_grid &ga = descs[a].grid, &gb = descs[b].grid;
            return std::tie(ga.nx, ga.ny, ga.dx, ga.dy) < std::tie(gb.nx, gb.ny, gb.dx, gb.dy);
        });
        e->eps0 = descs[0].rp.eps0;
        e->dump_every = descs[0].rp.dump_every;
        int scale = 60;
        for (int k : order) {
            const aw_run_desc& d = descs[k];
            AW_REQUIRE(d.species && d.nspecies > 0, "every run needs at least one species");
            AW_REQUIRE(d.n_modes >= 0 && d.n_modes <= 32, "n_modes must be in [0, 32]");
            AW_REQUIRE(d.rp.eps0 == e->eps0, "every run of an ensemble shares eps0");
            AW_REQUIRE(d.rp.dump_every == e->dump_every,
                       "every run of an ensemble shares rp.dump_every");
            aw_ensemble::Run r;
            r.user = k;
            r.g = d.grid;
            r.rp = d.rp;
            r.species.assign(d.species, d.species + d.nspecies);
            r.modes.assign(d.mode_amp, d.mode_amp + d.n_modes);
            const int64_t nc = (int64_t)d.grid.nx * d.grid.ny;
            double w0 = -1.0;
            for (const aw_species& sp : r.species) {
                AW_REQUIRE(sp.ppc > 0, "species ppc must be positive");
                r.n += (int64_t)sp.ppc * nc;
                const double w = sp.density * d.grid.dx * d.grid.dy / sp.ppc;
                AW_REQUIRE(w0 < 0 || w == w0, "ensemble runs carry one weight per run");
                w0 = w;
            }
            r.w_uniform = w0;
            r.has_b0 = d.has_b0 != 0;
            int tile = e->cfg.tile_x;
            if (tile <= 0) tile = r.n >= 64 * nc ? 8 : 16;
            r.bpt = 0;
            if (e->cfg.autotune && e->cfg.tile_x <= 0) {  // per run: its fastest tile and workgroups
                int tt = 0, tb = 0;
                autotune_run(d.grid, d.rp, d.species, d.nspecies, r.has_b0, tt, tb);
                if (tt > 0) {
                    tile = tt;
                    r.bpt = tb;
                }
            }
            r.ctx.reset(make_ctx(&d.grid, tile, tile));
            scale = std::min(scale, fixed_scale_for(r.species, r.rp, false));
            e->runs.push_back(std::move(r));
        }
        // offsets into the concatenated arrays
        int64_t pb = 0, cb = 0, ob = 0;
        int wg = 0;
        size_t lds = 0;
        for (auto& r : e->runs) {
            r.ctx->fix_log2 = scale;
            r.pbase = pb;
            r.cbase = cb;
            r.obase = ob;
            if (e->cfg.blocks_per_tile > 0) r.bpt = e->cfg.blocks_per_tile;
            else if (r.bpt <= 0) r.bpt = tile_bpt(*r.ctx, r.n);
            r.wg_begin = wg;
            r.nwg = r.ctx->ntiles * r.bpt;
            pb += r.n;
            cb += ncell(*r.ctx);
            ob += r.ctx->ntiles + 1;
            wg += r.nwg;
            lds = std::max(lds, push_lds_bytes(*r.ctx));
            e->uz = e->uz || uz_evolves(r.rp, r.has_b0);
        }
        AW_REQUIRE(pb < (int64_t)INT32_MAX, "ensemble particle count must fit int32");
        e->fix_log2 = scale;
        e->ntot = pb;
        e->nctot = cb;
        e->total_wg = wg;
        e->lds = lds;
        // FFT batches per grid size
        for (size_t k = 0; k < e->runs.size(); ++k) {
            const auto& r = e->runs[k];
            if (e->groups.empty() || e->runs[e->groups.back().first].g.nx != r.g.nx ||
                e->runs[e->groups.back().first].g.ny != r.g.ny ||
                e->runs[e->groups.back().first].g.dx != r.g.dx ||
                e->runs[e->groups.back().first].g.dy != r.g.dy)
                e->groups.push_back({(int)k, 0, r.cbase});
            ++e->groups.back().count;
        }
        for (const auto& gr : e->groups) enable_batch_fft(*e->runs[gr.first].ctx, gr.count);
        AW_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
        const size_t n = (size_t)std::max<int64_t>(1, pb);
        for (auto& buf : e->P)
            for (auto& a : buf) a.resize(n);
        e->tile_off.resize((size_t)ob);
        int64_t nmax = 1;
        for (auto& r : e->runs) nmax = std::max(nmax, r.n);
        e->perm.resize((size_t)nmax);
        e->acc.resize((size_t)cb);
        AW_HIP(hipMemset(e->acc.get(), 0, (size_t)cb * sizeof(unsigned long long)));
        e->strays.resize(e->runs.size());
        AW_HIP(hipMemset(e->strays.get(), 0, e->runs.size() * sizeof(unsigned long long)));
        e->rho.resize((size_t)cb);
        e->rin.resize((size_t)cb);
        {
            // the baseline a single aw_sim of the run uses (same scale, same species)
            std::vector<long long> h((size_t)cb);
            for (auto& r : e->runs) {
                double dens = 0.0;
                for (const auto& sp : r.species) dens += sp.density;
                r.ctx->rho_base = rho_baseline(r.rp.qm, dens, scale);
                std::fill(h.begin() + r.cbase, h.begin() + r.cbase + ncell(*r.ctx),
                          r.ctx->rho_base);
            }
            e->rbase.resize((size_t)cb);
            AW_HIP(hipMemcpy(e->rbase.get(), h.data(), h.size() * sizeof(long long),
                             hipMemcpyHostToDevice));
        }
        e->E.resize((size_t)2 * cb);
        e->ke_part.resize((size_t)std::max(1, wg));
        {
            std::vector<int32_t> h((size_t)std::max(1, wg), 0);
            for (size_t k = 0; k < e->runs.size(); ++k)
                std::fill(h.begin() + e->runs[k].wg_begin,
                          h.begin() + e->runs[k].wg_begin + e->runs[k].nwg, (int32_t)k);
            e->wg_run.resize(h.size());
            AW_HIP(hipMemcpy(e->wg_run.get(), h.data(), h.size() * sizeof(int32_t),
                             hipMemcpyHostToDevice));
        }
        const size_t R = e->runs.size();
        e->diag_stage.resize(4 * R);
        e->diag_ring.resize(4 * R * (size_t)e->diag_cap);
        std::vector<DiagRun> dr(R);
        for (size_t k = 0; k < R; ++k) {
            const auto& r = e->runs[k];
            dr[k] = DiagRun{r.wg_begin, r.nwg, (long long)r.cbase, (int)ncell(*r.ctx),
                            r.w_uniform};
        }
        e->diag_runs.resize(R);
        AW_HIP(hipMemcpy(e->diag_runs.get(), dr.data(), R * sizeof(DiagRun),
                         hipMemcpyHostToDevice));
        e->build_slots();
        AW_HIP(hipHostMalloc((void**)&e->stray_host,
                             e->runs.size() * sizeof(unsigned long long), hipHostMallocDefault));
        const size_t nside = std::min<size_t>(kSideStreams, e->groups.size() - 1);
        for (size_t i = 0; i < nside; ++i) {
            hipStream_t ss = nullptr;
            AW_HIP(hipStreamCreateWithFlags(&ss, hipStreamNonBlocking));
            e->side.push_back(ss);
        }
        for (size_t i = 0; i < (nside ? nside + 1 : 0); ++i) {
            hipEvent_t ev = nullptr;
            AW_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            e->fork_join.push_back(ev);
        }
        AW_HIP(hipEventCreateWithFlags(&e->stray_ev, hipEventDisableTiming));
        *out = e.release();
    });
}

aw_status aw_ensemble_destroy(aw_ensemble* e) {
    return guard("aw_ensemble_destroy", [&] { delete e; });
}

aw_status aw_ensemble_set_graph(aw_ensemble* e, int32_t on) {
    return guard("aw_ensemble_set_graph", [&] {
        AW_REQUIRE(e, "ensemble is NULL");
        if (e->dar) {
            AW_REQUIRE(!on, "Darwin ensembles step without HIP graphs");
            return;
        }
        AW_REQUIRE(!e->initialized, "set the graph mode before aw_ensemble_init");
        e->use_graph = on != 0;
    });
}

aw_status aw_ensemble_init(aw_ensemble* e) {
    return guard("aw_ensemble_init", [&] {
        AW_REQUIRE(e, "ensemble is NULL");
        AW_REQUIRE(!e->initialized && !e->dar_initialized,
                   "aw_ensemble_init called twice (create a new ensemble to start over)");
        if (e->dar) {
            darwin_batch_init(e->dar);
            e->dar_initialized = true;
            return;
        }
        e->init();
    });
}

aw_status aw_ensemble_step(aw_ensemble* e, int64_t n) {
    return guard("aw_ensemble_step", [&] {
        AW_REQUIRE(e, "ensemble is NULL");
        if (e->dar) darwin_batch_step(e->dar, n);
        else e->step(n);
    });
}

aw_status aw_ensemble_run(aw_ensemble* e, int64_t first, int64_t nsteps) {
    return guard("aw_ensemble_run", [&] {
        AW_REQUIRE(e, "ensemble is NULL");
        if (e->dar) {
            for (int64_t k = 0; k < nsteps; ++k) darwin_batch_step(e->dar, first + k);
            darwin_batch_sync(e->dar);
            return;
        }
        for (int64_t k = 0; k < nsteps; ++k) e->step(first + k);
        AW_HIP(hipStreamSynchronize(e->stream));
    });
}

aw_status aw_ensemble_synchronize(aw_ensemble* e) {
    return guard("aw_ensemble_synchronize", [&] {
        AW_REQUIRE(e, "ensemble is NULL");
        if (e->dar) darwin_batch_sync(e->dar);
        else AW_HIP(hipStreamSynchronize(e->stream));
    });
}

aw_status aw_ensemble_info(aw_ensemble* e, aw_ensemble_stats* out) {
    return guard("aw_ensemble_info", [&] {
        AW_REQUIRE(e && out, "NULL argument");
        if (e->dar) {
            darwin_batch_info(e->dar, out);
            return;
        }
        out->nruns = (int32_t)e->runs.size();
        out->ngroups = (int32_t)e->groups.size();
        out->particles = e->ntot;
        out->cells = e->nctot;
        out->workgroups = e->total_wg;
        out->sorts = e->nsorts;
        out->run_sorts = e->run_sorts;
        out->graph_launches = e->graph_launches;
        out->graph = e->use_graph ? 1 : 0;
    });
}

aw_status aw_ensemble_particles(aw_ensemble* e, int32_t run, aw_particles* out) {
    return guard("aw_ensemble_particles", [&] {
        AW_REQUIRE(e && out, "NULL argument");
        if (e->dar) {
            *out = darwin_batch_particles(e->dar, run);
            return;
        }
        const auto& r = e->by_user(run);
        *out = e->run_particles(r, r.cur);
    });
}

aw_status aw_ensemble_fields(aw_ensemble* e, int32_t run, aw_fields* out) {
    return guard("aw_ensemble_fields", [&] {
        AW_REQUIRE(e && out, "NULL argument");
        if (e->dar) {
            *out = darwin_batch_fields(e->dar, run);
            return;
        }
        const auto& r = e->by_user(run);
        std::memset(out, 0, sizeof(*out));
        out->Ex = e->E.get() + 2 * r.cbase;
        out->Ey = out->Ex + ncell(*r.ctx);
    });
}

aw_status aw_ensemble_rho(aw_ensemble* e, int32_t run, float** out) {
    return guard("aw_ensemble_rho", [&] {
        AW_REQUIRE(e && out, "NULL argument");
        if (e->dar) {
            *out = darwin_batch_rho(e->dar, run);
            return;
        }
        *out = e->rho.get() + e->by_user(run).cbase;
    });
}

aw_status aw_ensemble_history(aw_ensemble* e, int32_t run, aw_diag_sample* out, int64_t cap,
                              int64_t* count) {
    return guard("aw_ensemble_history", [&] {
        AW_REQUIRE(e && count, "NULL argument");
        if (!e->dar) e->flush_diag();
        const auto& h = e->dar ? darwin_batch_history(e->dar, run) : e->by_user(run).history;
        const int64_t nh = (int64_t)h.size();
        if (out)
            for (int64_t k = 0; k < std::min(cap, nh); ++k) out[k] = h[k];
        *count = nh;
    });
}

}  // extern "C"
