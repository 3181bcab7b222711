/*
 * aw.h -- C ABI of the MI355X-native ArcWarden 2D3V PIC hot path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns an aw_status (0 = success).  On failure aw_last_error() returns a
 * message naming the call, the file:line and the HIP / rocFFT error, the
 * same information the reference's CudaError carries
 * (include/pic/cuda_utils.hpp:29-77).
 *
 * Device memory passed in aw_particles / aw_fields / float* arguments is
 * CALLER-owned (hipMalloc'd or torch-allocated), as in the reference, whose
 * kernels receive non-owning DeviceViews (include/pic/device_array.hpp:25-40).
 * Every launch is stream-ordered on the aw_stream argument (a hipStream_t;
 * NULL = the legacy default stream).  An aw_ctx owns per-grid resources
 * (rocFFT plans, k-space workspace, sort scratch) -- the reference's
 * SpectralEngine + Particles scratch (include/pic/spectral.hpp:217-260,
 * include/pic/particles.hpp:460-469).  Contexts hold no process-wide state,
 * so several grids (ensembles) can live side by side.
 *
 * Conventions (reference include/pic/grid.hpp:14-29, particles.hpp:12-16):
 *   positions x,y in cell units, x in [0,nx), y in [0,ny);
 *   fields row-major [ny][nx], x contiguous;
 *   k-space: interleaved complex64 [ny][nx/2+1];
 *   u = gamma*v with gamma == 1.
 *
 * Charge / current accumulation is deterministic: every per-particle
 * contribution is quantised to a 2^-S fixed-point integer (S chosen per
 * context, aw_ctx_set_fixed_scale) and summed with integer atomics, so the
 * result does not depend on thread order and all deposit policies return
 * identical bits.
 */
#ifndef AW_H
#define AW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t aw_status;
#define AW_OK 0
#define AW_ERR_INVALID 1     /* bad argument / shape */
#define AW_ERR_HIP 2         /* HIP runtime error */
#define AW_ERR_FFT 3         /* rocFFT error */
#define AW_ERR_NOMEM 4       /* device allocation failed */
#define AW_ERR_UNSUPPORTED 5 /* option not implemented for this path */
#define AW_ERR_STATE 6       /* call out of order (e.g. step before init) */

typedef void* aw_stream; /* hipStream_t */

/* ---- plain data (mirrors of the reference's POD types) ------------------ */

/* Grid -- include/pic/grid.hpp:44-57 (nx,ny,dx,dy,Lx,Ly). */
typedef struct aw_grid {
    int32_t nx, ny;
    double dx, dy, Lx, Ly;
} aw_grid;

/* RunParams -- include/pic/config.hpp:127-313, member for member (bool as
 * int32).  The Yee-branch members (jfilter, bnd_*, deltaf and df_*, rel,
 * cold_nc, b0_*, ant_*) are read by the aw_maxwell_* engine. */
typedef struct aw_run_params {
    double dt, n0, vth, vd, wpe, wce;
    float B0[3];
    int32_t ppc;
    int64_t nsteps, dump_every;
    int32_t norm;
    double eps0, qm, c;
    int32_t ndc;
    int32_t jfilter, tile_sort, bnd_x, bnd_nd;
    double bnd_numax;
    int32_t deltaf;
    double df_tpar, df_tperp;
    int32_t rel;
    double cold_nc;
    int32_t b0_prof;
    double b0_a, b0_xc, b0_lre, b0_sref, b0_c[6];
    int32_t df_dist;
    double df_rho, df_kappa, df_taud;
    int32_t df_wprec;
    double ant_amp, ant_x0, ant_sigma, ant_w0, ant_trmp, ant_toff;
    int32_t pump;
    double pump_ex, pump_ey, pump_ez, pump_k0, pump_w0, pump_trmp, pump_toff;
    double weight, ax, ay, k_filter;
    uint64_t rng_seed;
    double perturb_amp;
    int32_t perturb_kx;
    int32_t two_stream, bump_on_tail;
    double beam_frac, beam_vd, beam_vth;
    int32_t noisy_load;
} aw_run_params;

/* Species -- include/pic/species.hpp:20-45. */
typedef struct aw_species {
    char name[32];
    double density;
    int32_t ppc;
    double uth[3], ufl[3];
    int32_t deltaf, dist;
    double lc_rho, lc_kappa, taud, wdnoise;
} aw_species;

/* Particle SoA (include/pic/particles.hpp:47-57): caller-owned device
 * arrays.  w == NULL means every particle carries weight w_uniform. */
typedef struct aw_particles {
    float *x, *y, *ux, *uy, *uz;
    float* w;
    int64_t n;
    double w_uniform;
} aw_particles;

/* Real-space fields (include/pic/fields.hpp:197-200 + Darwin members). */
typedef struct aw_fields {
    float *Ex, *Ey, *Ez, *Bx, *By, *Bz;
    float *ELx, *ELy, *ETx, *ETy, *ETz;
} aw_fields;

/* Real-space sources (include/pic/sources.hpp:106-113). */
typedef struct aw_sources {
    float* rho;
    float *Jx, *Jy, *Jz;
    float *dcux, *dcuy, *dcuz;
    float *amu0, *amu1, *amu2, *amu3;
} aw_sources;

/* DiagSample -- include/pic/diagnostics.hpp:36-44. */
typedef struct aw_diag_sample {
    int64_t step;
    double time, ke, ee, total, charge, max_e;
} aw_diag_sample;

/* Deposit policies (include/pic/config.hpp:58-66). */
#define AW_DEPOSIT_GLOBAL 0    /* AtomicGlobalDeposit */
#define AW_DEPOSIT_LDS_GRID 1  /* SharedTileDeposit: whole grid in LDS */
#define AW_DEPOSIT_BINNED 2    /* TiledBinnedDeposit: tile + halo in LDS */
#define AW_DEPOSIT_AUTO 3      /* pick per grid (the reference's charge()) */

/* Field models (include/pic/config.hpp:74). */
#define AW_MODEL_ES 0
#define AW_MODEL_DARWIN 1

/* Current deposit of the Darwin model (aw_sim_config.current_deposit). */
#define AW_CURRENT_CIC 0        /* charge_current: CIC J at x^{n+1} (the reference) */
#define AW_CURRENT_ESIRKEPOV 1  /* Esirkepov J of the move x^n -> x^{n+1} (staggered) */

/* aw_solve_darwin_el_b_ex flags */
#define AW_J_STAGGERED 1        /* Jx at (i+1/2, j), Jy at (i, j+1/2) */

/* ---- input deck (Deck / load_deck, include/pic/deck.hpp:31-274) ------------- */

#define AW_DECK_MAX_SPECIES 16
#define AW_DECK_MAX_PROBES 64

/* Deck, member for member; std::string members as fixed NUL-terminated
 * buffers (diag_enable: the module keys joined by single spaces), the
 * species list and the probe list as counted arrays. */
typedef struct aw_deck {
    int32_t nx, ny;
    double Lx, Ly;
    aw_run_params rp;
    int32_t nspecies;
    aw_species species[AW_DECK_MAX_SPECIES];
    int64_t dump_every;
    char outdir[256];
    int32_t darwin, yee;
    double dx_wpe_c, c_direct, wce, theta_deg;
    int32_t b0_direct, b0_xc_set;
    double bnoise;
    int32_t pump_enable, pump_M;
    double pump_w0, pump_ex0, pump_ey0, pump_ez0, pump_amp, pump_trmp, pump_toff;
    char diag_enable[256];
    int32_t nprobes;
    double probes[AW_DECK_MAX_PROBES];
    double tsnap;
    int32_t band_lo, band_hi, kt_stride, n_frames;
    char prefix[128];
} aw_deck;

/* load_deck (include/pic/deck.hpp:99-274): parses the INI deck at path and
 * applies the finalize step (c from dx_wpe_c, B0 from wce/theta_deg, pump
 * amplitudes and k0, B0-profile centre).  Host only (no GPU work).  Errors
 * carry the reference's messages ("deck: cannot open <path>", "deck: bad
 * section at line N", ...).  profile = dipole also fits rp.b0_c / b0_sref
 * (bg::fit_dipole, include/pic/background_b0.hpp:98-146): the even degree-5
 * polynomial in (s/s_ref)^2 of the dipole B/B_eq, checked to 2e-4. */
aw_status aw_deck_load(const char* path, aw_deck* out);

/* ---- errors -------------------------------------------------------------- */

/* Copies the calling thread's last error message (NUL-terminated). */
void aw_last_error(char* buf, size_t len);
/* Library version string, e.g. "arcwarden-amd 0.1 gfx950". */
const char* aw_version(void);

/* Device memory for C / C++ callers without the HIP headers: DeviceArray's
 * allocate / release / copies (include/pic/device_array.hpp:45-107).
 * aw_copy is a blocking copy between any two host or device pointers. */
aw_status aw_device_alloc(size_t bytes, void** out);
aw_status aw_device_free(void* p);
aw_status aw_copy(void* dst, const void* src, size_t bytes);
aw_status aw_memset_zero(void* p, size_t bytes, aw_stream s);
aw_status aw_stream_synchronize(aw_stream s);
/* cudaDeviceSynchronize for callers without the HIP headers: waits for all
 * work on the current device (the diagnostics modules read state this way). */
aw_status aw_device_synchronize(void);

/* Default-constructed RunParams / Species: the member initialisers of
 * include/pic/config.hpp:127-313 and include/pic/species.hpp:20-45. */
void aw_run_params_defaults(aw_run_params* rp);
void aw_species_defaults(aw_species* sp);

/* ---- context ------------------------------------------------------------- */

typedef struct aw_ctx aw_ctx;

/* Creates a context for one grid: rocFFT R2C/C2R plans of shape (ny,nx)
 * (include/pic/fft.hpp:26-35), k-space workspace, tile geometry tx*ty
 * (0 -> 16x16, clamped to the grid).  Replaces SpectralEngine(g)
 * (include/pic/spectral.hpp:217-227). */
aw_status aw_ctx_create(const aw_grid* g, int32_t tile_x, int32_t tile_y, aw_ctx** out);
aw_status aw_ctx_destroy(aw_ctx* ctx);
/* Fixed-point exponent S for rho (contributions quantised to 2^-S);
 * J / amu / dcu use S minus a per-quantity headroom.  Default S = 40. */
aw_status aw_ctx_set_fixed_scale(aw_ctx* ctx, int32_t scale_log2);
aw_status aw_ctx_tiles(const aw_ctx* ctx, int32_t* ntiles, int32_t* ntx, int32_t* tx,
                       int32_t* ty);

/* ---- integer layer: cells, tile bins, sort -------------------------------- */

/* cell[i] = floor(y)*nx + floor(x), clamped (include/pic/particles.hpp:310-314). */
aw_status aw_cell_index(aw_ctx* ctx, const aw_particles* p, int32_t* cell, aw_stream s);
/* Tile binning without moving particles -- Particles::build_tile_bins
 * (include/pic/particles.hpp:632-671): off[ntiles+1], idx[n].  The order
 * inside a tile is STABLE (ascending particle index), unlike the
 * reference's atomic-order slots (:149-152). */
aw_status aw_bin_tiles(aw_ctx* ctx, const aw_particles* p, int32_t* off, int32_t* idx,
                       aw_stream s);
/* Stable sort of particles by tile-major cell key (tile*tx*ty + in-tile
 * cell): perm[n] (perm[k] = index of k-th particle), off[ncellkeys+1]. */
aw_status aw_sort_cells(aw_ctx* ctx, const aw_particles* p, int32_t* perm, int32_t* off,
                        aw_stream s);
/* Physical tile sort -- Particles::sort_by_tile (include/pic/particles.hpp:
 * 678-726): stable by tile, SoA moved from src into dst (distinct buffers,
 * same n), tile offsets into off[ntiles+1]. */
aw_status aw_sort_by_tile(aw_ctx* ctx, const aw_particles* src, aw_particles* dst,
                          int32_t* off, aw_stream s);
/* particle_migrate_kernel (include/pic/particles.hpp:293-315): periodic wrap;
 * cell may be NULL. */
aw_status aw_migrate(aw_ctx* ctx, aw_particles* p, int32_t* cell, aw_stream s);

/* ---- loaders (setup) ------------------------------------------------------ */

/* Particles::initialize(species...) (include/pic/particles.hpp:533-559):
 * species k fills [base_k, base_k + ppc_k*nx*ny) of p (p->n must equal the
 * total).  Particle (cell c, slot s) of species k is drawn exactly as the
 * reference draws index base+c*ppc+s and stored at base + s*nx*ny + c.
 * mode_amp[m-1] (physical length) adds x += sum_m A_m sin(2 pi m x/Lx)
 * (config extension; n_modes may be 0). weights_out (nspecies doubles)
 * receives density*dx*dy/ppc per species. */
aw_status aw_load_species(aw_ctx* ctx, const aw_species* sp, int32_t nspecies,
                          const aw_run_params* rp, const double* mode_amp,
                          int32_t n_modes, aw_particles* p, double* weights_out,
                          aw_stream s);
/* Particles::initialize(rp) single-species legacy load with perturb_amp,
 * two_stream and bump_on_tail (include/pic/particles.hpp:213-290,521-527). */
aw_status aw_load_legacy(aw_ctx* ctx, const aw_run_params* rp, aw_particles* p,
                         aw_stream s);

/* ---- deposits ---------------------------------------------------------------- */

/* Depositor<Cfg>::charge (include/pic/depositor.hpp:34-153,440-496):
 * rho[c] += sum_p qm*w_p/(dx*dy) * W_cic.  Accumulates into rho (caller
 * zeroes it, as in the reference).  policy = AW_DEPOSIT_*. */
aw_status aw_deposit_rho(aw_ctx* ctx, const aw_particles* p, double qm, int32_t policy,
                         float* rho, aw_stream s);

/* No counterpart in the reference (its deposits add FP32 with atomics).  The
 * int64 fixed-point scale 2^S the engine uses for rho: S puts |qm| density at
 * 2^40 (clamped to [8, 60]), then is lowered until n particles of weight
 * wmax summed into one cell stay below 2^62, so no cell sum can wrap; a
 * configuration that would need S < 8 is AW_ERR_INVALID.  Host only.  For
 * callers that choose the scale of aw_deposit_rho_noalloc or
 * aw_ctx_set_fixed_scale. */
aw_status aw_rho_fixed_scale(const aw_grid* g, double qm, double density, double wmax,
                             int64_t n, int32_t* scale_log2);

/* No counterpart in the reference.  The charge deposit of aw_deposit_rho for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): it needs no context
 * and allocates nothing, accumulating into a static per-GPU device array of
 * at most 2^20 cells (8 MiB; larger grids are an AW_ERR_INVALID), then adding
 * the FP32 result into rho like aw_deposit_rho.  The same integer sums as
 * aw_deposit_rho at fixed-point scale 2^scale_log2 give the same bits.
 * policy: AW_DEPOSIT_GLOBAL, AW_DEPOSIT_LDS_GRID or AW_DEPOSIT_AUTO.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (several streams or host threads) race; calls on one stream are ordered. */
aw_status aw_deposit_rho_noalloc(const aw_grid* g, const aw_particles* p, double qm,
                                 int32_t policy, int32_t scale_log2, float* rho, aw_stream s);

/* Esirkepov charge-conserving current, the deposit half of k_push_esirkepov /
 * k_push_esirkepov_tiled with esirkepov_scatter + GlobalJSink
 * (include/pic/yee2d.hpp:124-132,310-389,401-469): for the moves
 * (x0[i], y0[i]) -> (p->x[i], p->y[i]) of less than one cell (cell units; p's
 * positions may be wrapped or unwrapped, the minimal image is used),
 * jx += staggered Jx at (i+1/2, j), jy += Jy at (i, j+1/2), jz += nodal
 * Jz = q w uz <S>_t/(dx dy), q = qm.  Accumulates (caller zeroes).  Charge
 * conservation holds against aw_deposit_rho of the two positions.  policy =
 * AW_DEPOSIT_* (BINNED bins by the old position).  Deterministic: int64
 * fixed point, every policy returns the same bits. */
aw_status aw_deposit_esirkepov(aw_ctx* ctx, const float* x0, const float* y0,
                               const aw_particles* p, double qm, double dt, int32_t policy,
                               float* jx, float* jy, float* jz, aw_stream s);

/* ---- push ---------------------------------------------------------------------- */

/* Pusher<Cfg>::boris (include/pic/pusher.hpp:112-129,265-275): CIC gather of
 * (Ex,Ey), Boris with uniform rp.B0 when has_b0, x += u*dt/dx.  No wrap
 * (call aw_migrate), as in the reference. */
aw_status aw_push_es(aw_ctx* ctx, aw_particles* p, const float* Ex, const float* Ey,
                     const aw_run_params* rp, int32_t has_b0, aw_stream s);
/* Pusher<Cfg>::half_step_back (include/pic/pusher.hpp:133-145,321-330). */
aw_status aw_half_step_back(aw_ctx* ctx, aw_particles* p, const float* Ex,
                            const float* Ey, const aw_run_params* rp, int32_t has_b0,
                            aw_stream s);
/* The MI355X hot path: one pass per particle that gathers E^n from an LDS
 * field tile, pushes (Boris, uniform B0), moves, wraps, and deposits
 * rho(x^{n+1}) into an LDS tile flushed once per workgroup.  Particles must
 * be tile-sorted (off from aw_sort_by_tile); particles that left their
 * tile's apron since the sort fall back to global gathers/atomics.
 * rho_next += rho(x^{n+1}).  Replaces the reference's
 * boris_push_tiled_kernel + particle_migrate_kernel + next step's
 * deposit_charge_tiled_kernel (include/pic/pusher.hpp:154-202;
 * particles.hpp:293-315; depositor.hpp:112-153).  strays (may be NULL)
 * receives the number of fallback particles. */
aw_status aw_push_deposit_es(aw_ctx* ctx, aw_particles* p, const int32_t* off,
                             const float* Ex, const float* Ey, const aw_run_params* rp,
                             int32_t has_b0, float* rho_next, int64_t* strays,
                             aw_stream s);

/* ---- spectral -------------------------------------------------------------------- */

/* SpectralEngine::r2c / c2r (include/pic/spectral.hpp:228-243): unnormalised
 * R2C; C2R applies 1/(nx*ny) and may overwrite its input. */
aw_status aw_fft_r2c(aw_ctx* ctx, const float* in, void* out_k, aw_stream s);
aw_status aw_fft_c2r(aw_ctx* ctx, void* in_k, float* out, aw_stream s);
/* ElectrostaticSpectralSolver::solve (include/pic/solver_es.hpp:85-106):
 * rho -> E = -grad phi, phi_k = rho_k/(eps0 k^2), DC and Nyquist zeroed. */
aw_status aw_solve_es(aw_ctx* ctx, const float* rho, float* Ex, float* Ey, double eps0,
                      aw_stream s);

/* DarwinSpectralSolver::solve_el_b (include/pic/solver_darwin.hpp:188-216) +
 * add_background_b0 (include/pic/pump.hpp:59-67): src4 = [rho|Jx|Jy|Jz]
 * (4*nx*ny floats, contiguous) -> out5 = [ELx|ELy|Bx|By|Bz], B = B_self + B0. */
aw_status aw_solve_darwin_el_b(aw_ctx* ctx, const float* src4, float* out5, double eps0,
                               double c, const float B0[3], aw_stream s);
/* aw_solve_darwin_el_b for a staggered current (flags = AW_J_STAGGERED): Jx
 * sampled at (i+1/2, j) and Jy at (i, j+1/2), as the Esirkepov deposit writes
 * them, are moved onto the nodes in k-space (J_k *= exp(-i k dx/2)) before B
 * is formed.  No reference counterpart: the reference's Darwin solver takes
 * the CIC (nodal) current only.  flags = 0 is aw_solve_darwin_el_b. */
aw_status aw_solve_darwin_el_b_ex(aw_ctx* ctx, const float* src4, float* out5, double eps0,
                                  double c, const float B0[3], int32_t flags, aw_stream s);
/* DarwinSpectralSolver::solve_et (include/pic/solver_darwin.hpp:235-263):
 * src7 = [dcux|dcuy|dcuz|amu0|amu1|amu2|amu3] -> out3 = [ETx|ETy|ETz]. */
aw_status aw_solve_darwin_et(aw_ctx* ctx, const float* src7, float* out3, double eps0,
                             double c, double n0, aw_stream s);

/* form_el / form_etotal + add_pump_field (include/pic/solver_darwin.hpp:134-144,
 * 220-231; include/pic/pump.hpp:27-79): f->Ex, Ey, Ez = E_L + E_T + pump(t)
 * from f->ELx, ELy and f->ETx, ETy, ETz (f->ETx NULL: E_L + pump only, the
 * field the dcu deposit gathers).  The pump is rp->pump_* (off when rp->pump
 * is 0). */
aw_status aw_form_e(aw_ctx* ctx, aw_fields* f, const aw_run_params* rp, double t,
                    aw_stream s);

/* ---- Darwin particle operators (any particle order) ----------------------------- */

/* Depositor<Cfg>::charge_current + deposit_amu (include/pic/depositor.hpp:160-270,
 * 369-422): with W the CIC weights and c_p = qm w_p/(dx dy),
 *   rho += c_p W, J += c_p (ux, uy, uz) W,
 *   amu0 += c_p (ux^2 - uy^2)/2 W, amu1 += c_p ux uy W, amu2 += c_p uz ux W,
 *   amu3 += c_p uz uy W
 * into out->rho, Jx, Jy, Jz, amu0..amu3; a NULL member is not deposited
 * (charge_current passes no amu, deposit_amu no rho / J).  Accumulates;
 * stream-ordered, no allocation after the context's first call. */
aw_status aw_deposit_rho_j_amu(aw_ctx* ctx, const aw_particles* p, double qm, aw_sources* out,
                               aw_stream s);
/* Depositor<Cfg>::deposit_dcu (include/pic/depositor.hpp:275-324,426-437):
 * dcu += qm^2 w/(dx dy) (E + u x B) W with E = f->Ex, Ey, Ez (E_L + pump) and
 * B = f->Bx, By, Bz gathered at the particle, into out->dcux, dcuy, dcuz.
 * Accumulates. */
aw_status aw_deposit_dcu(aw_ctx* ctx, const aw_particles* p, double qm, const aw_fields* f,
                         aw_sources* out, aw_stream s);
/* Pusher<Cfg>::boris_em + Particles::migrate (include/pic/pusher.hpp:211-258,
 * 300-317; particles.hpp:293-315): E = f->Ex, Ey, Ez and B = f->Bx, By, Bz
 * gathered at the particle, full Boris rotation, move, periodic wrap. */
aw_status aw_push_boris_em(aw_ctx* ctx, aw_particles* p, const aw_fields* f,
                           const aw_run_params* rp, aw_stream s);

/* ---- diagnostics ------------------------------------------------------------------- */

/* Diagnostics::compute (include/pic/diagnostics.hpp:58-108) as an on-device
 * FP64 reduction: KE = 1/2 sum w u^2, EE = 1/2 eps0 sum |E|^2 dx dy,
 * charge = sum rho dx dy, max|E|.  rho may be NULL (charge = 0).  Blocks
 * until the result is on the host. */
aw_status aw_reduce_energies(aw_ctx* ctx, const aw_particles* p, const float* Ex,
                             const float* Ey, const float* rho, double eps0,
                             aw_diag_sample* out, aw_stream s);

/* Module diagnostics (include/pic/diag) reduce on the device: */

/* f(x, v_par/v_th) counts over all particles, v_par = ux*bx + uz*bz (b the
 * unit B0 direction in the x-z plane): counts[iv*nxb + ix] += 1 (uint32,
 * the caller zeroes).  fp64 = 0: the binning of phase_hist_kernel
 * (include/pic/diag/module.hpp:58-70; FP32, x wrapped in cells,
 * ix = (int)(x/nx*nxb)); fp64 = 1: the binning of Snapshot's host loop
 * (include/pic/diag/snapshot.hpp:55-62; FP64, x*dx wrapped by Lx).  Both
 * compute iv = (int)((v_par - vlo)/(vhi - vlo)*nvb) and skip bins outside
 * [0, nxb) x [0, nvb) after the truncating cast, as the reference does. */
aw_status aw_phase_hist(const aw_particles* p, const aw_grid* g, int32_t nxb, int32_t nvb,
                        double bx, double bz, double vth, double vlo, double vhi, int32_t fp64,
                        uint32_t* counts, aw_stream s);
/* diag::sum_sq (include/pic/diag/module.hpp:41-46) on the device: FP64
 * sum of (a[i] - offset)^2 over n floats, in a fixed order
 * (deterministic).  Blocks until *out (host) holds the result. */
aw_status aw_sum_sq(const float* a, int64_t n, double offset, double* out, aw_stream s);
/* dst[k] = src[idx[k]] for k < m: PhaseFrames' fixed particle sample
 * (include/pic/diag/phase_frames.hpp:36-50) gathered on the device. */
aw_status aw_gather_f32(const float* src, const int32_t* idx, int64_t m, float* dst,
                        aw_stream s);

/* ---- simulation (Simulation<Cfg>, include/pic/simulation.hpp:38-140) ------------- */

typedef struct aw_sim aw_sim;

/* SimConfig (include/pic/config.hpp:81-98) as run-time values plus the
 * MI355X execution knobs.  Zero-initialise and set what you need. */
typedef struct aw_sim_config {
    int32_t has_b0;         /* SimConfig::has_b0 */
    int32_t deposit_policy; /* AW_DEPOSIT_* used for init deposits */
    int32_t field_model;    /* AW_MODEL_ES | AW_MODEL_DARWIN */
    int32_t tile_x, tile_y; /* 0 -> 8 (ES at >= 64 particles per cell), else 16 */
    int32_t sort_every;     /* re-sort cadence in steps (0 -> adaptive only) */
    double stray_frac;      /* re-sort when strays exceed this fraction (0 -> 0.02) */
    int32_t n_modes;        /* x-displacement modes applied at load */
    double mode_amp[32];    /* physical amplitude of mode m = 1..n_modes */
    int32_t legacy_load;    /* 1: Particles::initialize(rp) path */
    int32_t fixed_scale;    /* 0 -> chosen from the species densities */
    int32_t blocks_per_tile; /* workgroups sharing one tile (0 -> heuristic) */
    int32_t autotune;       /* 1: at aw_sim_init, time one step's particle
                               kernels (ES: the fused push+deposit; Darwin: the
                               dcu deposit + EM push) for tile 8/16/32 x 1..16
                               workgroups per tile and keep the fastest (cached
                               per grid, particle count and kernel variant for
                               the process); results are unchanged */
    int32_t current_deposit; /* Darwin model: AW_CURRENT_CIC (the reference) or
                               AW_CURRENT_ESIRKEPOV -- the next state's J is the
                               charge-conserving current of the push's move
                               (staggered, shifted onto the nodes in k-space);
                               rho and amu stay CIC.  The t = 0 moments are CIC. */
} aw_sim_config;

aw_status aw_sim_create(const aw_grid* g, const aw_run_params* rp, const aw_species* sp,
                        int32_t nspecies, const aw_sim_config* cfg, aw_sim** out);
aw_status aw_sim_destroy(aw_sim* sim);
/* Simulation::init (:56-77): load, t=0 field, leapfrog half step back. */
aw_status aw_sim_init(aw_sim* sim);
/* Simulation::step (:80-93); records a DiagSample when rp.dump_every says so
 * (dump_every <= 0: every step), as Diagnostics::maybe_compute (:111-120). */
aw_status aw_sim_step(aw_sim* sim, int64_t n);
/* Simulation::run (:95-98): steps 0..nsteps-1 then synchronises. */
aw_status aw_sim_run(aw_sim* sim, int64_t nsteps);
aw_status aw_sim_synchronize(aw_sim* sim);
aw_status aw_sim_stream(aw_sim* sim, aw_stream* out);
aw_status aw_sim_grid(aw_sim* sim, aw_grid* out);
/* Current particle buffers (they swap on re-sort; query after each step). */
aw_status aw_sim_particles(aw_sim* sim, aw_particles* out);
aw_status aw_sim_fields(aw_sim* sim, aw_fields* out);
aw_status aw_sim_sources(aw_sim* sim, aw_sources* out);
/* Diagnostics::history (:128): copies up to cap samples, returns count. */
aw_status aw_sim_history(aw_sim* sim, aw_diag_sample* out, int64_t cap, int64_t* count);
/* Diagnostics::compute now (blocking). */
aw_status aw_sim_diag_now(aw_sim* sim, int64_t step, aw_diag_sample* out);
/* Restart files of a Simulation (no reference counterpart beyond the file
 * layout above, which the reference defines but fills with headers only):
 * save writes the particles, the fixed-point sources the next step reads
 * (ES: rho; Darwin: rho, J, amu) and the ES field, with the header's step /
 * time / rng_seed and the deck text.  restore loads such a file into a
 * created (not initialised) simulation of the same grid, species and model;
 * stepping then continues with the same rho, fields and particle multiset bit
 * for bit as the run that saved it (the particles are re-sorted).  Plain runs
 * only (no Darwin batches). */
aw_status aw_sim_checkpoint_save(aw_sim* sim, const char* path, int64_t step, double time,
                                 const char* deck, const char* git_hash);
aw_status aw_sim_checkpoint_restore(aw_sim* sim, const char* path);
/* The caller wrote new particle values into the buffers of aw_sim_particles
 * (positions in [0, nx) x [0, ny)).  The reference's step deposits its
 * sources from the particles as they are (include/pic/simulation.hpp:80-93),
 * so tools that edit sim.particles() between steps keep working: the next
 * step re-sorts and takes rho (and, Darwin, J and amu) from the edited state. */
aw_status aw_sim_particles_changed(aw_sim* sim);
/* Forces a re-sort before the next step (tests; normally cadence/adaptive). */
aw_status aw_sim_request_sort(aw_sim* sim);
/* Event timing of the dominant particle kernel, recorded on the sim stream
 * while enabled: total ms, launches, sorts performed. */
aw_status aw_sim_set_profiling(aw_sim* sim, int32_t on);
aw_status aw_sim_kernel_time(aw_sim* sim, double* total_ms, int64_t* launches,
                             int64_t* sorts);
/* No reference counterpart (the reference fixes 16x16 tiles and picks blocks
 * per tile from a cached occupancy, include/pic/depositor.hpp:344-364): the
 * tile size and workgroups per tile in use, after the autotuner when
 * cfg.autotune was set. */
aw_status aw_sim_exec_config(aw_sim* sim, int32_t* tile_x, int32_t* tile_y,
                             int32_t* blocks_per_tile);

/* ---- checkpoint / restart (include/pic/checkpoint.hpp:1-131) ----------------------
 * The ARCW v1 file: "ARCW", u32 version, u64 rng_seed, i64 step, f64 time,
 * f64 S_d, f64 eps_L; u64-length deck text; u64-length git hash; u32 array
 * count and per array u32-length name, u32 dtype, u64 count; then the
 * payloads in manifest order (a schema-only file stops before them). */

#define AW_CK_F32 0u
#define AW_CK_F64 1u
#define AW_CK_I32 2u
#define AW_CK_U64 3u

typedef struct aw_checkpoint_header { /* CheckpointHeader (:40-52) */
    uint32_t version;                 /* read back; written as 1 */
    uint64_t rng_seed;
    int64_t step;
    double time, sd, eps_l;
} aw_checkpoint_header;

typedef struct aw_checkpoint_array {  /* CheckpointArray (:34-38) + its payload */
    const char* name;
    uint32_t dtype;                   /* AW_CK_* */
    uint64_t count;
    void* data;                       /* NULL: manifest entry only */
    int32_t on_device;                /* 1: device memory (staged through pinned host) */
} aw_checkpoint_array;

typedef struct aw_checkpoint aw_checkpoint;

/* checkpoint_write_header (:65-90) plus the payloads of every array with
 * data (host or device).  deck / git_hash may be NULL (empty). */
aw_status aw_checkpoint_write(const char* path, const aw_checkpoint_header* h, const char* deck,
                              const char* git_hash, const aw_checkpoint_array* arrays,
                              int32_t narrays);
/* checkpoint_read_header (:92-127): "checkpoint: bad magic", "checkpoint:
 * version N newer than reader", "checkpoint: short read" as in the
 * reference; payloads are located, not read. */
aw_status aw_checkpoint_open(const char* path, aw_checkpoint** out);
aw_status aw_checkpoint_close(aw_checkpoint* ck);
aw_status aw_checkpoint_info(const aw_checkpoint* ck, aw_checkpoint_header* h, int32_t* narrays,
                             int64_t* deck_len, int64_t* git_len);
/* which = 0: the deck text, 1: the git hash (cap >= length + 1). */
aw_status aw_checkpoint_text(const aw_checkpoint* ck, int32_t which, char* buf, int64_t cap);
aw_status aw_checkpoint_array_info(const aw_checkpoint* ck, int32_t i, char* name, int64_t cap,
                                   uint32_t* dtype, uint64_t* count, int32_t* present);
aw_status aw_checkpoint_read_array(const aw_checkpoint* ck, int32_t i, void* dst,
                                   int32_t on_device);

/* ---- full-Maxwell (Yee) field model (MaxwellSimulation,
 *      include/pic/simulation_maxwell.hpp:24-227; include/pic/yee2d.hpp) ------------- */

typedef struct aw_maxwell aw_maxwell;

/* The mesh arrays (device, nx*ny floats each, row-major): E, B (the wave
 * field; B0 and the pump are applied in the push), J of the last step, and
 * the cold-fluid velocities (NULL unless rp.cold_nc > 0).  Staggering:
 * Ex(i+1/2,j) Ey(i,j+1/2) Ez(i,j) Bx(i,j+1/2) By(i+1/2,j) Bz(i+1/2,j+1/2),
 * Jx, Jy, Jz on the E sites (include/pic/yee2d.hpp:1-31). */
typedef struct aw_yee_mesh {
    float *ex, *ey, *ez, *bx, *by, *bz, *jx, *jy, *jz;
    float *vcy, *vcz;
} aw_yee_mesh;

/* Delta-f marker weights (Particles::enable_deltaf, include/pic/particles.hpp:
 * 484-491): wd (always), wc (the Kahan compensation, rp.df_wprec = 1) and
 * wdd (FP64 accumulator, rp.df_wprec = 2); device arrays of n. */
typedef struct aw_deltaf_state {
    float* wd;
    float* wc;
    double* wdd;
} aw_deltaf_state;

/* MaxwellSimulation(g, rp): a zeroed mesh, the damping masks of rp.bnd_x and
 * the cold fluid of rp.cold_nc (ny = 1 only), on its own stream. */
aw_status aw_maxwell_create(const aw_grid* g, const aw_run_params* rp, aw_maxwell** out);
aw_status aw_maxwell_destroy(aw_maxwell* m);
aw_status aw_maxwell_stream(aw_maxwell* m, aw_stream* out);
/* Waits for the stream; AW_ERR_INVALID (once) when a step since the last call
 * had a particle outside the current's fixed-point scale: |uz/gamma| above
 * 2 max(1, c), or a weight above the largest the run's weights had when the
 * array was first stepped. */
aw_status aw_maxwell_synchronize(aw_maxwell* m);
aw_status aw_maxwell_mesh(aw_maxwell* m, aw_yee_mesh* out);
/* MaxwellSimulation::step / step_at (:51-110): B half step; staggered gather
 * (+ pump, + B0 or the rp.b0_prof mirror field), Boris (rp.rel: relativistic),
 * delta-f weight update (rp.deltaf, df may be NULL otherwise), move, walls
 * (rp.bnd_x); the charge-conserving current of the move (int64 fixed point,
 * independent of thread and particle order); periodic wrap; rp.jfilter
 * binomial passes; antenna and cold-fluid currents; B half step; E full
 * step; x damping.  The caller's particle arrays are updated in place; with
 * rp.tile_sort = N > 0 they are stably re-sorted into 16x16 tiles every N
 * steps (with the delta-f weight; df_wprec must be 0) and the push and the
 * current run as one kernel per tile (k_push_esirkepov_tiled,
 * include/pic/yee2d.hpp:401-469), else they are never reordered (large
 * grids with >= 8 particles per cell run the same kernel over tile bins
 * rebuilt every 8 steps).  step uses t = (steps so far) * dt. */
aw_status aw_maxwell_step(aw_maxwell* m, aw_particles* p, const aw_deltaf_state* df);
aw_status aw_maxwell_step_at(aw_maxwell* m, aw_particles* p, const aw_deltaf_state* df,
                             double tnow);
/* MaxwellSimulation::field_energy (:112-120): eps0/2 sum |E|^2 dA and
 * c^2/2 sum |B|^2 dA (FP64, deterministic; blocking). */
aw_status aw_maxwell_field_energy(aw_maxwell* m, double* we, double* wb);
/* MaxwellSimulation::residuals (:122-161): rms div B and the rms drift of the
 * Gauss residual div E - rho since the first call (rho the CIC node charge of
 * p, times wd when wd is non-NULL, smoothed like J).  Blocking. */
aw_status aw_maxwell_residuals(aw_maxwell* m, const aw_particles* p, const float* wd,
                               double* divb_rms, double* gauss_drift_rms);
/* The Fields the diagnostics modules read on the Yee branch (tools/arcsim.cu:
 * 121-142): Ex..Ez (wave + pump at time t), Bx..Bz (wave + B0), ELx = wave
 * Ex; the other members NULL.  The arrays live in the engine until the next
 * call. */
aw_status aw_maxwell_diag_fields(aw_maxwell* m, double t, aw_fields* out);
/* MaxwellSimulation::wd_stats (:166-181): sum wd, rms wd, max |wd| (FP64,
 * deterministic; blocking). */
aw_status aw_wd_stats(const float* wd, int64_t n, double* sum, double* rms, double* max_abs,
                      aw_stream s);
/* Particles::initialize_mirror (include/pic/particles.hpp:573-612): the
 * marker count of the mirror-equilibrium load of species q in the rp.b0_prof
 * background (ny = 1, uth[1] == uth[2]), then the load itself into p (p->n
 * must equal the count; p->w receives the equal weights when non-NULL). */
aw_status aw_mirror_count(const aw_grid* g, const aw_run_params* rp, const aw_species* q,
                          int64_t* n);
aw_status aw_load_mirror(const aw_grid* g, const aw_run_params* rp, const aw_species* q,
                         aw_particles* p, aw_stream s);

/* ---- batched ensembles ------------------------------------------------------------ */

/* No reference counterpart (the reference runs one Simulation per process):
 * many independent electrostatic runs stepped together.  Runs are grouped
 * by grid size; one step of the ensemble is one fixed-point conversion over
 * every run, one batched R2C / k-space / C2R per grid size and ONE fused
 * push+deposit launch over every run; optionally the step is captured once
 * in a HIP graph and replayed.  Each run keeps its own grid, RunParams and
 * species (one weight per run), with the same results as an aw_sim of the
 * same run.  Requirements: every run shares rp.eps0 and rp.dump_every. */
typedef struct aw_ensemble aw_ensemble;

typedef struct aw_run_desc {
    aw_grid grid;
    aw_run_params rp;
    const aw_species* species; /* read during aw_ensemble_create only */
    int32_t nspecies;
    int32_t has_b0;            /* SimConfig::has_b0 of this run */
    int32_t n_modes;
    double mode_amp[32];       /* x-displacement modes, as aw_sim_config */
} aw_run_desc;

typedef struct aw_ensemble_stats {
    int32_t nruns, ngroups;
    int64_t particles, cells;
    int32_t workgroups, graph;
    int64_t sorts, graph_launches;
    int64_t run_sorts; /* runs re-sorted, summed over the sorts (a sort on the
                          stray threshold re-sorts only the runs beyond it) */
} aw_ensemble_stats;

/* cfg: tile_x (0 -> 8 at >= 64 particles per cell, else 16), blocks_per_tile,
 * sort_every (0: re-sort only when strays exceed stray_frac), stray_frac
 * (0 -> 0.02), autotune (1: each run takes the tile side and workgroups per
 * tile that the autotuner picks for it stepped alone -- tuned once per
 * distinct grid / particle count / weights / B0 shape, cached for the
 * process; explicit tile_x / blocks_per_tile still win); the other members
 * are ignored.  The choice changes no result (order-free deposits).
 * aw_ensemble_init may be called once per ensemble.
 * cfg->field_model = AW_MODEL_DARWIN builds a Darwin ensemble: a uniform
 * batch whose runs share the grid, the species layout (count, ppc, density)
 * and every step parameter (dt, qm, eps0, c, n0, B0, pump, dump_every) and
 * differ in their loads (velocities, seed, mode amplitudes).  Each particle
 * kernel and each batched rocFFT call spans all runs (blockIdx.y = run); cfg
 * then also takes current_deposit and autotune, and graphs are not used.
 * Results equal an aw_sim of the same run bit for bit. */
aw_status aw_ensemble_create(const aw_run_desc* runs, int32_t nruns, const aw_sim_config* cfg,
                             aw_ensemble** out);
aw_status aw_ensemble_destroy(aw_ensemble* e);
/* 1: capture the step into HIP graphs at aw_ensemble_init and replay them. */
aw_status aw_ensemble_set_graph(aw_ensemble* e, int32_t on);
/* Loads every run, t=0 solve, leapfrog half step back. */
aw_status aw_ensemble_init(aw_ensemble* e);
/* Step n of every run (diagnostics when n % rp.dump_every == 0). */
aw_status aw_ensemble_step(aw_ensemble* e, int64_t n);
/* Steps first .. first+nsteps-1, then synchronises. */
aw_status aw_ensemble_run(aw_ensemble* e, int64_t first, int64_t nsteps);
aw_status aw_ensemble_synchronize(aw_ensemble* e);
aw_status aw_ensemble_info(aw_ensemble* e, aw_ensemble_stats* out);
/* Per run (index in the caller's order): its particles, fields (ES: Ex/Ey;
 * Darwin: every aw_fields grid), rho^n and DiagSample history (blocking). */
aw_status aw_ensemble_particles(aw_ensemble* e, int32_t run, aw_particles* out);
aw_status aw_ensemble_fields(aw_ensemble* e, int32_t run, aw_fields* out);
aw_status aw_ensemble_rho(aw_ensemble* e, int32_t run, float** out);
aw_status aw_ensemble_history(aw_ensemble* e, int32_t run, aw_diag_sample* out, int64_t cap,
                              int64_t* count);

/* ---- adjoint (gradients of the field energy) -------------------------------------- */

/* No reference counterpart (the reference computes no gradients): the
 * discrete adjoint of the electrostatic run, unmagnetised or with B0 along z
 * (rp.B0[2]; the Boris rotation is transposed in the reverse sweep).  The
 * forward run is
 * this library's ES step (fused push + fixed-point deposit, spectral Poisson
 * solve) from a species load with x-displacement modes
 * x += sum_m A_m sin(m 2 pi x / Lx) (physical units, m = 1..n_modes); the
 * loss is EE(N) = eps0/2 sum |E^N|^2 dx dy.  The reverse sweep recomputes
 * each segment of checkpoint_every steps from its checkpoint.  One particle
 * weight for all species; uz does not enter the run. */
typedef struct aw_adjoint aw_adjoint;

/* The reverse sweep's pieces, on caller arrays (no reference counterpart),
 * for callers that write their own reverse loops:
 *   aw_adj_push_reverse: the transpose of one push (Pusher::boris,
 *     include/pic/pusher.hpp:35-80,112-129) at positions p (x^n) with the
 *     step's field Ex, Ey: lu += (dt/dx) lx (the move; not for half_back),
 *     le = h (I + R^T) lu and lu = R^T lu (the Boris update, h = qm dt_eff/2,
 *     R the rotation of B0 z when has_b0, dt_eff = dt or -dt/2 for the half
 *     step back), lx += (de/dx)^T le; writes le (lex, ley);
 *   aw_adj_gather_t: lE += sum_p W(x_p) le_p, the transpose of the CIC gather
 *     (a deposit of per-particle vectors; int64 fixed point with a scale from
 *     the measured max |le|, deterministic); accumulates into lEx, lEy;
 *   aw_adj_deposit_t: lx += qm w_p/(dx dy) grad_x sum_k W_k lrho_k, the
 *     transpose of the charge deposit (Depositor::charge) with respect to the
 *     positions.
 * Positions in cells; gradients with respect to x in cells. */
aw_status aw_adj_push_reverse(aw_ctx* ctx, const aw_particles* p, const float* Ex,
                              const float* Ey, const aw_run_params* rp, int32_t has_b0,
                              int32_t half_back, float* lx, float* ly, float* lux, float* luy,
                              float* lex, float* ley, aw_stream s);
aw_status aw_adj_gather_t(aw_ctx* ctx, const aw_particles* p, const float* lex, const float* ley,
                          float* lEx, float* lEy, aw_stream s);
aw_status aw_adj_deposit_t(aw_ctx* ctx, const aw_particles* p, const float* lrho, double qm,
                           float* lx, float* ly, aw_stream s);

/* tile: 0 -> 8 at >= 64 particles per cell, else 16. */
aw_status aw_adj_create(const aw_grid* g, const aw_run_params* rp, const aw_species* sp,
                        int32_t nspecies, int32_t n_modes, int32_t checkpoint_every,
                        int32_t tile, aw_adjoint** out);
aw_status aw_adj_destroy(aw_adjoint* a);
/* Loads with amplitudes mode_amp[n_modes] and drifts ufl[3*nspecies] (NULL:
 * the species' own), runs nsteps steps keeping checkpoints, returns the loss. */
aw_status aw_adj_forward(aw_adjoint* a, const double* mode_amp, const double* ufl,
                         int64_t nsteps, double* loss);
/* dL/dA_m (n_modes values) and dL/dufl (3 per species; the z entries are 0). */
aw_status aw_adj_backward(aw_adjoint* a, double* grad_modes, double* grad_ufl);
/* After aw_adj_backward: dL/d(x, y, ux, uy) of the initial state per particle
 * (load order, x and y in cells; device arrays of n floats, NULL skips). */
aw_status aw_adj_state_grad(aw_adjoint* a, float* gx, float* gy, float* gux, float* guy);
/* After aw_adj_forward: the initial state (x, y, ux, uy before the leapfrog
 * half step back, load order) into out (out->n = particle count). */
aw_status aw_adj_initial_state(aw_adjoint* a, aw_particles* out);

#ifdef __cplusplus
}
#endif
#endif /* AW_H */
