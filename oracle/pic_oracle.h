/*
 * pic_oracle.h -- CPU restatement of ArcWarden's 2D3V PIC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * entry point and the bench's cpu_baseline leg load it.  The product path
 * (libarcwarden_hip.so) never links, loads or calls it.
 *
 * Two flavours of every particle operation:
 *   *_f64  -- the reference's mathematics evaluated in double precision from
 *             the same FP32 inputs (the FP64 oracle for the 1e-6 gates);
 *   *_f32  -- the exact FP32 operation sequence the HIP kernels execute
 *             (explicit fmaf, no contraction), so integer outputs (cells,
 *             tile keys, permutations) and particle state can be compared
 *             bit for bit against a scalar host run.
 *
 * Conventions follow the reference: positions in cell units, row-major
 * [ny][nx] fields with x fastest (include/pic/grid.hpp:14-29), R2C half
 * spectrum [ny][nx/2+1] (include/pic/grid.hpp:20-27).
 */
#ifndef PIC_ORACLE_H
#define PIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t nx, ny;
    double  dx, dy;
} orc_grid;

/* uniform-B0 Boris parameters, precomputed in double and rounded once */
typedef struct {
    double qm, dt;
    float  B0[3];
    int32_t has_b0;
} orc_push_params;

/* ---- integer layer: cells, tiles, stable counting sort ------------------ */
int32_t orc_cell_of(const orc_grid* g, float x, float y);
void orc_cell_index(const orc_grid* g, const float* x, const float* y, int64_t n,
                    int32_t* cell);
void orc_tile_ids(const orc_grid* g, int32_t tx, int32_t ty, const float* x,
                  const float* y, int64_t n, uint32_t* tile);
void orc_cell_keys(const orc_grid* g, int32_t tx, int32_t ty, const float* x,
                   const float* y, int64_t n, uint32_t* key);
void orc_counting_sort(const uint32_t* key, int64_t n, uint32_t nkeys,
                       int32_t* perm, int32_t* off);
void orc_wrap_f32(const orc_grid* g, float* x, float* y, int64_t n);
void orc_migrate_f32(const orc_grid* g, float* x, float* y, int32_t* cell, int64_t n);

/* ---- deposit ------------------------------------------------------------ */
void orc_deposit_rho_f64(const orc_grid* g, const float* x, const float* y,
                         const float* w, double w_uniform, int64_t n, double qm,
                         double* rho);
void orc_deposit_rho_fix(const orc_grid* g, const float* x, const float* y,
                         const float* w, double w_uniform, int64_t n, double qm,
                         int32_t scale_log2, int64_t* acc);
void orc_fix_to_float(const int64_t* acc, int64_t ncell, int32_t scale_log2, float* out);
void orc_deposit_rho_j_f64(const orc_grid* g, const float* x, const float* y,
                           const float* ux, const float* uy, const float* uz,
                           const float* w, double w_uniform, int64_t n, double qm,
                           double* rho, double* jx, double* jy, double* jz);
void orc_deposit_amu_f64(const orc_grid* g, const float* x, const float* y,
                         const float* ux, const float* uy, const float* uz,
                         const float* w, double w_uniform, int64_t n, double qm,
                         double* amu0, double* amu1, double* amu2, double* amu3);
void orc_deposit_dcu_f64(const orc_grid* g, const float* x, const float* y,
                         const float* ux, const float* uy, const float* uz,
                         const float* w, double w_uniform, int64_t n, double qm,
                         const float* Ex, const float* Ey, const float* Ez,
                         const float* Bx, const float* By, const float* Bz,
                         double* dcx, double* dcy, double* dcz);

/* ---- push --------------------------------------------------------------- */
void orc_push_es_f32(const orc_grid* g, const orc_push_params* pp, float* x, float* y,
                     float* ux, float* uy, float* uz, int64_t n, const float* Ex,
                     const float* Ey, int32_t wrap);
void orc_push_es_f64(const orc_grid* g, const orc_push_params* pp, double* x, double* y,
                     double* ux, double* uy, double* uz, int64_t n, const double* Ex,
                     const double* Ey, int32_t wrap);
void orc_half_back_f32(const orc_grid* g, const orc_push_params* pp, const float* x,
                       const float* y, float* ux, float* uy, float* uz, int64_t n,
                       const float* Ex, const float* Ey);
void orc_push_em_f32(const orc_grid* g, double qm, double dt, float* x, float* y, float* ux,
                     float* uy, float* uz, int64_t n, const float* Ex, const float* Ey,
                     const float* Ez, const float* Bx, const float* By, const float* Bz);
void orc_push_em_f64(const orc_grid* g, double qm, double dt, double* x, double* y,
                     double* ux, double* uy, double* uz, int64_t n,
                     const double* Ex, const double* Ey, const double* Ez,
                     const double* Bx, const double* By, const double* Bz, int32_t wrap);

/* ---- spectral (double) --------------------------------------------------- */
void orc_fft2_r2c(int32_t nx, int32_t ny, const double* in, double* out_re_im);
void orc_fft2_c2r(int32_t nx, int32_t ny, const double* in_re_im, double* out);
void orc_dft2_r2c_brute(int32_t nx, int32_t ny, const double* in, double* out_re_im);
void orc_solve_es_f64(const orc_grid* g, double eps0, const double* rho, double* Ex,
                      double* Ey);
void orc_solve_darwin_b_f64(const orc_grid* g, double eps0, double c, const double* jx,
                            const double* jy, const double* jz, double* bx, double* by,
                            double* bz, int32_t staggered);
void orc_solve_darwin_et_f64(const orc_grid* g, double eps0, double c, double n0,
                             const double* dcx, const double* dcy, const double* dcz,
                             const double* a0, const double* a1, const double* a2,
                             const double* a3, double* etx, double* ety, double* etz);

/* ---- Esirkepov (charge-conserving current) ------------------------------ */
void orc_esirkepov_f64(const orc_grid* g, const float* x0, const float* y0,
                       const float* x1, const float* y1, const float* uz,
                       const float* w, double w_uniform, int64_t n, double qm, double dt,
                       double* jx, double* jy, double* jz);
void orc_rho_nodes_f64(const orc_grid* g, const float* x, const float* y, const float* w,
                       double w_uniform, int64_t n, double qm, double* rho);

/* ---- loaders -------------------------------------------------------------- */
double orc_radical_inverse(uint32_t i, uint32_t base);
uint32_t orc_hash_u32(uint32_t v);
double orc_rng_uniform(int64_t t, uint32_t stream, uint64_t seed);
double orc_erfinv(double y);
float orc_quiet_u(int64_t t, uint32_t base, double uth, double ufl);
typedef struct {
    int32_t ppc, index;
    int64_t base, count;
    double  weight, uth[3], ufl[3];
    int32_t noisy;
    uint64_t seed;
} orc_species_init;
void orc_load_species(const orc_grid* g, const orc_species_init* sp, float* x, float* y,
                      float* ux, float* uy, float* uz);
void orc_apply_modes_f32(const orc_grid* g, float* x, int64_t n, const double* amp_cells,
                         int32_t nmodes);

/* ---- diagnostics (double) -------------------------------------------------- */
/* ---- full-Maxwell (Yee) branch ------------------------------------------ */
typedef struct {
    double qm, dt, c;
    float B0[3];
    int32_t rel, pump, deltaf, df_wprec, df_dist, b0_prof, bnd_x, bnd_nd;
    double pump_ex, pump_ey, pump_ez, pump_k0, pump_w0, pump_trmp, pump_toff;
    double df_tpar, df_tperp, df_kappa, df_rho, df_taud;
    double b0_a, b0_xc, b0_sref, b0_c[6];
    double bnd_numax;
} orc_yee_params;
void orc_yee_faraday_f32(const orc_grid* g, float* F, float h);
void orc_yee_ampere_f32(const orc_grid* g, float* F, float c2, float dt);
void orc_yee_binomial_f32(const orc_grid* g, const float* src, float* dst);
float orc_b0_profile(const orc_yee_params* p, float xph);
void orc_yee_antenna(const orc_grid* g, double amp, double x0, double sigma, double w0,
                     double trmp, double toff, double t, float* F);
void orc_yee_cold_fluid_f32(const orc_grid* g, const orc_yee_params* p, double cold_nc,
                            float* F, float* vcy, float* vcz);
void orc_yee_damp_f32(const orc_grid* g, int32_t nd, double numax, double dt, float* F,
                      float* vcy, float* vcz);
void orc_yee_push_f32(const orc_grid* g, const orc_yee_params* p, double tnow, const float* F,
                      float* x, float* y, float* ux, float* uy, float* uz, const float* w,
                      float w_uniform, float* wd, float* wc, double* wdd, float* x0o,
                      float* y0o, float* vzo, float* wdep, int64_t n);
void orc_yee_energy_f64(const orc_grid* g, const float* F, double c2, double* out2);
double orc_yee_divb_rms(const orc_grid* g, const float* F);
void orc_yee_div_e(const orc_grid* g, const float* F, double* out);
int64_t orc_mirror_offsets(const orc_grid* g, const orc_yee_params* p, int32_t ppc,
                           double uth_par, double uth_perp, int32_t dist, double lc_rho,
                           double lc_kappa, int32_t* off);
void orc_phase_hist(const orc_grid* g, const float* x, const float* ux, const float* uz,
                    int64_t n, int32_t nxb, int32_t nvb, double bx, double bz, double vth,
                    double vlo, double vhi, int32_t fp64, uint32_t* counts);
void orc_energies(const orc_grid* g, const float* ux, const float* uy, const float* uz,
                  const float* w, double w_uniform, int64_t n, const float* Ex,
                  const float* Ey, const float* rho, double eps0, double* out5);

/* ---- whole ES step, same-algorithm FP32 path (CPU baseline) --------------- */
int32_t orc_es_step_f32(const orc_grid* g, const orc_push_params* pp, double eps0,
                        float* x, float* y, float* ux, float* uy, float* uz, int64_t n,
                        double w_uniform, int32_t scale_log2, float* rho, float* Ex,
                        float* Ey, int32_t nthreads);
int32_t orc_max_threads(void);

/* ---- electrostatic run in FP64 and its discrete adjoint (no B0) ---------- */
void orc_solve_es_adj_f64(const orc_grid* g, double eps0, const double* lEx,
                          const double* lEy, double* lrho);
double orc_es_adjoint_f64(const orc_grid* g, double qm, double dt, double eps0, double b0z,
                          double w_uniform, const double* x0, const double* y0,
                          const double* ux0, const double* uy0, int64_t n, int32_t nsteps,
                          double* gx, double* gy, double* gux, double* guy, int32_t nthreads);

#ifdef __cplusplus
}
#endif
#endif
