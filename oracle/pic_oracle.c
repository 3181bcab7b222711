/*
 * pic_oracle.c -- CPU restatement of ArcWarden's 2D3V PIC hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see pic_oracle.h).  Each function names the
 * reference file:line whose behaviour it restates.  Build with
 * -ffp-contract=off so the FP32 "same algorithm" functions evaluate exactly
 * the operation sequence written here (fused multiply-adds are explicit
 * fmaf calls, as in the HIP kernels).
 */
#include "pic_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TWO_PI 6.28318530717958647692

/* ------------------------------------------------------------------------ */
/* integer layer                                                             */
/* ------------------------------------------------------------------------ */

/* Owning cell of a wrapped position: floor + clamp of the x==nx rounding
 * corner (include/pic/particles.hpp:310-314). */
int32_t orc_cell_of(const orc_grid* g, float x, float y) {
    int32_t ci = (int32_t)floorf(x), cj = (int32_t)floorf(y);
    if (ci >= g->nx) ci = g->nx - 1;
    if (cj >= g->ny) cj = g->ny - 1;
    if (ci < 0) ci = 0;
    if (cj < 0) cj = 0;
    return cj * g->nx + ci;
}

void orc_cell_index(const orc_grid* g, const float* x, const float* y, int64_t n,
                    int32_t* cell) {
    for (int64_t i = 0; i < n; ++i) cell[i] = orc_cell_of(g, x[i], y[i]);
}

/* Tile id of a cell (include/pic/particles.hpp:74-77 tile_of):
 * (cj/ty)*ntx + ci/tx with ntx = ceil(nx/tx). */
void orc_tile_ids(const orc_grid* g, int32_t tx, int32_t ty, const float* x,
                  const float* y, int64_t n, uint32_t* tile) {
    const int32_t ntx = (g->nx + tx - 1) / tx;
    for (int64_t i = 0; i < n; ++i) {
        const int32_t c = orc_cell_of(g, x[i], y[i]);
        const int32_t ci = c % g->nx, cj = c / g->nx;
        tile[i] = (uint32_t)((cj / ty) * ntx + ci / tx);
    }
}

/* Tile-major cell key: tile*(tx*ty) + in-tile row-major cell. */
void orc_cell_keys(const orc_grid* g, int32_t tx, int32_t ty, const float* x,
                   const float* y, int64_t n, uint32_t* key) {
    const int32_t ntx = (g->nx + tx - 1) / tx;
    for (int64_t i = 0; i < n; ++i) {
        const int32_t c = orc_cell_of(g, x[i], y[i]);
        const int32_t ci = c % g->nx, cj = c / g->nx;
        const uint32_t t = (uint32_t)((cj / ty) * ntx + ci / tx);
        key[i] = t * (uint32_t)(tx * ty) + (uint32_t)((cj % ty) * tx + (ci % tx));
    }
}

/* Stable counting sort.  The reference's counting sort
 * (include/pic/particles.hpp:84-153) assigns slots in atomic order; the
 * build defines the permutation as the STABLE one, so it is unique.
 * perm[k] = index of the k-th particle in key order; off[nkeys+1]. */
void orc_counting_sort(const uint32_t* key, int64_t n, uint32_t nkeys,
                       int32_t* perm, int32_t* off) {
    int64_t* cnt = (int64_t*)calloc((size_t)nkeys + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) cnt[key[i] + 1]++;
    for (uint32_t k = 0; k < nkeys; ++k) cnt[k + 1] += cnt[k];
    if (off)
        for (uint32_t k = 0; k <= nkeys; ++k) off[k] = (int32_t)cnt[k];
    for (int64_t i = 0; i < n; ++i) perm[cnt[key[i]]++] = (int32_t)i;
    free(cnt);
}

/* Periodic wrap of one coordinate with the reference's semantics
 * (include/pic/particles.hpp:301-306): fmodf, +n if negative, and the
 * "+= wrap rounds to n exactly" guard.  The one-period fast paths are exact
 * rewrites (fmodf(v,n)==v-n on [n,2n), ==v on (-n,0)). */
static inline float wrap_coord(float v, float n) {
    if (v >= n) {
        v = (v < 2.0f * n) ? v - n : fmodf(v, n);
    } else if (v < 0.0f) {
        if (!(v > -n)) v = fmodf(v, n);
        if (v < 0.0f) v += n;
    }
    if (v >= n) v = 0.0f;
    return v;
}

void orc_wrap_f32(const orc_grid* g, float* x, float* y, int64_t n) {
    const float fnx = (float)g->nx, fny = (float)g->ny;
    for (int64_t i = 0; i < n; ++i) {
        x[i] = wrap_coord(x[i], fnx);
        y[i] = wrap_coord(y[i], fny);
    }
}

/* particle_migrate_kernel (include/pic/particles.hpp:293-315). */
void orc_migrate_f32(const orc_grid* g, float* x, float* y, int32_t* cell, int64_t n) {
    orc_wrap_f32(g, x, y, n);
    if (cell) orc_cell_index(g, x, y, n, cell);
}

/* ------------------------------------------------------------------------ */
/* CIC deposit                                                               */
/* ------------------------------------------------------------------------ */

static inline int32_t wrapi(int32_t i, int32_t n) {
    if (i >= n) i -= n;
    if (i < 0) i += n;
    return i;
}

/* cic_stencil (include/pic/grid.hpp:140-155) with weights in double. */
typedef struct {
    int32_t c[4];
    double  w[4];
} cic_d;

static inline cic_d cic_double(const orc_grid* g, float x, float y) {
    cic_d s;
    const int32_t i0 = (int32_t)floorf(x), j0 = (int32_t)floorf(y);
    const double fx = (double)x - (double)i0, fy = (double)y - (double)j0;
    const double gx = 1.0 - fx, gy = 1.0 - fy;
    const int32_t a = wrapi(i0, g->nx), b = wrapi(i0 + 1, g->nx);
    const int32_t c = wrapi(j0, g->ny), d = wrapi(j0 + 1, g->ny);
    s.c[0] = c * g->nx + a; s.w[0] = gx * gy;
    s.c[1] = c * g->nx + b; s.w[1] = fx * gy;
    s.c[2] = d * g->nx + a; s.w[2] = gx * fy;
    s.c[3] = d * g->nx + b; s.w[3] = fx * fy;
    return s;
}

static inline double coef_of(const float* w, double w_uniform, int64_t i, double qm,
                             double area) {
    return w ? qm * (double)w[i] / area : qm * w_uniform / area;
}

/* deposit_charge_kernel (include/pic/depositor.hpp:34-58): rho += q*w/(dx*dy)*W. */
void orc_deposit_rho_f64(const orc_grid* g, const float* x, const float* y,
                         const float* w, double w_uniform, int64_t n, double qm,
                         double* rho) {
    const double area = g->dx * g->dy;
    for (int64_t i = 0; i < n; ++i) {
        const cic_d s = cic_double(g, x[i], y[i]);
        const double coef = coef_of(w, w_uniform, i, qm, area);
        for (int k = 0; k < 4; ++k) rho[s.c[k]] += coef * s.w[k];
    }
}

/* Same deposit, accumulated the way the HIP kernels do: every contribution
 * coef*W_k (one double rounding) is quantised to a 2^-scale_log2 fixed-point
 * integer and summed exactly, so the total is independent of order. */
void orc_deposit_rho_fix(const orc_grid* g, const float* x, const float* y,
                         const float* w, double w_uniform, int64_t n, double qm,
                         int32_t scale_log2, int64_t* acc) {
    const double area = g->dx * g->dy;
    const double scale = ldexp(1.0, scale_log2);
    for (int64_t i = 0; i < n; ++i) {
        const cic_d s = cic_double(g, x[i], y[i]);
        const double coef = coef_of(w, w_uniform, i, qm, area);
        for (int k = 0; k < 4; ++k) acc[s.c[k]] += (int64_t)rint(coef * s.w[k] * scale);
    }
}

void orc_fix_to_float(const int64_t* acc, int64_t ncell, int32_t scale_log2, float* out) {
    const double inv = ldexp(1.0, -scale_log2);
    for (int64_t c = 0; c < ncell; ++c) out[c] = (float)((double)acc[c] * inv);
}

/* deposit_rho_j_tiled_kernel (include/pic/depositor.hpp:160-221):
 * rho and J = coef*W*u (gamma == 1). */
void orc_deposit_rho_j_f64(const orc_grid* g, const float* x, const float* y,
                           const float* ux, const float* uy, const float* uz,
                           const float* w, double w_uniform, int64_t n, double qm,
                           double* rho, double* jx, double* jy, double* jz) {
    const double area = g->dx * g->dy;
    for (int64_t i = 0; i < n; ++i) {
        const cic_d s = cic_double(g, x[i], y[i]);
        const double coef = coef_of(w, w_uniform, i, qm, area);
        for (int k = 0; k < 4; ++k) {
            const double cw = coef * s.w[k];
            rho[s.c[k]] += cw;
            jx[s.c[k]] += cw * ux[i];
            jy[s.c[k]] += cw * uy[i];
            jz[s.c[k]] += cw * uz[i];
        }
    }
}

/* deposit_amu_kernel (include/pic/depositor.hpp:230-270):
 * amu0=(vx^2-vy^2)/2, amu1=vx vy, amu2=vz vx, amu3=vz vy, times coef*W. */
void orc_deposit_amu_f64(const orc_grid* g, const float* x, const float* y,
                         const float* ux, const float* uy, const float* uz,
                         const float* w, double w_uniform, int64_t n, double qm,
                         double* amu0, double* amu1, double* amu2, double* amu3) {
    const double area = g->dx * g->dy;
    for (int64_t i = 0; i < n; ++i) {
        const cic_d s = cic_double(g, x[i], y[i]);
        const double coef = coef_of(w, w_uniform, i, qm, area);
        const double vx = ux[i], vy = uy[i], vz = uz[i];
        const double a0 = 0.5 * (vx * vx - vy * vy), a1 = vx * vy;
        const double a2 = vz * vx, a3 = vz * vy;
        for (int k = 0; k < 4; ++k) {
            const double cw = coef * s.w[k];
            amu0[s.c[k]] += cw * a0;
            amu1[s.c[k]] += cw * a1;
            amu2[s.c[k]] += cw * a2;
            amu3[s.c[k]] += cw * a3;
        }
    }
}

static inline double gather_d(const float* f, const cic_d* s) {
    return s->w[0] * f[s->c[0]] + s->w[1] * f[s->c[1]] + s->w[2] * f[s->c[2]] +
           s->w[3] * f[s->c[3]];
}

/* deposit_dcu_kernel (include/pic/depositor.hpp:275-324):
 * dcu += qm^2 w/area * (E + v x B) at the particle, CIC-scattered. */
void orc_deposit_dcu_f64(const orc_grid* g, const float* x, const float* y,
                         const float* ux, const float* uy, const float* uz,
                         const float* w, double w_uniform, int64_t n, double qm,
                         const float* Ex, const float* Ey, const float* Ez,
                         const float* Bx, const float* By, const float* Bz,
                         double* dcx, double* dcy, double* dcz) {
    const double area = g->dx * g->dy;
    for (int64_t i = 0; i < n; ++i) {
        const cic_d s = cic_double(g, x[i], y[i]);
        const double ex = gather_d(Ex, &s), ey = gather_d(Ey, &s), ez = gather_d(Ez, &s);
        const double bx = gather_d(Bx, &s), by = gather_d(By, &s), bz = gather_d(Bz, &s);
        const double vx = ux[i], vy = uy[i], vz = uz[i];
        const double coef = (w ? qm * qm * (double)w[i] : qm * qm * w_uniform) / area;
        const double ax = coef * (ex + (vy * bz - vz * by));
        const double ay = coef * (ey + (vz * bx - vx * bz));
        const double az = coef * (ez + (vx * by - vy * bx));
        for (int k = 0; k < 4; ++k) {
            dcx[s.c[k]] += ax * s.w[k];
            dcy[s.c[k]] += ay * s.w[k];
            dcz[s.c[k]] += az * s.w[k];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Boris push                                                                */
/* ------------------------------------------------------------------------ */

/* Rotation vectors for a uniform B0 (include/pic/pusher.hpp:58-75):
 * t = (q/m)(dt_eff/2) B0, s = 2t/(1+|t|^2); evaluated in double, rounded once. */
typedef struct {
    float qmh, t[3], s[3];
} boris_consts;

static boris_consts boris_make(double qm, double dt_eff, const float B0[3], int has_b0) {
    boris_consts b;
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

typedef struct {
    int32_t c[4];
    float   w[4];
} cic_f;

/* cic_stencil (include/pic/grid.hpp:140-155), FP32 weights. */
static inline cic_f cic_float(const orc_grid* g, float x, float y) {
    cic_f s;
    const int32_t i0 = (int32_t)floorf(x), j0 = (int32_t)floorf(y);
    const float fx = x - (float)i0, fy = y - (float)j0;
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    const int32_t a = wrapi(i0, g->nx), b = wrapi(i0 + 1, g->nx);
    const int32_t c = wrapi(j0, g->ny), d = wrapi(j0 + 1, g->ny);
    s.c[0] = c * g->nx + a; s.w[0] = gx * gy;
    s.c[1] = c * g->nx + b; s.w[1] = fx * gy;
    s.c[2] = d * g->nx + a; s.w[2] = gx * fy;
    s.c[3] = d * g->nx + b; s.w[3] = fx * fy;
    return s;
}

/* gather_E (include/pic/pusher.hpp:35-43) as an explicit fma chain. */
static inline float gather_f(const float* f, const cic_f* s) {
    return fmaf(s->w[3], f[s->c[3]],
                fmaf(s->w[2], f[s->c[2]], fmaf(s->w[1], f[s->c[1]], s->w[0] * f[s->c[0]])));
}

/* boris_velocity_update (include/pic/pusher.hpp:47-80); the ES path has no
 * Ez, so uz is only rotated. */
static inline void boris_es_f32(const boris_consts* b, int has_b0, float ex, float ey,
                                float* pux, float* puy, float* puz) {
    float ux = *pux, uy = *puy, uz = *puz;
    ux = fmaf(b->qmh, ex, ux);
    uy = fmaf(b->qmh, ey, uy);
    if (has_b0) {
        const float tx = b->t[0], ty = b->t[1], tz = b->t[2];
        const float sx = b->s[0], sy = b->s[1], sz = b->s[2];
        const float upx = fmaf(uy, tz, fmaf(-uz, ty, ux));
        const float upy = fmaf(uz, tx, fmaf(-ux, tz, uy));
        const float upz = fmaf(ux, ty, fmaf(-uy, tx, uz));
        ux = fmaf(upy, sz, fmaf(-upz, sy, ux));
        uy = fmaf(upz, sx, fmaf(-upx, sz, uy));
        uz = fmaf(upx, sy, fmaf(-upy, sx, uz));
    }
    ux = fmaf(b->qmh, ex, ux);
    uy = fmaf(b->qmh, ey, uy);
    *pux = ux; *puy = uy; *puz = uz;
}

/* boris_push_kernel (include/pic/pusher.hpp:112-129): gather, Boris over dt,
 * move x += ux*dt/dx in cell units; wrap!=0 also applies the migrate wrap. */
void orc_push_es_f32(const orc_grid* g, const orc_push_params* pp, float* x, float* y,
                     float* ux, float* uy, float* uz, int64_t n, const float* Ex,
                     const float* Ey, int32_t wrap) {
    const boris_consts b = boris_make(pp->qm, pp->dt, pp->B0, pp->has_b0);
    const float cx = (float)(pp->dt / g->dx), cy = (float)(pp->dt / g->dy);
    const float fnx = (float)g->nx, fny = (float)g->ny;
    for (int64_t i = 0; i < n; ++i) {
        const cic_f s = cic_float(g, x[i], y[i]);
        const float ex = gather_f(Ex, &s), ey = gather_f(Ey, &s);
        boris_es_f32(&b, pp->has_b0, ex, ey, &ux[i], &uy[i], &uz[i]);
        float xn = fmaf(ux[i], cx, x[i]), yn = fmaf(uy[i], cy, y[i]);
        if (wrap) {
            xn = wrap_coord(xn, fnx);
            yn = wrap_coord(yn, fny);
        }
        x[i] = xn;
        y[i] = yn;
    }
}

/* boris_half_back_kernel (include/pic/pusher.hpp:133-145): velocity-only
 * Boris over -dt/2 for the leapfrog start. */
void orc_half_back_f32(const orc_grid* g, const orc_push_params* pp, const float* x,
                       const float* y, float* ux, float* uy, float* uz, int64_t n,
                       const float* Ex, const float* Ey) {
    const boris_consts b = boris_make(pp->qm, -0.5 * pp->dt, pp->B0, pp->has_b0);
    for (int64_t i = 0; i < n; ++i) {
        const cic_f s = cic_float(g, x[i], y[i]);
        const float ex = gather_f(Ex, &s), ey = gather_f(Ey, &s);
        boris_es_f32(&b, pp->has_b0, ex, ey, &ux[i], &uy[i], &uz[i]);
    }
}

static inline double wrap_d(double v, double n) {
    v = fmod(v, n);
    if (v < 0.0) v += n;
    if (v >= n) v = 0.0;
    return v;
}

static inline cic_d cic_double_d(const orc_grid* g, double x, double y) {
    cic_d s;
    const int32_t i0 = (int32_t)floor(x), j0 = (int32_t)floor(y);
    const double fx = x - i0, fy = y - j0, gx = 1.0 - fx, gy = 1.0 - fy;
    const int32_t a = wrapi(i0, g->nx), b = wrapi(i0 + 1, g->nx);
    const int32_t c = wrapi(j0, g->ny), d = wrapi(j0 + 1, g->ny);
    s.c[0] = c * g->nx + a; s.w[0] = gx * gy;
    s.c[1] = c * g->nx + b; s.w[1] = fx * gy;
    s.c[2] = d * g->nx + a; s.w[2] = gx * fy;
    s.c[3] = d * g->nx + b; s.w[3] = fx * fy;
    return s;
}

static inline double gather_dd(const double* f, const cic_d* s) {
    return s->w[0] * f[s->c[0]] + s->w[1] * f[s->c[1]] + s->w[2] * f[s->c[2]] +
           s->w[3] * f[s->c[3]];
}

/* Boris rotation about t, s (include/pic/pusher.hpp:85-97) in double. */
static inline void rotate_d(double* u, const double t[3], const double s[3]) {
    const double upx = u[0] + (u[1] * t[2] - u[2] * t[1]);
    const double upy = u[1] + (u[2] * t[0] - u[0] * t[2]);
    const double upz = u[2] + (u[0] * t[1] - u[1] * t[0]);
    u[0] += upy * s[2] - upz * s[1];
    u[1] += upz * s[0] - upx * s[2];
    u[2] += upx * s[1] - upy * s[0];
}

/* FP64 restatement of the ES push (include/pic/pusher.hpp:47-80,112-129). */
void orc_push_es_f64(const orc_grid* g, const orc_push_params* pp, double* x, double* y,
                     double* ux, double* uy, double* uz, int64_t n, const double* Ex,
                     const double* Ey, int32_t wrap) {
    const double qmh = pp->qm * 0.5 * pp->dt;
    double t[3], s[3], t2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        t[k] = pp->has_b0 ? qmh * (double)pp->B0[k] : 0.0;
        t2 += t[k] * t[k];
    }
    for (int k = 0; k < 3; ++k) s[k] = 2.0 * t[k] / (1.0 + t2);
    for (int64_t i = 0; i < n; ++i) {
        const cic_d st = cic_double_d(g, x[i], y[i]);
        const double ex = gather_dd(Ex, &st), ey = gather_dd(Ey, &st);
        double u[3] = {ux[i] + qmh * ex, uy[i] + qmh * ey, uz[i]};
        if (pp->has_b0) rotate_d(u, t, s);
        u[0] += qmh * ex;
        u[1] += qmh * ey;
        ux[i] = u[0]; uy[i] = u[1]; uz[i] = u[2];
        x[i] += u[0] * pp->dt / g->dx;
        y[i] += u[1] * pp->dt / g->dy;
        if (wrap) {
            x[i] = wrap_d(x[i], g->nx);
            y[i] = wrap_d(y[i], g->ny);
        }
    }
}

/* em_push_tiled_kernel (include/pic/pusher.hpp:211-258) + boris_update_full
 * (:103-109): gathered E(3), B(3); full Boris; move. FP64 restatement. */
void orc_push_em_f64(const orc_grid* g, double qm, double dt, double* x, double* y,
                     double* ux, double* uy, double* uz, int64_t n,
                     const double* Ex, const double* Ey, const double* Ez,
                     const double* Bx, const double* By, const double* Bz, int32_t wrap) {
    const double qmh = qm * 0.5 * dt;
    for (int64_t i = 0; i < n; ++i) {
        const cic_d st = cic_double_d(g, x[i], y[i]);
        const double e[3] = {gather_dd(Ex, &st), gather_dd(Ey, &st), gather_dd(Ez, &st)};
        const double b[3] = {gather_dd(Bx, &st), gather_dd(By, &st), gather_dd(Bz, &st)};
        double u[3] = {ux[i] + qmh * e[0], uy[i] + qmh * e[1], uz[i] + qmh * e[2]};
        double t[3], s[3], t2 = 0.0;
        for (int k = 0; k < 3; ++k) { t[k] = qmh * b[k]; t2 += t[k] * t[k]; }
        for (int k = 0; k < 3; ++k) s[k] = 2.0 * t[k] / (1.0 + t2);
        rotate_d(u, t, s);
        for (int k = 0; k < 3; ++k) u[k] += qmh * e[k];
        ux[i] = u[0]; uy[i] = u[1]; uz[i] = u[2];
        x[i] += u[0] * dt / g->dx;
        y[i] += u[1] * dt / g->dy;
        if (wrap) {
            x[i] = wrap_d(x[i], g->nx);
            y[i] = wrap_d(y[i], g->ny);
        }
    }
}

/* em_push_tiled_kernel (include/pic/pusher.hpp:211-258) with boris_update_full
 * (:103-109) as the FP32 sequence the HIP kernel executes (explicit fmaf);
 * E, B are the gathered grids, move x += u dt/dx, periodic wrap. */
void orc_push_em_f32(const orc_grid* g, double qm, double dt, float* x, float* y, float* ux,
                     float* uy, float* uz, int64_t n, const float* Ex, const float* Ey,
                     const float* Ez, const float* Bx, const float* By, const float* Bz) {
    const float qmh = (float)(qm * 0.5 * dt);
    const float cx = (float)(dt / g->dx), cy = (float)(dt / g->dy);
    const float fnx = (float)g->nx, fny = (float)g->ny;
    for (int64_t i = 0; i < n; ++i) {
        const cic_f s = cic_float(g, x[i], y[i]);
        const float ex = gather_f(Ex, &s), ey = gather_f(Ey, &s), ez = gather_f(Ez, &s);
        const float bx = gather_f(Bx, &s), by = gather_f(By, &s), bz = gather_f(Bz, &s);
        float u0 = ux[i], u1 = uy[i], u2 = uz[i];
        u0 = fmaf(qmh, ex, u0);
        u1 = fmaf(qmh, ey, u1);
        u2 = fmaf(qmh, ez, u2);
        const float tx = qmh * bx, ty = qmh * by, tz = qmh * bz;
        const float t2 = fmaf(tz, tz, fmaf(ty, ty, tx * tx));
        const float sf = 2.0f / (1.0f + t2);
        const float sx = sf * tx, sy = sf * ty, sz = sf * tz;
        const float upx = fmaf(u1, tz, fmaf(-u2, ty, u0));
        const float upy = fmaf(u2, tx, fmaf(-u0, tz, u1));
        const float upz = fmaf(u0, ty, fmaf(-u1, tx, u2));
        u0 = fmaf(upy, sz, fmaf(-upz, sy, u0));
        u1 = fmaf(upz, sx, fmaf(-upx, sz, u1));
        u2 = fmaf(upx, sy, fmaf(-upy, sx, u2));
        u0 = fmaf(qmh, ex, u0);
        u1 = fmaf(qmh, ey, u1);
        u2 = fmaf(qmh, ez, u2);
        ux[i] = u0; uy[i] = u1; uz[i] = u2;
        x[i] = wrap_coord(fmaf(u0, cx, x[i]), fnx);
        y[i] = wrap_coord(fmaf(u1, cy, y[i]), fny);
    }
}

/* ------------------------------------------------------------------------ */
/* FFT (double, mixed radix, unnormalised like cuFFT / rocFFT)               */
/* ------------------------------------------------------------------------ */

static int smallest_factor(int n) {
    for (int p = 2; p * p <= n; ++p)
        if (n % p == 0) return p;
    return n;
}

/* Recursive decimation in time on the smallest prime factor: exact DFT
 * definition X_k = sum_n x_n e^{sign*2 pi i n k / N}. */
static void fft1(double complex* a, int n, int sign) {
    if (n <= 1) return;
    const int p = smallest_factor(n), m = n / p;
    double complex* tmp = (double complex*)malloc(sizeof(double complex) * (size_t)n);
    for (int r = 0; r < p; ++r)
        for (int k = 0; k < m; ++k) tmp[r * m + k] = a[k * p + r];
    for (int r = 0; r < p; ++r) fft1(tmp + r * m, m, sign);
    for (int k = 0; k < n; ++k) {
        double complex acc = 0.0;
        for (int r = 0; r < p; ++r) {
            const double ang = sign * TWO_PI * (double)(((long long)r * k) % n) / n;
            acc += tmp[r * m + (k % m)] * (cos(ang) + I * sin(ang));
        }
        a[k] = acc;
    }
    free(tmp);
}

static void fft2_complex(int nx, int ny, double complex* a, int sign) {
    double complex* col = (double complex*)malloc(sizeof(double complex) * (size_t)ny);
    for (int j = 0; j < ny; ++j) fft1(a + (size_t)j * nx, nx, sign);
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) col[j] = a[(size_t)j * nx + i];
        fft1(col, ny, sign);
        for (int j = 0; j < ny; ++j) a[(size_t)j * nx + i] = col[j];
    }
    free(col);
}

/* R2C: real [ny][nx] -> half spectrum [ny][nx/2+1] (re, im interleaved);
 * forward sign -1 (include/pic/fft.hpp:26-35 / grid.hpp:14-29 layout). */
void orc_fft2_r2c(int32_t nx, int32_t ny, const double* in, double* out) {
    const int nkx = nx / 2 + 1;
    double complex* a = (double complex*)malloc(sizeof(double complex) * (size_t)nx * ny);
    for (size_t k = 0; k < (size_t)nx * ny; ++k) a[k] = in[k];
    fft2_complex(nx, ny, a, -1);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nkx; ++i) {
            out[2 * ((size_t)j * nkx + i)] = creal(a[(size_t)j * nx + i]);
            out[2 * ((size_t)j * nkx + i) + 1] = cimag(a[(size_t)j * nx + i]);
        }
    free(a);
}

/* C2R: Hermitian half spectrum -> real field, unnormalised (sign +1). */
void orc_fft2_c2r(int32_t nx, int32_t ny, const double* in, double* out) {
    const int nkx = nx / 2 + 1;
    double complex* a = (double complex*)malloc(sizeof(double complex) * (size_t)nx * ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double complex v;
            if (i < nkx) {
                v = in[2 * ((size_t)j * nkx + i)] + I * in[2 * ((size_t)j * nkx + i) + 1];
            } else {
                const int jj = (ny - j) % ny, ii = nx - i;
                v = in[2 * ((size_t)jj * nkx + ii)] - I * in[2 * ((size_t)jj * nkx + ii) + 1];
            }
            a[(size_t)j * nx + i] = v;
        }
    fft2_complex(nx, ny, a, +1);
    for (size_t k = 0; k < (size_t)nx * ny; ++k) out[k] = creal(a[k]);
    free(a);
}

/* O(N^2) DFT straight from the definition; pins the FFT on tiny grids. */
void orc_dft2_r2c_brute(int32_t nx, int32_t ny, const double* in, double* out) {
    const int nkx = nx / 2 + 1;
    for (int q = 0; q < ny; ++q)
        for (int p = 0; p < nkx; ++p) {
            double re = 0.0, im = 0.0;
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const double ang =
                        -TWO_PI * ((double)p * i / nx + (double)q * j / ny);
                    re += in[(size_t)j * nx + i] * cos(ang);
                    im += in[(size_t)j * nx + i] * sin(ang);
                }
            out[2 * ((size_t)q * nkx + p)] = re;
            out[2 * ((size_t)q * nkx + p) + 1] = im;
        }
}

/* KGrid (include/pic/grid.hpp:95-126): kx = i*2pi/Lx; ky wraps past ny/2. */
static inline void kvec(const orc_grid* g, int i, int j, double* kx, double* ky) {
    const double Lx = g->nx * g->dx, Ly = g->ny * g->dy;
    const int m = (j <= g->ny / 2) ? j : j - g->ny;
    *kx = i * (TWO_PI / Lx);
    *ky = m * (TWO_PI / Ly);
}

static inline int is_nyquist(const orc_grid* g, int i, int j) {
    return ((g->nx % 2 == 0) && i == g->nx / 2) || ((g->ny % 2 == 0) && j == g->ny / 2);
}

/* es_poisson_field_kernel (include/pic/solver_es.hpp:39-78) + the C2R
 * normalisation (include/pic/spectral.hpp:97-101,235-243), all in double. */
void orc_solve_es_f64(const orc_grid* g, double eps0, const double* rho, double* Ex,
                      double* Ey) {
    const int nx = g->nx, ny = g->ny, nkx = nx / 2 + 1;
    const size_t nk = (size_t)nkx * ny;
    double* rk = (double*)malloc(sizeof(double) * 2 * nk);
    double* exk = (double*)malloc(sizeof(double) * 2 * nk);
    double* eyk = (double*)malloc(sizeof(double) * 2 * nk);
    orc_fft2_r2c(nx, ny, rho, rk);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nkx; ++i) {
            const size_t t = (size_t)j * nkx + i;
            double kx, ky;
            kvec(g, i, j, &kx, &ky);
            const double k2 = kx * kx + ky * ky;
            const double gr = k2 > 0.0 ? 1.0 / (eps0 * k2) : 0.0;
            const double pr = rk[2 * t] * gr, pi = rk[2 * t + 1] * gr;
            if (is_nyquist(g, i, j)) {
                exk[2 * t] = exk[2 * t + 1] = eyk[2 * t] = eyk[2 * t + 1] = 0.0;
            } else {
                exk[2 * t] = kx * pi;  exk[2 * t + 1] = -kx * pr;
                eyk[2 * t] = ky * pi;  eyk[2 * t + 1] = -ky * pr;
            }
        }
    orc_fft2_c2r(nx, ny, exk, Ex);
    orc_fft2_c2r(nx, ny, eyk, Ey);
    const double s = 1.0 / ((double)nx * ny);
    for (size_t c = 0; c < (size_t)nx * ny; ++c) { Ex[c] *= s; Ey[c] *= s; }
    free(rk); free(exk); free(eyk);
}

/* darwin_b_field_kernel (include/pic/solver_darwin.hpp:38-75):
 * B_k = i mu0 (k x J)/k^2, mu0 = 1/(eps0 c^2); DC and Nyquist zeroed.
 * staggered: jx holds samples at (i+1/2, j) and jy at (i, j+1/2) (Esirkepov
 * current); a sample offset by half a cell carries the factor exp(i k d/2)
 * in its DFT, so J_k *= exp(-i k d/2) recovers the nodal spectrum. */
void orc_solve_darwin_b_f64(const orc_grid* g, double eps0, double c, const double* jx,
                            const double* jy, const double* jz, double* bx, double* by,
                            double* bz, int32_t staggered) {
    const int nx = g->nx, ny = g->ny, nkx = nx / 2 + 1;
    const size_t nk = (size_t)nkx * ny;
    double* J[3];
    double* B[3];
    for (int q = 0; q < 3; ++q) {
        J[q] = (double*)malloc(sizeof(double) * 2 * nk);
        B[q] = (double*)malloc(sizeof(double) * 2 * nk);
    }
    orc_fft2_r2c(nx, ny, jx, J[0]);
    orc_fft2_r2c(nx, ny, jy, J[1]);
    orc_fft2_r2c(nx, ny, jz, J[2]);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nkx; ++i) {
            const size_t t = (size_t)j * nkx + i;
            double kx, ky;
            kvec(g, i, j, &kx, &ky);
            const double k2 = kx * kx + ky * ky;
            if (is_nyquist(g, i, j) || k2 == 0.0) {
                for (int q = 0; q < 3; ++q) B[q][2 * t] = B[q][2 * t + 1] = 0.0;
                continue;
            }
            const double G = 1.0 / (eps0 * c * c * k2);
            double jxr = J[0][2 * t], jxi = J[0][2 * t + 1];
            double jyr = J[1][2 * t], jyi = J[1][2 * t + 1];
            if (staggered) {
                const double ax = -0.5 * kx * g->dx, ay = -0.5 * ky * g->dy;
                const double xr = jxr * cos(ax) - jxi * sin(ax);
                const double xi = jxr * sin(ax) + jxi * cos(ax);
                const double yr = jyr * cos(ay) - jyi * sin(ay);
                const double yi = jyr * sin(ay) + jyi * cos(ay);
                jxr = xr; jxi = xi; jyr = yr; jyi = yi;
            }
            const double jzr = J[2][2 * t], jzi = J[2][2 * t + 1];
            B[0][2 * t] = -G * ky * jzi;  B[0][2 * t + 1] = G * ky * jzr;
            B[1][2 * t] = G * kx * jzi;   B[1][2 * t + 1] = -G * kx * jzr;
            const double cr = kx * jyr - ky * jxr, ci = kx * jyi - ky * jxi;
            B[2][2 * t] = -G * ci;        B[2][2 * t + 1] = G * cr;
        }
    orc_fft2_c2r(nx, ny, B[0], bx);
    orc_fft2_c2r(nx, ny, B[1], by);
    orc_fft2_c2r(nx, ny, B[2], bz);
    const double s = 1.0 / ((double)nx * ny);
    for (size_t k = 0; k < (size_t)nx * ny; ++k) { bx[k] *= s; by[k] *= s; bz[k] *= s; }
    for (int q = 0; q < 3; ++q) { free(J[q]); free(B[q]); }
}

/* darwin_et_kernel (include/pic/solver_darwin.hpp:83-129): dJ/dt = dcu - i S
 * with S from amu, transverse projection, E_T = -G_et (dJ/dt)_T,
 * G_et = 1/(eps0 c^2 k^2 + n0). */
void orc_solve_darwin_et_f64(const orc_grid* g, double eps0, double c, double n0,
                             const double* dcx, const double* dcy, const double* dcz,
                             const double* a0, const double* a1, const double* a2,
                             const double* a3, double* etx, double* ety, double* etz) {
    const int nx = g->nx, ny = g->ny, nkx = nx / 2 + 1;
    const size_t nk = (size_t)nkx * ny;
    const double* srcs[7] = {dcx, dcy, dcz, a0, a1, a2, a3};
    double* K[7];
    double* E[3];
    for (int q = 0; q < 7; ++q) {
        K[q] = (double*)malloc(sizeof(double) * 2 * nk);
        orc_fft2_r2c(nx, ny, srcs[q], K[q]);
    }
    for (int q = 0; q < 3; ++q) E[q] = (double*)malloc(sizeof(double) * 2 * nk);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nkx; ++i) {
            const size_t t = (size_t)j * nkx + i;
            double kx, ky;
            kvec(g, i, j, &kx, &ky);
            const double k2 = kx * kx + ky * ky;
            if (is_nyquist(g, i, j) || k2 == 0.0) {
                for (int q = 0; q < 3; ++q) E[q][2 * t] = E[q][2 * t + 1] = 0.0;
                continue;
            }
            const double G = 1.0 / (eps0 * c * c * k2 + n0);
#define RE(q) K[q][2 * t]
#define IM(q) K[q][2 * t + 1]
            const double Sxr = kx * RE(3) + ky * RE(4), Sxi = kx * IM(3) + ky * IM(4);
            const double Syr = kx * RE(4) - ky * RE(3), Syi = kx * IM(4) - ky * IM(3);
            const double Szr = kx * RE(5) + ky * RE(6), Szi = kx * IM(5) + ky * IM(6);
            double Jxr = RE(0) + Sxi, Jxi = IM(0) - Sxr;
            double Jyr = RE(1) + Syi, Jyi = IM(1) - Syr;
            const double Jzr = RE(2) + Szi, Jzi = IM(2) - Szr;
#undef RE
#undef IM
            const double inv = 1.0 / k2;
            const double kdr = (kx * Jxr + ky * Jyr) * inv, kdi = (kx * Jxi + ky * Jyi) * inv;
            Jxr -= kx * kdr; Jxi -= kx * kdi;
            Jyr -= ky * kdr; Jyi -= ky * kdi;
            E[0][2 * t] = -G * Jxr; E[0][2 * t + 1] = -G * Jxi;
            E[1][2 * t] = -G * Jyr; E[1][2 * t + 1] = -G * Jyi;
            E[2][2 * t] = -G * Jzr; E[2][2 * t + 1] = -G * Jzi;
        }
    orc_fft2_c2r(nx, ny, E[0], etx);
    orc_fft2_c2r(nx, ny, E[1], ety);
    orc_fft2_c2r(nx, ny, E[2], etz);
    const double s = 1.0 / ((double)nx * ny);
    for (size_t k = 0; k < (size_t)nx * ny; ++k) { etx[k] *= s; ety[k] *= s; etz[k] *= s; }
    for (int q = 0; q < 7; ++q) free(K[q]);
    for (int q = 0; q < 3; ++q) free(E[q]);
}

/* ------------------------------------------------------------------------ */
/* Esirkepov                                                                 */
/* ------------------------------------------------------------------------ */

/* s1_shape4 (include/pic/yee2d.hpp:126-132): node n at ib-1+n, S = max(0,1-|x-p|). */
static inline void shape4_d(double x, int ib, double S[4]) {
    for (int k = 0; k < 4; ++k) {
        const double d = fabs(x - (double)(ib - 1 + k));
        S[k] = d < 1.0 ? 1.0 - d : 0.0;
    }
}

static inline int wrapfar(int i, int n) {
    i %= n;
    if (i < 0) i += n;
    return i;
}

/* esirkepov_scatter (include/pic/yee2d.hpp:310-355) with the union stencil
 * ib = floor(min(x0,x1)) (:383-384); Jx at (i+1/2,j), Jy at (i,j+1/2), Jz nodal. */
void orc_esirkepov_f64(const orc_grid* g, const float* x0, const float* y0,
                       const float* x1, const float* y1, const float* uz,
                       const float* w, double w_uniform, int64_t n, double qm, double dt,
                       double* jx, double* jy, double* jz) {
    for (int64_t p = 0; p < n; ++p) {
        const double qw = qm * (w ? (double)w[p] : w_uniform);
        const double X0 = x0[p], Y0 = y0[p];
        double X1 = x1[p], Y1 = y1[p];
        /* x1 wrapped by the periodic boundary: take its minimal image (the
         * reference deposits before migrate, so its x1 is never wrapped and
         * this is the identity there) */
        if (X1 - X0 > 0.5 * g->nx) X1 -= g->nx;
        else if (X0 - X1 > 0.5 * g->nx) X1 += g->nx;
        if (Y1 - Y0 > 0.5 * g->ny) Y1 -= g->ny;
        else if (Y0 - Y1 > 0.5 * g->ny) Y1 += g->ny;
        const int ib = (int)floor(fmin(X0, X1)), jb = (int)floor(fmin(Y0, Y1));
        double Sx0[4], Sx1[4], Sy0[4], Sy1[4], DSx[4], DSy[4];
        shape4_d(X0, ib, Sx0); shape4_d(X1, ib, Sx1);
        shape4_d(Y0, jb, Sy0); shape4_d(Y1, jb, Sy1);
        for (int k = 0; k < 4; ++k) { DSx[k] = Sx1[k] - Sx0[k]; DSy[k] = Sy1[k] - Sy0[k]; }
        const double cJx = -qw / (dt * g->dy), cJy = -qw / (dt * g->dx);
        const double cJz = qw * (double)uz[p] / (g->dx * g->dy);
        for (int b = 0; b < 4; ++b) {
            const double sy = Sy0[b] + 0.5 * DSy[b];
            double acc = 0.0;
            for (int a = 0; a < 4; ++a) {
                acc += DSx[a] * sy;
                jx[wrapfar(jb - 1 + b, g->ny) * g->nx + wrapfar(ib - 1 + a, g->nx)] += cJx * acc;
            }
        }
        for (int a = 0; a < 4; ++a) {
            const double sx = Sx0[a] + 0.5 * DSx[a];
            double acc = 0.0;
            for (int b = 0; b < 4; ++b) {
                acc += DSy[b] * sx;
                jy[wrapfar(jb - 1 + b, g->ny) * g->nx + wrapfar(ib - 1 + a, g->nx)] += cJy * acc;
            }
        }
        for (int b = 0; b < 4; ++b)
            for (int a = 0; a < 4; ++a) {
                const double wz = Sx0[a] * Sy0[b] + 0.5 * (DSx[a] * Sy0[b] + Sx0[a] * DSy[b]) +
                                  DSx[a] * DSy[b] / 3.0;
                jz[wrapfar(jb - 1 + b, g->ny) * g->nx + wrapfar(ib - 1 + a, g->nx)] += cJz * wz;
            }
    }
}

/* k_rho_nodes (include/pic/yee2d.hpp:491-507): CIC node charge. */
void orc_rho_nodes_f64(const orc_grid* g, const float* x, const float* y, const float* w,
                       double w_uniform, int64_t n, double qm, double* rho) {
    orc_deposit_rho_f64(g, x, y, w, w_uniform, n, qm, rho);
}

/* ------------------------------------------------------------------------ */
/* loaders                                                                   */
/* ------------------------------------------------------------------------ */

/* radical_inverse (include/pic/particles.hpp:190-194). */
double orc_radical_inverse(uint32_t i, uint32_t base) {
    const double inv = 1.0 / (double)base;
    double f = inv, r = 0.0;
    while (i > 0) {
        r += (double)(i % base) * f;
        i /= base;
        f *= inv;
    }
    return r;
}

/* hash_u32 / rng_uniform (include/pic/particles.hpp:200-210). */
uint32_t orc_hash_u32(uint32_t v) {
    v ^= v >> 16; v *= 0x7feb352dU;
    v ^= v >> 15; v *= 0x846ca68bU;
    v ^= v >> 16;
    return v;
}

double orc_rng_uniform(int64_t t, uint32_t stream, uint64_t seed) {
    const uint32_t h = orc_hash_u32((uint32_t)t * 0x9e3779b9U + stream * 0x85ebca6bU +
                                    (uint32_t)seed);
    return ((double)h + 0.5) * (1.0 / 4294967296.0);
}

/* Inverse error function: Winitzki's closed-form approximation (S. Winitzki,
 * "A handy approximation for the error function and its inverse", 2008,
 * a = 0.147) as the start, then Halley steps on erf(x) - y (erfc near |y|=1). */
double orc_erfinv(double y) {
    if (y <= -1.0) return -HUGE_VAL;
    if (y >= 1.0) return HUGE_VAL;
    if (y == 0.0) return 0.0;
    const double a = 0.147, pi = 3.14159265358979323846;
    const double ln = log((1.0 - y) * (1.0 + y));
    const double t1 = 2.0 / (pi * a) + 0.5 * ln;
    double x = sqrt(sqrt(t1 * t1 - ln / a) - t1);
    if (y < 0.0) x = -x;
    const double two_over_sqrtpi = 1.12837916709551257390;
    for (int it = 0; it < 4; ++it) {
        /* f = erf(x) - y, written through erfc where erf(x) is close to +-1 */
        double f;
        if (fabs(y) < 0.5) f = erf(x) - y;
        else if (y > 0.0) f = (1.0 - y) - erfc(x);
        else f = erfc(-x) - (1.0 + y);
        const double fp = two_over_sqrtpi * exp(-x * x);
        if (fp == 0.0) break;
        x -= f / (fp + x * f);
    }
    return x;
}

/* One quiet Maxwellian component: ufl + sqrt(2) uth erfinv(2 q - 1) with
 * q = radical_inverse(t+1, base) (include/pic/particles.hpp:270-289). */
float orc_quiet_u(int64_t t, uint32_t base, double uth, double ufl) {
    const double q = orc_radical_inverse((uint32_t)t + 1, base);
    return (float)(ufl + 1.41421356237309515 * uth * orc_erfinv(2.0 * q - 1.0));
}

/* x-displacement modes applied at load: x += sum_m A_m sin(m k1 x), k1 =
 * 2 pi/nx in cell units, A_m in cells, then the periodic wrap.  This is the
 * multi-mode form of the reference's single perturbation
 * (include/pic/particles.hpp:239-245, perturb_amp sin(perturb_kx k1 x)),
 * evaluated in double and rounded once. */
void orc_apply_modes_f32(const orc_grid* g, float* x, int64_t n, const double* amp_cells,
                         int32_t nmodes) {
    if (nmodes <= 0) return;
    const double k1 = TWO_PI / (double)g->nx;
    for (int64_t i = 0; i < n; ++i) {
        double d = 0.0;
        for (int32_t m = 0; m < nmodes; ++m) d += amp_cells[m] * sin((m + 1) * k1 * (double)x[i]);
        x[i] = wrap_coord((float)((double)x[i] + d), (float)g->nx);
    }
}

/* species_init_kernel (include/pic/particles.hpp:333-379).  The draw for
 * l = c*ppc + s (t = base + l) is the reference's; the particle is STORED at
 * slot base + s*ncell + c so that a stable tile sort leaves consecutive
 * particles of a tile in consecutive cells. */
void orc_load_species(const orc_grid* g, const orc_species_init* sp, float* x, float* y,
                      float* ux, float* uy, float* uz) {
    const int64_t ncell = (int64_t)g->nx * g->ny;
    const double r2 = 1.41421356237309515;
    for (int64_t l = 0; l < sp->count; ++l) {
        const int64_t t = sp->base + l;
        const int32_t c = (int32_t)(l / sp->ppc), s = (int32_t)(l % sp->ppc);
        const int32_t i = c % g->nx, j = c / g->nx;
        float fx, fy;
        if (sp->noisy) {
            fx = (float)orc_rng_uniform(t, 0, sp->seed);
            fy = (float)orc_rng_uniform(t, 1, sp->seed);
        } else {
            const float px = (float)sp->index * 0.6180339887498949f;
            const float py = (float)sp->index * 0.7548776662466927f;
            fx = (float)orc_radical_inverse((uint32_t)s + 1, 2) + px; fx -= floorf(fx);
            fy = (float)orc_radical_inverse((uint32_t)s + 1, 3) + py; fy -= floorf(fy);
        }
        float xv = (float)i + fx, yv = (float)j + fy;
        if (xv >= (float)g->nx) xv = nextafterf((float)g->nx, 0.0f);
        if (yv >= (float)g->ny) yv = nextafterf((float)g->ny, 0.0f);
        const double qx = sp->noisy ? orc_rng_uniform(t, 2, sp->seed)
                                    : orc_radical_inverse((uint32_t)t + 1, 5);
        const double qy = sp->noisy ? orc_rng_uniform(t, 3, sp->seed)
                                    : orc_radical_inverse((uint32_t)t + 1, 7);
        const double qz = sp->noisy ? orc_rng_uniform(t, 4, sp->seed)
                                    : orc_radical_inverse((uint32_t)t + 1, 11);
        const int64_t slot = sp->base + (int64_t)s * ncell + c;
        x[slot] = xv;
        y[slot] = yv;
        ux[slot] = (float)(sp->ufl[0] + r2 * sp->uth[0] * orc_erfinv(2.0 * qx - 1.0));
        uy[slot] = (float)(sp->ufl[1] + r2 * sp->uth[1] * orc_erfinv(2.0 * qy - 1.0));
        uz[slot] = (float)(sp->ufl[2] + r2 * sp->uth[2] * orc_erfinv(2.0 * qz - 1.0));
    }
}

/* ------------------------------------------------------------------------ */
/* diagnostics                                                               */
/* ------------------------------------------------------------------------ */

/* f(x, v_par/v_th) counts.  fp64 = 0: phase_hist_kernel
 * (include/pic/diag/module.hpp:58-70) -- FP32, x wrapped in cells,
 * vp = (ux*bx + uz*bz)/vth, ix = (int)(x/nx*nxb); fp64 = 1: Snapshot's host
 * histogram (include/pic/diag/snapshot.hpp:55-62) -- FP64, x*dx wrapped by
 * Lx, ix = (int)(xx/Lx*nxb).  iv = (int)((vp - vlo)/(vhi - vlo)*nvb); the
 * casts truncate toward zero and out-of-window bins are skipped.
 * counts[iv*nxb + ix] += 1. */
void orc_phase_hist(const orc_grid* g, const float* x, const float* ux, const float* uz,
                    int64_t n, int32_t nxb, int32_t nvb, double bx, double bz, double vth,
                    double vlo, double vhi, int32_t fp64, uint32_t* counts) {
    const double Lx = g->nx * g->dx;
    for (int64_t i = 0; i < n; ++i) {
        int ix, iv;
        if (fp64) {
            double xx = (double)x[i] * g->dx;
            if (xx < 0.0) xx += Lx;
            else if (xx >= Lx) xx -= Lx;
            const double vp = ((double)ux[i] * bx + (double)uz[i] * bz) / vth;
            ix = (int)(xx / Lx * (double)nxb);
            iv = (int)((vp - vlo) / (vhi - vlo) * (double)nvb);
        } else {
            const float fnx = (float)g->nx;
            float xx = x[i];
            if (xx < 0.f) xx += fnx;
            else if (xx >= fnx) xx -= fnx;
            const float vp = (ux[i] * (float)bx + uz[i] * (float)bz) / (float)vth;
            ix = (int)(xx / fnx * (float)nxb);
            iv = (int)((vp - (float)vlo) / ((float)vhi - (float)vlo) * (float)nvb);
        }
        if (ix >= 0 && ix < nxb && iv >= 0 && iv < nvb) counts[(int64_t)iv * nxb + ix] += 1u;
    }
}


/* ------------------------------------------------------------------------ */
/* full-Maxwell (Yee) branch: FP32 restatement of the mesh updates and the   */
/* push, FP64 reductions                                                     */
/* ------------------------------------------------------------------------ */

/* F holds nine nx*ny grids: ex ey ez bx by bz jx jy jz.  Staggering
 * include/pic/yee2d.hpp:7-9; periodic one-period wrap. */
static inline int yat(const orc_grid* g, int i, int j) {
    if (i < 0) i += g->nx; else if (i >= g->nx) i -= g->nx;
    if (j < 0) j += g->ny; else if (j >= g->ny) j -= g->ny;
    return j * g->nx + i;
}

/* k_faraday (include/pic/yee2d.hpp:70-81): B += h (-curl E); every update
 * reads the E of the previous state, so the order of cells is free. */
void orc_yee_faraday_f32(const orc_grid* g, float* F, float h) {
    const int64_t nc = (int64_t)g->nx * g->ny;
    const float* ex = F; const float* ey = F + nc; const float* ez = F + 2 * nc;
    float* bx = F + 3 * nc; float* by = F + 4 * nc; float* bz = F + 5 * nc;
    const float dx = (float)g->dx, dy = (float)g->dy;
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) {
            const int c = j * g->nx + i;
            const float e = ez[c];
            const float dez_y = (ez[yat(g, i, j + 1)] - e) / dy;
            const float dez_x = (ez[yat(g, i + 1, j)] - e) / dx;
            const float dey_x = (ey[yat(g, i + 1, j)] - ey[c]) / dx;
            const float dex_y = (ex[yat(g, i, j + 1)] - ex[c]) / dy;
            bx[c] = bx[c] - h * dez_y;
            by[c] = by[c] + h * dez_x;
            bz[c] = bz[c] - h * (dey_x - dex_y);
        }
}

/* k_ampere (include/pic/yee2d.hpp:84-97): E += dt (c^2 curl B - J). */
void orc_yee_ampere_f32(const orc_grid* g, float* F, float c2, float dt) {
    const int64_t nc = (int64_t)g->nx * g->ny;
    float* ex = F; float* ey = F + nc; float* ez = F + 2 * nc;
    const float* bx = F + 3 * nc; const float* by = F + 4 * nc; const float* bz = F + 5 * nc;
    const float* jx = F + 6 * nc; const float* jy = F + 7 * nc; const float* jz = F + 8 * nc;
    const float dx = (float)g->dx, dy = (float)g->dy;
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) {
            const int c = j * g->nx + i;
            const float b = bz[c];
            const float dbz_y = (b - bz[yat(g, i, j - 1)]) / dy;
            const float dbz_x = (b - bz[yat(g, i - 1, j)]) / dx;
            const float dby_x = (by[c] - by[yat(g, i - 1, j)]) / dx;
            const float dbx_y = (bx[c] - bx[yat(g, i, j - 1)]) / dy;
            ex[c] = ex[c] + dt * (c2 * dbz_y - jx[c]);
            ey[c] = ey[c] + dt * (-c2 * dbz_x - jy[c]);
            ez[c] = ez[c] + dt * (c2 * (dby_x - dbx_y) - jz[c]);
        }
}

/* k_antenna (include/pic/yee2d.hpp:582-607): the rotating transverse current
 * column g(i) (cos, sin)(w0 t), g Gaussian in x around x0 (cut at 4 sigma)
 * with a trapezoidal envelope, added to Jy, Jz.  Double arithmetic as the
 * kernel; exp, cos and sin are the host's. */
void orc_yee_antenna(const orc_grid* g, double amp, double x0, double sigma, double w0,
                     double trmp, double toff, double t, float* F) {
    const int64_t nc = (int64_t)g->nx * g->ny;
    float* jy = F + 7 * nc; float* jz = F + 8 * nc;
    if (toff > 0 && t >= toff) return;
    double ramp = 1.0;
    if (trmp > 0) {
        if (t < trmp) ramp = t / trmp;
        else if (toff > 0 && t > toff - trmp) ramp = (toff - t) / trmp;
    }
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) {
            const double d = i - x0;
            if (fabs(d) > 4.0 * sigma) continue;
            const double a = amp * ramp * exp(-d * d / (2.0 * sigma * sigma));
            const int c = j * g->nx + i;
            jy[c] += (float)(a * cos(w0 * t));
            jz[c] += (float)(a * sin(w0 * t));
        }
}

/* k_cold_fluid (include/pic/yee2d.hpp:536-560): the linearised cold electron
 * fluid on the x nodes (ny = 1): half E kick, rotation by -qm B0(x) dt, half
 * E kick, then J += qm nc v.  FP32 as the kernel; sinf and cosf are the
 * host's (the kernel's sincosf is the device's). */
void orc_yee_cold_fluid_f32(const orc_grid* g, const orc_yee_params* p, double cold_nc,
                            float* F, float* vcy, float* vcz) {
    const int64_t nc = (int64_t)g->nx * g->ny;
    const float* ey = F + nc; const float* ez = F + 2 * nc;
    float* jy = F + 7 * nc; float* jz = F + 8 * nc;
    const float dx = (float)g->dx;
    const float h = (float)(0.5 * p->dt * p->qm);
    const float jc = (float)(p->qm * cold_nc);
    for (int64_t c = 0; c < nc; ++c) {
        const int i = (int)(c % g->nx);
        const float Ey = ey[c], Ez = ez[c];
        float uy = vcy[c] + h * Ey;
        float uz = vcz[c] + h * Ez;
        const float b0 = p->b0_prof ? orc_b0_profile(p, i * dx) : p->B0[0];
        const float th = (float)(-p->qm * p->dt) * b0;
        const float sn = sinf(th), cs = cosf(th);
        const float uy2 = uy * cs - uz * sn;
        const float uz2 = uy * sn + uz * cs;
        uy = uy2 + h * Ey;
        uz = uz2 + h * Ez;
        vcy[c] = uy;
        vcz[c] = uz;
        jy[c] += jc * uy;
        jz[c] += jc * uz;
    }
}

/* The x damping layers: masks exp(-numax d^2 dt), d the depth into the
 * nd-cell layer over nd, at integer x (Ey, Ez, Bx, the cold fluid) and
 * half-integer x (Ex, By, Bz) (include/pic/simulation_maxwell.hpp:189-203),
 * applied as k_damp_x (include/pic/yee2d.hpp:514-526).  vcy may be NULL. */
static float damp_mask(const orc_grid* g, int nd, double numax, double dt, double xi) {
    double d = 0.0;
    if (xi < nd) d = (nd - xi) / (double)nd;
    else if (xi > g->nx - nd) d = (xi - (g->nx - nd)) / (double)nd;
    if (d <= 0.0) return 1.0f;
    return (float)exp(-numax * d * d * dt);
}

void orc_yee_damp_f32(const orc_grid* g, int32_t nd, double numax, double dt, float* F,
                      float* vcy, float* vcz) {
    const int64_t nc = (int64_t)g->nx * g->ny;
    for (int64_t c = 0; c < nc; ++c) {
        const int i = (int)(c % g->nx);
        const float a = damp_mask(g, nd, numax, dt, i), b = damp_mask(g, nd, numax, dt, i + 0.5);
        if (a == 1.f && b == 1.f) continue;
        F[nc + c] *= a;      /* ey */
        F[2 * nc + c] *= a;  /* ez */
        F[3 * nc + c] *= a;  /* bx */
        F[c] *= b;           /* ex */
        F[4 * nc + c] *= b;  /* by */
        F[5 * nc + c] *= b;  /* bz */
        if (vcy) {
            vcy[c] *= a;
            vcz[c] *= a;
        }
    }
}

/* one [1,2,1]x[1,2,1]/16 pass (include/pic/yee2d.hpp:494-508) */
void orc_yee_binomial_f32(const orc_grid* g, const float* src, float* dst) {
    const float w[3] = {0.25f, 0.5f, 0.25f};
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) {
            float acc = 0.f;
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di)
                    acc = fmaf(w[di + 1] * w[dj + 1], src[yat(g, i + di, j + dj)], acc);
            dst[j * g->nx + i] = acc;
        }
}

/* CIC interpolation of samples at (i + ox, j + oy) (gather_stag,
 * include/pic/yee2d.hpp:102-112), as a fixed fma chain */
static inline float ygather(const orc_grid* g, const float* a, float x, float y, float ox,
                            float oy) {
    const float xs = x - ox, ys = y - oy;
    const float fi = floorf(xs), fj = floorf(ys);
    const int i0 = wrapfar((int)fi, g->nx), j0 = wrapfar((int)fj, g->ny);
    const int i1 = i0 + 1 == g->nx ? 0 : i0 + 1, j1 = j0 + 1 == g->ny ? 0 : j0 + 1;
    const float fx = xs - fi, fy = ys - fj, gx = 1.f - fx, gy = 1.f - fy;
    const float v00 = a[j0 * g->nx + i0], v10 = a[j0 * g->nx + i1];
    const float v01 = a[j1 * g->nx + i0], v11 = a[j1 * g->nx + i1];
    return fmaf(fx * fy, v11, fmaf(gx * fy, v01, fmaf(fx * gy, v10, (gx * gy) * v00)));
}

static inline void yrotate(float* u, float bx, float by, float bz, float h) {
    const float tx = h * bx, ty = h * by, tz = h * bz;
    const float t2 = fmaf(tz, tz, fmaf(ty, ty, tx * tx));
    const float sf = 2.0f / (1.0f + t2);
    const float sx = sf * tx, sy = sf * ty, sz = sf * tz;
    const float upx = fmaf(u[1], tz, fmaf(-u[2], ty, u[0]));
    const float upy = fmaf(u[2], tx, fmaf(-u[0], tz, u[1]));
    const float upz = fmaf(u[0], ty, fmaf(-u[1], tx, u[2]));
    u[0] = fmaf(upy, sz, fmaf(-upz, sy, u[0]));
    u[1] = fmaf(upz, sx, fmaf(-upx, sz, u[1]));
    u[2] = fmaf(upx, sy, fmaf(-upy, sx, u[2]));
}

/* background_b0.hpp:28-55 */
float orc_b0_profile(const orc_yee_params* p, float xph) {
    const float d = xph - (float)p->b0_xc;
    if (p->b0_prof == 2) {
        const float sn = d / (float)p->b0_sref, u = sn * sn;
        float q = (float)p->b0_c[5];
        for (int k = 4; k >= 0; --k) q = q * u + (float)p->b0_c[k];
        return p->B0[0] * q;
    }
    return p->B0[0] * (1.f + (float)p->b0_a * d * d);
}

static float b0_profile_dx(const orc_yee_params* p, float xph) {
    const float d = xph - (float)p->b0_xc;
    if (p->b0_prof == 2) {
        const float sn = d / (float)p->b0_sref, u = sn * sn;
        float dp = 5.f * (float)p->b0_c[5];
        for (int k = 4; k >= 1; --k) dp = dp * u + (float)k * (float)p->b0_c[k];
        return p->B0[0] * dp * 2.f * d / (float)(p->b0_sref * p->b0_sref);
    }
    return p->B0[0] * 2.f * (float)p->b0_a * d;
}

static double yramp(double t, double trmp, double toff) {
    if (t >= toff) return 0.0;
    if (t < trmp) return t / trmp;
    if (t > toff - trmp) return (toff - t) / trmp;
    return 1.0;
}

/* yee_advance_particle (include/pic/yee2d.hpp:145-302) in the library's FP32
 * operation order: staggered gather of E (+ pump) and the wave B (+ B0 or the
 * mirror field), kick, rotate with qmh / gamma_mid (rel), kick, delta-f weight
 * update at the time-centred u, move (unwrapped), x walls.  Writes u, x, y,
 * wd (wc, wdd), and x0, y0 (old positions), vz (uz/gamma, when non-NULL) and
 * wdep (w wd, when non-NULL). */
void orc_yee_push_f32(const orc_grid* g, const orc_yee_params* p, double tnow, const float* F,
                      float* x, float* y, float* ux, float* uy, float* uz, const float* w,
                      float w_uniform, float* wd, float* wc, double* wdd, float* x0o,
                      float* y0o, float* vzo, float* wdep, int64_t n) {
    const int64_t nc = (int64_t)g->nx * g->ny;
    const float *ex = F, *ey = F + nc, *ez = F + 2 * nc;
    const float *bx = F + 3 * nc, *by = F + 4 * nc, *bz = F + 5 * nc;
    const float dx = (float)g->dx, dy = (float)g->dy, fnx = (float)g->nx;
    for (int64_t t = 0; t < n; ++t) {
        const float x0 = x[t], y0 = y[t];
        float Ex = ygather(g, ex, x0, y0, 0.5f, 0.f);
        float Ey = ygather(g, ey, x0, y0, 0.f, 0.5f);
        float Ez = ygather(g, ez, x0, y0, 0.f, 0.f);
        if (p->pump && tnow < p->pump_toff) {
            const double th = p->pump_k0 * (double)x0 * (double)dx - p->pump_w0 * tnow;
            const double r = yramp(tnow, p->pump_trmp, p->pump_toff);
            Ex += (float)(p->pump_ex * cos(th) * r);
            Ey += (float)(-p->pump_ey * sin(th) * r);
            Ez += (float)(p->pump_ez * cos(th) * r);
        }
        const float dBx = ygather(g, bx, x0, y0, 0.f, 0.5f);
        const float dBy = ygather(g, by, x0, y0, 0.5f, 0.f);
        const float dBz = ygather(g, bz, x0, y0, 0.5f, 0.5f);
        float u[3] = {ux[t], uy[t], uz[t]};
        const float uo[3] = {u[0], u[1], u[2]};
        const float qmh = (float)(p->qm * 0.5 * p->dt);
        u[0] = fmaf(qmh, Ex, u[0]);
        u[1] = fmaf(qmh, Ey, u[1]);
        u[2] = fmaf(qmh, Ez, u[2]);
        const float gmid = p->rel ? 1.f / sqrtf(1.f + fmaf(u[2], u[2], fmaf(u[1], u[1], u[0] * u[0])))
                                  : 1.f;
        if (p->b0_prof) {
            const float xph = x0 * dx;
            const float b0 = orc_b0_profile(p, xph);
            const float mc = b0_profile_dx(p, xph) / (2.f * b0 * (float)p->qm);
            yrotate(u, dBx + b0, dBy + mc * u[2], dBz - mc * u[1], qmh * gmid);
        } else {
            yrotate(u, dBx + p->B0[0], dBy + p->B0[1], dBz + p->B0[2], qmh * gmid);
        }
        u[0] = fmaf(qmh, Ex, u[0]);
        u[1] = fmaf(qmh, Ey, u[1]);
        u[2] = fmaf(qmh, Ez, u[2]);
        const float gnew = p->rel ? 1.f / sqrtf(1.f + fmaf(u[2], u[2], fmaf(u[1], u[1], u[0] * u[0])))
                                  : 1.f;
        if (p->deltaf) {
            const float uxc = 0.5f * (u[0] + uo[0]), uyc = 0.5f * (u[1] + uo[1]),
                        uzc = 0.5f * (u[2] + uo[2]);
            const float gc = p->rel ? 1.f / sqrtf(1.f + fmaf(uzc, uzc, fmaf(uyc, uyc, uxc * uxc)))
                                    : 1.f;
            const float Fx = Ex + gc * (uyc * dBz - uzc * dBy);
            const float Fy = Ey + gc * (uzc * dBx - uxc * dBz);
            const float Fz = Ez + gc * (uxc * dBy - uyc * dBx);
            float tperp = (float)p->df_tperp, t2 = (float)(p->df_kappa * p->df_tperp);
            if (p->b0_prof) {
                const float b = orc_b0_profile(p, x0 * dx) / p->B0[0];
                const float par = (1.f - 1.f / b) / (float)p->df_tpar;
                tperp = 1.f / (par + (1.f / b) / (float)p->df_tperp);
                t2 = 1.f / (par + (1.f / b) / (float)(p->df_kappa * p->df_tperp));
            }
            float perp = 1.f / tperp;
            if (p->df_dist == 1) {
                const float u2p = fmaf(uzc, uzc, uyc * uyc);
                const float gg = (float)p->df_rho * expf(-0.5f * u2p * (1.f / t2 - 1.f / tperp));
                const float omg = fmaxf(1.f - gg, 1e-3f);
                perp = fmaxf((1.f / tperp - gg / t2) / omg, -10.f / tperp);
            }
            const float S = Fx * uxc / (float)p->df_tpar + (Fy * uyc + Fz * uzc) * perp;
            const float relax = p->df_taud > 0.0 ? (float)(1.0 - p->dt / p->df_taud) : 1.f;
            const float hi = 1.f - 1e-4f, lo = -8.f;
            float wn;
            if (p->df_wprec == 2) {
                const double wo = wdd[t];
                double v = (wo + p->dt * p->qm * (1.0 - wo) * (double)S) * (double)relax;
                v = fmin(fmax(v, (double)lo), (double)hi);
                wdd[t] = v;
                wn = (float)v;
            } else if (p->df_wprec == 1) {
                const float wo = wd[t];
                const float inc = (float)(p->dt * p->qm) * (1.f - wo) * S;
                const float yk = inc - wc[t];
                const float sum = wo + yk;
                wc[t] = (sum - wo) - yk;
                wn = fminf(fmaxf(sum * relax, lo), hi);
            } else {
                const float wo = wd[t];
                wn = fminf(fmaxf((wo + (float)(p->dt * p->qm) * (1.f - wo) * S) * relax, lo), hi);
            }
            wd[t] = wn;
            if (wdep) wdep[t] = (w ? w[t] : w_uniform) * wn;
        }
        float x1 = x0 + (float)((double)(u[0] * gnew) * p->dt / (double)dx);
        const float y1 = y0 + (float)((double)(u[1] * gnew) * p->dt / (double)dy);
        float us[3] = {u[0], u[1], u[2]};
        if (p->bnd_x) {
            if (x1 < 0.f) { x1 = -x1; us[0] = -u[0]; }
            else if (x1 >= fnx) { x1 = 2.f * fnx - x1; us[0] = -u[0]; }
            if (x1 >= fnx) x1 = nextafterf(fnx, 0.f);
            if (p->bnd_x == 2) {
                const float nd = (float)p->bnd_nd;
                float d = 0.f;
                if (x1 < nd) d = (nd - x1) / nd;
                else if (x1 > fnx - nd) d = (x1 - (fnx - nd)) / nd;
                if (d > 0.f) {
                    const float f = expf((float)(-p->bnd_numax * p->dt) * d * d);
                    us[1] = u[1] * f;
                    us[2] = u[2] * f;
                }
            }
        }
        ux[t] = us[0]; uy[t] = us[1]; uz[t] = us[2];
        if (vzo) vzo[t] = u[2] * gnew;
        x0o[t] = x0; y0o[t] = y0;
        x[t] = x1; y[t] = y1;
    }
}

/* k_energy (include/pic/yee2d.hpp:648-665): {eps0/2 |E|^2 dA, c^2/2 |B|^2 dA}
 * in FP64 over the FP32 mesh values; dA and c^2 are the mesh's FP32 copies
 * (YeeViews dxp, dyp, c2; yee2d.hpp:676-677) widened to FP64. */
void orc_yee_energy_f64(const orc_grid* g, const float* F, double c2, double* out2) {
    const int64_t nc = (int64_t)g->nx * g->ny;
    const double dA = (double)(float)g->dx * (double)(float)g->dy;
    const double c2f = (double)(float)c2;
    double we = 0.0, wb = 0.0;
    for (int64_t c = 0; c < nc; ++c) {
        double e = 0.0, b = 0.0;
        for (int q = 0; q < 3; ++q) e += (double)F[q * nc + c] * F[q * nc + c];
        for (int q = 3; q < 6; ++q) b += (double)F[q * nc + c] * F[q * nc + c];
        we += 0.5 * e * dA;
        wb += 0.5 * c2f * b * dA;
    }
    out2[0] = we;
    out2[1] = wb;
}

/* rms div B on the Yee mesh (k_div_stats, include/pic/yee2d.hpp:608-626):
 * the cell divergence in FP32 with the mesh's FP32 cell sizes, squared and
 * summed in FP64 */
double orc_yee_divb_rms(const orc_grid* g, const float* F) {
    const int64_t nc = (int64_t)g->nx * g->ny;
    const float* bx = F + 3 * nc; const float* by = F + 4 * nc;
    const float dxf = (float)g->dx, dyf = (float)g->dy;
    double s = 0.0;
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) {
            const int c = j * g->nx + i;
            const float d = (bx[yat(g, i + 1, j)] - bx[c]) / dxf +
                            (by[yat(g, i, j + 1)] - by[c]) / dyf;
            s += (double)d * d;
        }
    return sqrt(s / (double)nc);
}

/* div E at the nodes (include/pic/yee2d.hpp:614-623), FP64 */
void orc_yee_div_e(const orc_grid* g, const float* F, double* out) {
    const int64_t nc = (int64_t)g->nx * g->ny;
    const float* ex = F; const float* ey = F + nc;
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) {
            const int c = j * g->nx + i;
            out[c] = ((double)ex[c] - ex[yat(g, i - 1, j)]) / g->dx +
                     ((double)ey[c] - ey[yat(g, i, j - 1)]) / g->dy;
        }
}

/* marker offsets of the mirror load (include/pic/particles.hpp:585-597) */
int64_t orc_mirror_offsets(const orc_grid* g, const orc_yee_params* p, int32_t ppc,
                           double uth_par, double uth_perp, int32_t dist, double lc_rho,
                           double lc_kappa, int32_t* off) {
    const double tpa = uth_par * uth_par, tpe = uth_perp * uth_perp;
    int64_t tot = 0;
    for (int c = 0; c < g->nx; ++c) {
        const double b = (double)orc_b0_profile(p, (c + 0.5f) * (float)g->dx) / (double)p->B0[0];
        const double t1 = 1.0 / ((1.0 - 1.0 / b) / tpa + (1.0 / b) / tpe);
        double nfac = t1 / tpe;
        if (dist == 1) {
            const double t2 = 1.0 / ((1.0 - 1.0 / b) / tpa + (1.0 / b) / (lc_kappa * tpe));
            nfac = (t1 - lc_rho * t2) / (tpe * (1.0 - lc_rho * lc_kappa));
        }
        off[c] = (int32_t)tot;
        tot += lround(ppc * nfac);
    }
    off[g->nx] = (int32_t)tot;
    return tot;
}

/* Diagnostics::compute (include/pic/diagnostics.hpp:58-108):
 * out5 = {ke, ee, total, charge, max|E|}. */
void orc_energies(const orc_grid* g, const float* ux, const float* uy, const float* uz,
                  const float* w, double w_uniform, int64_t n, const float* Ex,
                  const float* Ey, const float* rho, double eps0, double* out5) {
    double ke = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double u2 = (double)ux[i] * ux[i] + (double)uy[i] * uy[i] + (double)uz[i] * uz[i];
        ke += (w ? (double)w[i] : w_uniform) * u2;
    }
    ke *= 0.5;
    double se = 0.0, sr = 0.0, me = 0.0;
    const int64_t nc = (int64_t)g->nx * g->ny;
    for (int64_t c = 0; c < nc; ++c) {
        const double e2 = (double)Ex[c] * Ex[c] + (double)Ey[c] * Ey[c];
        se += e2;
        if (e2 > me) me = e2;
        if (rho) sr += rho[c];
    }
    out5[0] = ke;
    out5[1] = 0.5 * eps0 * se * g->dx * g->dy;
    out5[2] = out5[0] + out5[1];
    out5[3] = sr * g->dx * g->dy;
    out5[4] = sqrt(me);
}

/* ------------------------------------------------------------------------ */
/* whole ES step on the host (CPU baseline)                                  */
/* ------------------------------------------------------------------------ */

int32_t orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* One ES step in the order of Simulation::step (include/pic/simulation.hpp:
 * 80-93): rho (already deposited for x^n) -> E -> push+move+wrap -> deposit
 * rho(x^{n+1}).  FP32 particle state, same operation sequence as the HIP
 * path; per-thread private fixed-point grids reduced in thread order. */
int32_t orc_es_step_f32(const orc_grid* g, const orc_push_params* pp, double eps0,
                        float* x, float* y, float* ux, float* uy, float* uz, int64_t n,
                        double w_uniform, int32_t scale_log2, float* rho, float* Ex,
                        float* Ey, int32_t nthreads) {
    const int64_t nc = (int64_t)g->nx * g->ny;
    double* rd = (double*)malloc(sizeof(double) * (size_t)nc);
    double* exd = (double*)malloc(sizeof(double) * (size_t)nc);
    double* eyd = (double*)malloc(sizeof(double) * (size_t)nc);
    if (nc <= 0 || !rd || !exd || !eyd) {
        free(rd); free(exd); free(eyd);
        return -1;
    }
    for (int64_t c = 0; c < nc; ++c) rd[c] = rho[c];
    orc_solve_es_f64(g, eps0, rd, exd, eyd);
    for (int64_t c = 0; c < nc; ++c) { Ex[c] = (float)exd[c]; Ey[c] = (float)eyd[c]; }
    free(rd); free(exd); free(eyd);
    if (nthreads < 1) nthreads = 1;
    int64_t* acc = (int64_t*)calloc((size_t)nc * (size_t)nthreads, sizeof(int64_t));
    const boris_consts b = boris_make(pp->qm, pp->dt, pp->B0, pp->has_b0);
    const float cx = (float)(pp->dt / g->dx), cy = (float)(pp->dt / g->dy);
    const float fnx = (float)g->nx, fny = (float)g->ny;
    const double coef = pp->qm * w_uniform / (g->dx * g->dy);
    const double scale = ldexp(1.0, scale_log2);
#pragma omp parallel num_threads(nthreads)
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
#else
        const int tid = 0;
#endif
        int64_t* my = acc + (size_t)tid * (size_t)nc;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const cic_f s = cic_float(g, x[i], y[i]);
            const float ex = gather_f(Ex, &s), ey = gather_f(Ey, &s);
            boris_es_f32(&b, pp->has_b0, ex, ey, &ux[i], &uy[i], &uz[i]);
            const float xn = wrap_coord(fmaf(ux[i], cx, x[i]), fnx);
            const float yn = wrap_coord(fmaf(uy[i], cy, y[i]), fny);
            x[i] = xn;
            y[i] = yn;
            const cic_d d = cic_double(g, xn, yn);
            for (int k = 0; k < 4; ++k) my[d.c[k]] += (int64_t)rint(coef * d.w[k] * scale);
        }
    }
    const double inv = ldexp(1.0, -scale_log2);
    for (int64_t c = 0; c < nc; ++c) {
        int64_t s = 0;
        for (int t = 0; t < nthreads; ++t) s += acc[(size_t)t * nc + c];
        rho[c] = (float)((double)s * inv);
    }
    free(acc);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* electrostatic forward in FP64 and its discrete adjoint (no B0)            */
/* ------------------------------------------------------------------------ */

/* CIC weights and their x / y derivatives (cell units) at (x, y). */
typedef struct {
    int32_t c[4];
    double w[4], dwx[4], dwy[4];
} cic_grad;

static inline cic_grad cic_grad_d(const orc_grid* g, double x, double y) {
    cic_grad s;
    const int32_t i0 = (int32_t)floor(x), j0 = (int32_t)floor(y);
    const double fx = x - i0, fy = y - j0, gx = 1.0 - fx, gy = 1.0 - fy;
    const int32_t a = wrapi(i0, g->nx), b = wrapi(i0 + 1, g->nx);
    const int32_t c = wrapi(j0, g->ny), d = wrapi(j0 + 1, g->ny);
    s.c[0] = c * g->nx + a; s.w[0] = gx * gy; s.dwx[0] = -gy; s.dwy[0] = -gx;
    s.c[1] = c * g->nx + b; s.w[1] = fx * gy; s.dwx[1] = gy;  s.dwy[1] = -fx;
    s.c[2] = d * g->nx + a; s.w[2] = gx * fy; s.dwx[2] = -fy; s.dwy[2] = gx;
    s.c[3] = d * g->nx + b; s.w[3] = fx * fy; s.dwx[3] = fy;  s.dwy[3] = fx;
    return s;
}

/* grid[c] += sum_p v_p W_c(x_p) over particles, FP64, per-thread grids. */
static void scatter_dd(const orc_grid* g, const double* x, const double* y, const double* v,
                       double v_const, int64_t n, double* grid, int nthreads) {
    const int64_t nc = (int64_t)g->nx * g->ny;
    double* priv = (double*)calloc((size_t)nc * (size_t)nthreads, sizeof(double));
#pragma omp parallel num_threads(nthreads)
    {
#ifdef _OPENMP
        double* my = priv + (size_t)omp_get_thread_num() * (size_t)nc;
#else
        double* my = priv;
#endif
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n; ++p) {
            const cic_grad s = cic_grad_d(g, x[p], y[p]);
            const double vp = v ? v[p] : v_const;
            for (int k = 0; k < 4; ++k) my[s.c[k]] += vp * s.w[k];
        }
    }
    for (int t = 0; t < nthreads; ++t)
        for (int64_t c = 0; c < nc; ++c) grid[c] += priv[(size_t)t * nc + c];
    free(priv);
}

/* Adjoint of the ES solve: lrho = S^T (lEx, lEy), S^T = IFFT(conj(H) FFT(.)),
 * H = -i k /(eps0 k^2) with DC and Nyquist zero (the transpose of
 * orc_solve_es_f64 under the real inner product). */
void orc_solve_es_adj_f64(const orc_grid* g, double eps0, const double* lEx,
                          const double* lEy, double* lrho) {
    const int nx = g->nx, ny = g->ny, nkx = nx / 2 + 1;
    const size_t nk = (size_t)nkx * ny;
    double* ax = (double*)malloc(sizeof(double) * 2 * nk);
    double* ay = (double*)malloc(sizeof(double) * 2 * nk);
    double* r = (double*)malloc(sizeof(double) * 2 * nk);
    orc_fft2_r2c(nx, ny, lEx, ax);
    orc_fft2_r2c(nx, ny, lEy, ay);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nkx; ++i) {
            const size_t t = (size_t)j * nkx + i;
            double kx, ky;
            kvec(g, i, j, &kx, &ky);
            const double k2 = kx * kx + ky * ky;
            const double gr = k2 > 0.0 ? 1.0 / (eps0 * k2) : 0.0;
            if (is_nyquist(g, i, j)) {
                r[2 * t] = r[2 * t + 1] = 0.0;
            } else {  /* i G (kx ax + ky ay) */
                r[2 * t] = -gr * (kx * ax[2 * t + 1] + ky * ay[2 * t + 1]);
                r[2 * t + 1] = gr * (kx * ax[2 * t] + ky * ay[2 * t]);
            }
        }
    orc_fft2_c2r(nx, ny, r, lrho);
    const double s = 1.0 / ((double)nx * ny);
    for (size_t c = 0; c < (size_t)nx * ny; ++c) lrho[c] *= s;
    free(ax); free(ay); free(r);
}

/* The ES run of Simulation<Cfg> without B0, in FP64: rho^0 = D(x^0),
 * E^0 = S rho^0, u^{-1/2} = u^0 - (qm dt/2) e(x^0); then per step
 * u += qm dt e(x^n, E^n), x = wrap(x + u dt/dx), rho = D(x), E = S rho;
 * loss = EE(N) = eps0/2 sum |E^N|^2 dx dy.  With gx != NULL also the
 * discrete adjoint: gx, gy = dL/dx^0, dL/dy^0 (cell units), gux, guy =
 * dL/du^0.  States x^n and fields E^n are kept for the reverse sweep. */
/* Boris step for B0 along z in FP64: u' = R(u + h e) + h e, R the rotation
 * with t = h B0z, s = 2t/(1 + t^2): R = [[1 - t s, s], [-s, 1 - t s]]
 * (include/pic/pusher.hpp:35-80 with only the z component of B0). */
typedef struct { double h, c, s; } rot_z;

static rot_z rot_z_make(double h, double b0z) {
    const double t = h * b0z, sf = 2.0 * t / (1.0 + t * t);
    rot_z r = {h, 1.0 - t * sf, sf};
    return r;
}

static inline void boris_z_d(double* ux, double* uy, double e0, double e1, const rot_z* r) {
    const double vx = *ux + r->h * e0, vy = *uy + r->h * e1;
    *ux = r->c * vx + r->s * vy + r->h * e0;
    *uy = -r->s * vx + r->c * vy + r->h * e1;
}

/* its transpose: le += h (I + R^T) lu, lu = R^T lu */
static inline void boris_z_adj_d(double* lx, double* ly, double* le0, double* le1,
                                 const rot_z* r) {
    const double tx = r->c * *lx - r->s * *ly, ty = r->s * *lx + r->c * *ly;
    *le0 += r->h * (*lx + tx);
    *le1 += r->h * (*ly + ty);
    *lx = tx;
    *ly = ty;
}

double orc_es_adjoint_f64(const orc_grid* g, double qm, double dt, double eps0, double b0z,
                          double w_uniform, const double* x0, const double* y0,
                          const double* ux0, const double* uy0, int64_t n, int32_t nsteps,
                          double* gx, double* gy, double* gux, double* guy, int32_t nthreads) {
    if (nthreads < 1) nthreads = 1;
    const int64_t nc = (int64_t)g->nx * g->ny;
    const double area = g->dx * g->dy, coef = qm * w_uniform / area;
    const double a = qm * dt, cx = dt / g->dx, cy = dt / g->dy;
    (void)a;
    const rot_z R1 = rot_z_make(qm * 0.5 * dt, b0z);          /* the step */
    const rot_z Rh = rot_z_make(qm * 0.5 * (-0.5 * dt), b0z);  /* the half step back */
    const int adj = gx != NULL;
    const int64_t keep = adj ? (int64_t)nsteps + 1 : 1;
    double* X = (double*)malloc(sizeof(double) * (size_t)(n * keep));
    double* Y = (double*)malloc(sizeof(double) * (size_t)(n * keep));
    double* EX = (double*)malloc(sizeof(double) * (size_t)(nc * keep));
    double* EY = (double*)malloc(sizeof(double) * (size_t)(nc * keep));
    double* ux = (double*)malloc(sizeof(double) * (size_t)n);
    double* uy = (double*)malloc(sizeof(double) * (size_t)n);
    double* rho = (double*)malloc(sizeof(double) * (size_t)nc);
    memcpy(X, x0, sizeof(double) * (size_t)n);
    memcpy(Y, y0, sizeof(double) * (size_t)n);
    memcpy(ux, ux0, sizeof(double) * (size_t)n);
    memcpy(uy, uy0, sizeof(double) * (size_t)n);
    const double fnx = (double)g->nx, fny = (double)g->ny;
    for (int32_t s = 0; s <= nsteps; ++s) {
        const int64_t ks = adj ? s : 0;
        double* xs = X + n * ks;
        double* ysv = Y + n * ks;
        memset(rho, 0, sizeof(double) * (size_t)nc);
        scatter_dd(g, xs, ysv, NULL, coef, n, rho, nthreads);
        double* ex = EX + nc * ks;
        double* ey = EY + nc * ks;
        orc_solve_es_f64(g, eps0, rho, ex, ey);
        if (s == nsteps) break;
        double* xn = adj ? X + n * (ks + 1) : X;
        double* yn = adj ? Y + n * (ks + 1) : Y;
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t p = 0; p < n; ++p) {
            const cic_grad c = cic_grad_d(g, xs[p], ysv[p]);
            double e0 = 0.0, e1 = 0.0;
            for (int k = 0; k < 4; ++k) { e0 += c.w[k] * ex[c.c[k]]; e1 += c.w[k] * ey[c.c[k]]; }
            /* step 0 starts with the half step back, u^{-1/2} = R'(u^0 + h' e) + h' e
             * (Pusher::half_step_back, dt -> -dt/2), then every step is
             * u^{n+1/2} = R(u^{n-1/2} + h e) + h e with the Boris rotation R of B0 z */
            if (s == 0) boris_z_d(&ux[p], &uy[p], e0, e1, &Rh);
            boris_z_d(&ux[p], &uy[p], e0, e1, &R1);
            xn[p] = wrap_d(xs[p] + cx * ux[p], fnx);
            yn[p] = wrap_d(ysv[p] + cy * uy[p], fny);
        }
    }
    const double* exN = EX + nc * (adj ? nsteps : 0);
    const double* eyN = EY + nc * (adj ? nsteps : 0);
    double loss = 0.0;
    for (int64_t c = 0; c < nc; ++c) loss += exN[c] * exN[c] + eyN[c] * eyN[c];
    loss *= 0.5 * eps0 * area;
    if (adj) {
        double* lEx = (double*)malloc(sizeof(double) * (size_t)nc);
        double* lEy = (double*)malloc(sizeof(double) * (size_t)nc);
        double* lrho = (double*)malloc(sizeof(double) * (size_t)nc);
        double* lex = (double*)malloc(sizeof(double) * (size_t)n);
        double* ley = (double*)malloc(sizeof(double) * (size_t)n);
        for (int64_t c = 0; c < nc; ++c) { lEx[c] = eps0 * area * exN[c]; lEy[c] = eps0 * area * eyN[c]; }
        orc_solve_es_adj_f64(g, eps0, lEx, lEy, lrho);
        const double* xN = X + n * nsteps;
        const double* yN = Y + n * nsteps;
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t p = 0; p < n; ++p) {
            const cic_grad c = cic_grad_d(g, xN[p], yN[p]);
            double sx = 0.0, sy = 0.0;
            for (int k = 0; k < 4; ++k) { sx += c.dwx[k] * lrho[c.c[k]]; sy += c.dwy[k] * lrho[c.c[k]]; }
            gx[p] = coef * sx;
            gy[p] = coef * sy;
            gux[p] = 0.0;
            guy[p] = 0.0;
        }
        for (int32_t s = nsteps - 1; s >= 0; --s) {
            const double* xs = X + n * s;
            const double* ysv = Y + n * s;
            const double* ex = EX + nc * s;
            const double* ey = EY + nc * s;
#pragma omp parallel for num_threads(nthreads) schedule(static)
            for (int64_t p = 0; p < n; ++p) {
                gux[p] += cx * gx[p];
                guy[p] += cy * gy[p];
                /* transpose of u' = R(u + h e) + h e: le = h (I + R^T) lu', lu = R^T lu' */
                double le0 = 0.0, le1 = 0.0;
                boris_z_adj_d(&gux[p], &guy[p], &le0, &le1, &R1);
                if (s == 0) boris_z_adj_d(&gux[p], &guy[p], &le0, &le1, &Rh);
                const cic_grad c = cic_grad_d(g, xs[p], ysv[p]);
                double dxx = 0, dxy = 0, dyx = 0, dyy = 0;  /* d e_comp / d pos */
                for (int k = 0; k < 4; ++k) {
                    dxx += c.dwx[k] * ex[c.c[k]]; dxy += c.dwy[k] * ex[c.c[k]];
                    dyx += c.dwx[k] * ey[c.c[k]]; dyy += c.dwy[k] * ey[c.c[k]];
                }
                gx[p] += le0 * dxx + le1 * dyx;
                gy[p] += le0 * dxy + le1 * dyy;
                lex[p] = le0;
                ley[p] = le1;
            }
            memset(lEx, 0, sizeof(double) * (size_t)nc);
            memset(lEy, 0, sizeof(double) * (size_t)nc);
            scatter_dd(g, xs, ysv, lex, 0.0, n, lEx, nthreads);
            scatter_dd(g, xs, ysv, ley, 0.0, n, lEy, nthreads);
            orc_solve_es_adj_f64(g, eps0, lEx, lEy, lrho);
#pragma omp parallel for num_threads(nthreads) schedule(static)
            for (int64_t p = 0; p < n; ++p) {
                const cic_grad c = cic_grad_d(g, xs[p], ysv[p]);
                double sx = 0.0, sy = 0.0;
                for (int k = 0; k < 4; ++k) { sx += c.dwx[k] * lrho[c.c[k]]; sy += c.dwy[k] * lrho[c.c[k]]; }
                gx[p] += coef * sx;
                gy[p] += coef * sy;
            }
        }
        free(lEx); free(lEy); free(lrho); free(lex); free(ley);
    }
    free(X); free(Y); free(EX); free(EY); free(ux); free(uy); free(rho);
    return loss;
}
