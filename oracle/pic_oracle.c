/*
 * pic_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see pic_oracle.h).
 *
 * Clean-room FP64 restatement of vvoidpointer/PIC3D's per-step loop on a
 * single domain.  Expressions keep the reference's association order
 * (build with -ffp-contract=off) so that, single-threaded, results agree
 * bit-for-bit with the reference compiled from its own sources.
 *
 * Single-domain reading of the reference's MPI layer (one rank, Cartesian
 * dims 1x1x1): a periodic axis has itself as both neighbours, an open axis
 * has MPI_PROC_NULL (-2 under OpenMPI, which the reference's "-2" tests
 * assume).  Field halos are wrapped X then Y then Z (the reference posts the
 * six faces concurrently, which leaves edge/corner ghosts timing-dependent;
 * the sequential order gives the exact periodic image).
 */
#include "pic_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PIC_LD_FN static inline
#include "../pic3d_amd/csrc/pic_loader_impl.h"

typedef struct {
    double q, m;
    int immobile, declared;
    int64_t n, cap;
    double *x, *y, *z, *xo, *yo, *zo, *ux, *uy, *uz, *w;
    int64_t* id;
    double *sx, *sy, *sz; /* spin (precessed when spin is on and m == m_e) */
    int64_t spin_errors;
} orc_species;

struct orc_ctx {
    pic_params p;
    int m, hm;
    int n[3];  /* cells per axis */
    int fe[3]; /* E/B extents n + 2m */
    int je[3]; /* J/rho extents n + 2m + 2 */
    int64_t fsize, jsize;
    double* F[6]; /* ex ey ez bx by bz */
    double* J[3]; /* jx jy jz */
    double* rho[PIC_MAX_SPECIES];
    orc_species sp[PIC_MAX_SPECIES];
    int nthreads;
    double* tbuf; /* per-thread J + rho scratch, nthreads * 4 * jsize */
    double rho_bg;
    int64_t oob;
    /* Gauss diagnostics: mobile charge at the last clear, recorded residual */
    double* rho_prev;
    double* gauss_ref;
    int rho_prev_valid, gauss_ref_valid;
    int xmax_face; /* 1: the x-max write at the last dual column n+m-1 */
    /* spin precession, particle.cpp:318-368 */
    int spin_on;
    double spin_mass, e_schwinger;
    double* gtab;
    int ngtab;
};

/* E/B index for buffer indices (k, j, i) in [0, n+2m-1]. */
#define FI(c, k, j, i) ((((int64_t)(k)) * (c)->fe[1] + (j)) * (c)->fe[0] + (i))
/* J/rho index for buffer indices (k, j, i) in [-1, n+2m]. */
#define JI(c, k, j, i) ((((int64_t)(k) + 1) * (c)->je[1] + ((j) + 1)) * (c)->je[0] + ((i) + 1))

enum { EX = 0, EY, EZ, BX, BY, BZ };

/* ------------------------------------------------------------------ */
/* shape functions: particle.cpp:8-19 (quadratic), 21-48 (quartic)     */
/* ------------------------------------------------------------------ */

void orc_shape2(double x, double* W) {
    int ic = (int)(x + 0.5);
    double d = x - (double)ic;
    W[0] = 0.5 * (0.5 - d) * (0.5 - d);
    W[1] = 0.75 - d * d;
    W[2] = 0.5 * (0.5 + d) * (0.5 + d);
}

/* quartic B-spline coefficients, constants.cpp:84-85 */
static const double A0 = 55.0 / 96.0, A1 = 5.0 / 24.0, A2 = 5.0 / 4.0, A3 = 5.0 / 6.0,
                    B0 = 115.0 / 192.0, B2 = 5.0 / 8.0;

void orc_shape4(double x, double* W) {
    int ii = (int)(x + 0.5);
    double d = x - (double)(ii - 2); /* distance from node ii-2, in [1.5, 2.5] */
    double t, t2;
    t = d - 2.5;
    t2 = t * t;
    W[0] = t2 * t2 / 24.0;
    t = d - 1.0;
    t2 = t * t;
    W[1] = -t2 * t2 / 6.0 + t2 * t * A3 - t2 * A2 + t * A1 + A0;
    t = d - 2.0;
    t2 = t * t;
    W[2] = t2 * t2 / 4.0 - t2 * B2 + B0;
    t = d - 3.0;
    t2 = t * t;
    W[3] = -t2 * t2 / 6.0 - t2 * t * A3 - t2 * A2 - t * A1 + A0;
    t = d - 1.5;
    t2 = t * t;
    W[4] = t2 * t2 / 24.0;
}

static inline void shape(int m, double x, double* W) {
    if (m == 2) orc_shape2(x, W); else orc_shape4(x, W);
}

/* ------------------------------------------------------------------ */
/* context                                                            */
/* ------------------------------------------------------------------ */

/* The state was set from outside the time loop: the Gauss residual's
 * rho_prev and reference no longer apply. */
static void gauss_invalidate(orc_ctx* c) {
    c->rho_prev_valid = 0;
    c->gauss_ref_valid = 0;
}

orc_ctx* orc_create(const pic_params* p, int nthreads) {
    if (p->mem != 2 && p->mem != 4) return NULL;
    orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    c->p = *p;
    c->m = p->mem;
    c->hm = p->mem / 2;
    c->n[0] = p->nx;
    c->n[1] = p->ny;
    c->n[2] = p->nz;
    for (int a = 0; a < 3; a++) {
        c->fe[a] = c->n[a] + 2 * c->m;
        c->je[a] = c->n[a] + 2 * c->m + 2;
    }
    c->fsize = (int64_t)c->fe[0] * c->fe[1] * c->fe[2];
    c->jsize = (int64_t)c->je[0] * c->je[1] * c->je[2];
    for (int f = 0; f < 6; f++) c->F[f] = (double*)calloc((size_t)c->fsize, sizeof(double));
    for (int f = 0; f < 3; f++) c->J[f] = (double*)calloc((size_t)c->jsize, sizeof(double));
    for (int s = 0; s < PIC_MAX_SPECIES; s++)
        c->rho[s] = (double*)calloc((size_t)c->jsize, sizeof(double));
    c->rho_prev = (double*)calloc((size_t)c->jsize, sizeof(double));
    c->gauss_ref = (double*)calloc((size_t)c->jsize, sizeof(double));
    c->nthreads = nthreads < 1 ? 1 : nthreads;
    return c;
}

static void species_free(orc_species* s) {
    free(s->x); free(s->y); free(s->z); free(s->xo); free(s->yo); free(s->zo);
    free(s->ux); free(s->uy); free(s->uz); free(s->w); free(s->id);
    free(s->sx); free(s->sy); free(s->sz);
    s->x = s->y = s->z = s->xo = s->yo = s->zo = s->ux = s->uy = s->uz = s->w = NULL;
    s->sx = s->sy = s->sz = NULL;
    s->id = NULL;
    s->n = s->cap = 0;
    s->spin_errors = 0;
}

void orc_destroy(orc_ctx* c) {
    if (!c) return;
    free(c->gtab);
    for (int f = 0; f < 6; f++) free(c->F[f]);
    for (int f = 0; f < 3; f++) free(c->J[f]);
    for (int s = 0; s < PIC_MAX_SPECIES; s++) {
        free(c->rho[s]);
        species_free(&c->sp[s]);
    }
    free(c->tbuf);
    free(c->rho_prev);
    free(c->gauss_ref);
    free(c);
}

int orc_species_set(orc_ctx* c, int s, double charge, double mass, int immobile) {
    if (s < 0 || s >= PIC_MAX_SPECIES) return PIC_ERR_ARG;
    gauss_invalidate(c);
    c->sp[s].q = charge;
    c->sp[s].m = mass;
    c->sp[s].immobile = immobile;
    c->sp[s].declared = 1;
    return PIC_OK;
}

void orc_set_background(orc_ctx* c, double rho_bg) {
    c->rho_bg = rho_bg;
    gauss_invalidate(c);
}

static void species_alloc(orc_species* s, int64_t n) {
    species_free(s);
    size_t b = (size_t)(n > 0 ? n : 1) * sizeof(double);
    s->x = (double*)malloc(b); s->y = (double*)malloc(b); s->z = (double*)malloc(b);
    s->xo = (double*)malloc(b); s->yo = (double*)malloc(b); s->zo = (double*)malloc(b);
    s->ux = (double*)malloc(b); s->uy = (double*)malloc(b); s->uz = (double*)malloc(b);
    s->w = (double*)malloc(b);
    s->sx = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    s->sy = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    s->sz = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    s->id = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    s->cap = n;
    s->n = n;
}

int orc_particles_set(orc_ctx* c, int s, int64_t n, const double* x, const double* y,
                      const double* z, const double* ux, const double* uy, const double* uz,
                      const double* w, const int64_t* id) {
    if (s < 0 || s >= PIC_MAX_SPECIES || n < 0) return PIC_ERR_ARG;
    gauss_invalidate(c);
    orc_species* S = &c->sp[s];
    species_alloc(S, n);
    size_t b = (size_t)n * sizeof(double);
    if (n) {
        memcpy(S->x, x, b); memcpy(S->y, y, b); memcpy(S->z, z, b);
        memcpy(S->xo, x, b); memcpy(S->yo, y, b); memcpy(S->zo, z, b);
        memcpy(S->ux, ux, b); memcpy(S->uy, uy, b); memcpy(S->uz, uz, b);
        memcpy(S->w, w, b);
        memcpy(S->id, id, (size_t)n * sizeof(int64_t));
    }
    return PIC_OK;
}

int64_t orc_num_particles(const orc_ctx* c, int s) { return c->sp[s].n; }

void orc_particles_get(const orc_ctx* c, int s, double* x, double* y, double* z, double* ux,
                       double* uy, double* uz, double* w, int64_t* id, double* x_old,
                       double* y_old, double* z_old) {
    const orc_species* S = &c->sp[s];
    size_t b = (size_t)S->n * sizeof(double);
    if (x) memcpy(x, S->x, b);
    if (y) memcpy(y, S->y, b);
    if (z) memcpy(z, S->z, b);
    if (ux) memcpy(ux, S->ux, b);
    if (uy) memcpy(uy, S->uy, b);
    if (uz) memcpy(uz, S->uz, b);
    if (w) memcpy(w, S->w, b);
    if (id) memcpy(id, S->id, (size_t)S->n * sizeof(int64_t));
    if (x_old) memcpy(x_old, S->xo, b);
    if (y_old) memcpy(y_old, S->yo, b);
    if (z_old) memcpy(z_old, S->zo, b);
}

static double* comp_ptr(const orc_ctx* c, int comp, int64_t* size) {
    if (comp >= PIC_EX && comp <= PIC_BZ) { *size = c->fsize; return c->F[comp]; }
    if (comp >= PIC_JX && comp <= PIC_JZ) { *size = c->jsize; return c->J[comp - PIC_JX]; }
    if (comp >= PIC_RHO0 && comp <= PIC_RHO3) { *size = c->jsize; return c->rho[comp - PIC_RHO0]; }
    *size = 0;
    return NULL;
}

void orc_field_get(const orc_ctx* c, int comp, double* out) {
    int64_t n;
    double* p = comp_ptr(c, comp, &n);
    if (p) memcpy(out, p, (size_t)n * sizeof(double));
}

void orc_field_set(orc_ctx* c, int comp, const double* in) {
    int64_t n;
    gauss_invalidate(c);
    double* p = comp_ptr(c, comp, &n);
    if (p) memcpy(p, in, (size_t)n * sizeof(double));
}

int64_t orc_out_of_bounds(const orc_ctx* c) { return c->oob; }

/* ------------------------------------------------------------------ */
/* field halo wrap: exchange_boundary_data, mpi_utils.cpp:9-60         */
/* ------------------------------------------------------------------ */

/* Periodic axis a: left ghosts [0, m-1] <- [N, N+m-1], right ghosts
 * [N+m, N+2m-1] <- [m, 2m-1], over the full extent of the other axes. */
static void wrap_axis(const orc_ctx* c, double* F, int a) {
    const int m = c->m, N = c->n[a];
    const int e0 = c->fe[0], e1 = c->fe[1], e2 = c->fe[2];
    if (a == 0) {
        for (int k = 0; k < e2; k++)
            for (int j = 0; j < e1; j++)
                for (int g = 0; g < m; g++) {
                    F[FI(c, k, j, g)] = F[FI(c, k, j, N + g)];
                    F[FI(c, k, j, m + N + g)] = F[FI(c, k, j, m + g)];
                }
    } else if (a == 1) {
        for (int k = 0; k < e2; k++)
            for (int g = 0; g < m; g++)
                for (int i = 0; i < e0; i++) {
                    F[FI(c, k, g, i)] = F[FI(c, k, N + g, i)];
                    F[FI(c, k, m + N + g, i)] = F[FI(c, k, m + g, i)];
                }
    } else {
        for (int g = 0; g < m; g++)
            for (int j = 0; j < e1; j++)
                for (int i = 0; i < e0; i++) {
                    F[FI(c, g, j, i)] = F[FI(c, N + g, j, i)];
                    F[FI(c, m + N + g, j, i)] = F[FI(c, m + g, j, i)];
                }
    }
}

void orc_field_wrap(orc_ctx* c, int comp) {
    for (int a = 0; a < 3; a++)
        if (c->p.periodic[a]) wrap_axis(c, c->F[comp], a);
}

/* ------------------------------------------------------------------ */
/* Yee half steps: field.cpp:56-117                                    */
/* ------------------------------------------------------------------ */

/* E += coef (curl B - J); coef = dt_half (constants.cpp:42) in the split
 * step, dt in update_electromagnetic_fields (field.cpp:24-54) */
static void e_step(orc_ctx* c, double coef) {
    const int m = c->m;
    const double dth = coef;
    const double dx = c->p.dx, dy = c->p.dy, dz = c->p.dz;
    double *ex = c->F[EX], *ey = c->F[EY], *ez = c->F[EZ];
    const double *bx = c->F[BX], *by = c->F[BY], *bz = c->F[BZ];
    const double *jx = c->J[0], *jy = c->J[1], *jz = c->J[2];
#pragma omp parallel for collapse(2) num_threads(c->nthreads) schedule(static)
    for (int k = m; k < m + c->n[2]; k++)
        for (int j = m; j < m + c->n[1]; j++)
            for (int i = m; i < m + c->n[0]; i++) {
                int64_t o = FI(c, k, j, i);
                ex[o] += dth * ((bz[o] - bz[FI(c, k, j - 1, i)]) / dy -
                                (by[o] - by[FI(c, k - 1, j, i)]) / dz - jx[JI(c, k, j, i)]);
                ey[o] += dth * ((bx[o] - bx[FI(c, k - 1, j, i)]) / dz -
                                (bz[o] - bz[FI(c, k, j, i - 1)]) / dx - jy[JI(c, k, j, i)]);
                ez[o] += dth * ((by[o] - by[FI(c, k, j, i - 1)]) / dx -
                                (bx[o] - bx[FI(c, k, j - 1, i)]) / dy - jz[JI(c, k, j, i)]);
            }
    orc_field_wrap(c, EX);
    orc_field_wrap(c, EY);
    orc_field_wrap(c, EZ);
}

static void b_step(orc_ctx* c, double coef) {
    const int m = c->m;
    const double dth = coef;
    const double dx = c->p.dx, dy = c->p.dy, dz = c->p.dz;
    const double *ex = c->F[EX], *ey = c->F[EY], *ez = c->F[EZ];
    double *bx = c->F[BX], *by = c->F[BY], *bz = c->F[BZ];
#pragma omp parallel for collapse(2) num_threads(c->nthreads) schedule(static)
    for (int k = m; k < m + c->n[2]; k++)
        for (int j = m; j < m + c->n[1]; j++)
            for (int i = m; i < m + c->n[0]; i++) {
                int64_t o = FI(c, k, j, i);
                bx[o] -= dth * ((ez[FI(c, k, j + 1, i)] - ez[o]) / dy -
                                (ey[FI(c, k + 1, j, i)] - ey[o]) / dz);
                by[o] -= dth * ((ex[FI(c, k + 1, j, i)] - ex[o]) / dz -
                                (ez[FI(c, k, j, i + 1)] - ez[o]) / dx);
                bz[o] -= dth * ((ey[FI(c, k, j, i + 1)] - ey[o]) / dx -
                                (ex[FI(c, k, j + 1, i)] - ex[o]) / dy);
            }
    orc_field_wrap(c, BX);
    orc_field_wrap(c, BY);
    orc_field_wrap(c, BZ);
}

void orc_fields_first(orc_ctx* c) { e_step(c, c->p.dt / 2.0); b_step(c, c->p.dt / 2.0); }
void orc_fields_second(orc_ctx* c) { b_step(c, c->p.dt / 2.0); e_step(c, c->p.dt / 2.0); }
/* update_electromagnetic_fields, field.cpp:24-54 */
void orc_fields_full(orc_ctx* c) { e_step(c, c->p.dt); b_step(c, c->p.dt); }

/* ------------------------------------------------------------------ */
/* laser source + Silver-Mueller: boundary.cpp:8-99                    */
/* ------------------------------------------------------------------ */

/* generate_laser_pulse, boundary.cpp:8-27 */
static double laser_envelope(const pic_params* P, double time, double position) {
    double t = time - position - P->laser_start_time;
    double env = 0.0;
    if (t >= 0.0 && t <= 1.0 * P->laser_pulse_duration) env = sin(P->pi * t / P->laser_pulse_duration);
    return P->laser_amplitude * env;
}

/* Gpulse, boundary.cpp:29-40 */
double orc_gpulse(const orc_ctx* c, double time, double x, double y, double z) {
    const pic_params* P = &c->p;
    double k = 2.0 * P->pi / P->l0;
    double zr = 0.5 * k * P->waist * P->waist;
    double xe = x - P->xfocus;
    double gouy = atan(xe / zr);
    double w = P->waist * sqrt(1 + pow(xe / zr, 2));
    double ph = (xe / zr) * (y * y + z * z) / w / w - gouy + k * xe - P->omega * time;
    return -1.0 * laser_envelope(P, time, x) * P->waist / w * exp(-(y * y + z * z) / (w * w)) *
           cos(ph + P->phase);
}

void orc_set_xmax_face(orc_ctx* c, int on) { c->xmax_face = on != 0; }

void orc_laser_boundary(orc_ctx* c, int step) { orc_radiating_boundary(c, step, 1, c->p.laser_right); }

void orc_radiating_boundary(orc_ctx* c, int step, int is_left, int is_right) {
    const pic_params* P = &c->p;
    if (P->periodic[0]) return; /* x neighbours exist: rank_x* != -2 */
    const int m = c->m, N = c->n[0];
    const double dt = P->dt, dx = P->dx, dy = P->dy, dz = P->dz;
    const double time = step * dt;
    const double alpha = 1 / (1 + dt / dx);
    const double beta = 1 - dt / dx;
    double *ey = c->F[EY], *ez = c->F[EZ], *bx = c->F[BX], *by = c->F[BY], *bz = c->F[BZ];
    const double *jy = c->J[1], *jz = c->J[2];
    const int pol = P->laser_polarization;
    for (int k = m; k < m + c->n[2]; k++)
        for (int j = m; j < m + c->n[1]; j++) {
            /* transverse coordinates use the buffer index (boundary.cpp:61-62) */
            double yy = ((double)j + 0.0) * dy - P->yfocus;
            double zz = ((double)k + 0.0) * dz - P->zfocus;
            double Sl = 0.0, Pl = 0.0, Sr = 0.0, Pr = 0.0;
            if (is_left) {
                if (pol == 1 || pol == 3) Sl = orc_gpulse(c, time, 0.0, yy, zz);
                if (pol == 2 || pol == 3) Pl = orc_gpulse(c, time, 0.0, yy, zz);
            }
            if (is_right) {
                if (pol == 1 || pol == 3) Sr = orc_gpulse(c, time, P->lx, yy, zz);
                if (pol == 2 || pol == 3) Pr = orc_gpulse(c, time, P->lx, yy, zz);
            }
            /* x-min ghost column m-1 */
            bz[FI(c, k, j, m - 1)] =
                alpha * (4 * Sl - 2 * ey[FI(c, k, j, m)] -
                         dt / dz * (bx[FI(c, k, j, m)] - bx[FI(c, k - 1, j, m)]) -
                         beta * bz[FI(c, k, j, m)] - dt * jy[JI(c, k, j, m)]);
            by[FI(c, k, j, m - 1)] =
                alpha * (-4 * Pl + 2 * ez[FI(c, k, j, m)] +
                         dt / dy * (bx[FI(c, k, j, m)] - bx[FI(c, k, j - 1, m)]) -
                         beta * by[FI(c, k, j, m)] + dt * jz[JI(c, k, j, m)]);
            /* x-max: column local_xcells+1, reading column local_xcells
             * (boundary.cpp:92-96; an interior column when m > 2).  With
             * xmax_face, the mirror of the x-min write: the last dual column
             * w = n+m-1 from E, Bx and J at the primal column w next to it
             * and By, Bz one dual column in. */
            const int w = c->xmax_face ? N + m - 1 : N + 1;
            const int re = c->xmax_face ? w : w - 1, rb = w - 1;
            bz[FI(c, k, j, w)] =
                alpha * (-4 * Sr + 2 * ey[FI(c, k, j, re)] -
                         dt / dz * (bx[FI(c, k, j, re)] - bx[FI(c, k - 1, j, re)]) -
                         beta * bz[FI(c, k, j, rb)] + dt * jy[JI(c, k, j, re)]);
            by[FI(c, k, j, w)] =
                alpha * (4 * Pr - 2 * ez[FI(c, k, j, re)] +
                         dt / dy * (bx[FI(c, k, j, re)] - bx[FI(c, k, j - 1, re)]) -
                         beta * by[FI(c, k, j, rb)] - dt * jz[JI(c, k, j, re)]);
        }
}

/* ------------------------------------------------------------------ */
/* currents: clear (particle.cpp:576-592), fold (mpi_utils.cpp:232-417) */
/* ------------------------------------------------------------------ */

void orc_clear_currents(orc_ctx* c) {
    /* keep the mobile charge (species order) for the Gauss diagnostics */
    for (int64_t q = 0; q < c->jsize; q++) {
        double sum = 0.0;
        for (int s = 0; s < PIC_MAX_SPECIES; s++)
            if (c->sp[s].declared && !c->sp[s].immobile) sum += c->rho[s][q];
        c->rho_prev[q] = sum;
    }
    c->rho_prev_valid = 1;
    for (int f = 0; f < 3; f++) memset(c->J[f], 0, (size_t)c->jsize * sizeof(double));
    for (int s = 0; s < PIC_MAX_SPECIES; s++) memset(c->rho[s], 0, (size_t)c->jsize * sizeof(double));
}

/* One periodic axis of the accumulate phase: the (m+1)-layer ghost slabs
 * [-1, m-1] and [N+m, N+2m] are added onto the interior slabs
 * [N-1, N+m-1] and [m, 2m], across the full J extent of the other axes. */
static void fold_accumulate(const orc_ctx* c, double* D, int a) {
    const int m = c->m, N = c->n[a];
    const int st = m, end = m + N - 1;
    const int lo = -1, hi0 = c->je[0] - 1, hi1 = c->je[1] - 1, hi2 = c->je[2] - 1;
    if (a == 0) {
        for (int k = lo; k < hi2; k++)
            for (int j = lo; j < hi1; j++)
                for (int i = 0; i <= m; i++) {
                    double from_right = D[JI(c, k, j, end + 1 + i)];
                    double from_left = D[JI(c, k, j, -1 + i)];
                    D[JI(c, k, j, st + i)] += from_right;
                    D[JI(c, k, j, end - m + i)] += from_left;
                }
    } else if (a == 1) {
        for (int k = lo; k < hi2; k++)
            for (int j = 0; j <= m; j++)
                for (int i = lo; i < hi0; i++) {
                    double from_right = D[JI(c, k, end + 1 + j, i)];
                    double from_left = D[JI(c, k, -1 + j, i)];
                    D[JI(c, k, st + j, i)] += from_right;
                    D[JI(c, k, end - m + j, i)] += from_left;
                }
    } else {
        for (int k = 0; k <= m; k++)
            for (int j = lo; j < hi1; j++)
                for (int i = lo; i < hi0; i++) {
                    double from_right = D[JI(c, end + 1 + k, j, i)];
                    double from_left = D[JI(c, -1 + k, j, i)];
                    D[JI(c, st + k, j, i)] += from_right;
                    D[JI(c, end - m + k, j, i)] += from_left;
                }
    }
}

/* Sync phase: ghosts [-1, m-1] <- [N-1, N+m-1], [N+m, N+2m] <- [m, 2m]. */
static void fold_sync(const orc_ctx* c, double* D, int a) {
    const int m = c->m, N = c->n[a];
    const int st = m, end = m + N - 1;
    const int lo = -1, hi0 = c->je[0] - 1, hi1 = c->je[1] - 1, hi2 = c->je[2] - 1;
    if (a == 0) {
        for (int k = lo; k < hi2; k++)
            for (int j = lo; j < hi1; j++)
                for (int i = 0; i <= m; i++) {
                    D[JI(c, k, j, -1 + i)] = D[JI(c, k, j, end - m + i)];
                    D[JI(c, k, j, end + 1 + i)] = D[JI(c, k, j, st + i)];
                }
    } else if (a == 1) {
        for (int k = lo; k < hi2; k++)
            for (int j = 0; j <= m; j++)
                for (int i = lo; i < hi0; i++) {
                    D[JI(c, k, -1 + j, i)] = D[JI(c, k, end - m + j, i)];
                    D[JI(c, k, end + 1 + j, i)] = D[JI(c, k, st + j, i)];
                }
    } else {
        for (int k = 0; k <= m; k++)
            for (int j = lo; j < hi1; j++)
                for (int i = lo; i < hi0; i++) {
                    D[JI(c, -1 + k, j, i)] = D[JI(c, end - m + k, j, i)];
                    D[JI(c, end + 1 + k, j, i)] = D[JI(c, st + k, j, i)];
                }
    }
}

static void fold_one(const orc_ctx* c, double* D) {
    for (int a = 0; a < 3; a++)
        if (c->p.periodic[a]) fold_accumulate(c, D, a);
    for (int a = 0; a < 3; a++)
        if (c->p.periodic[a]) fold_sync(c, D, a);
}

/* exchange_boundary_current_fixed of one array: comp 6..8 = jx..jz,
 * 9.. = rho of species comp-9 */
void orc_fold_one(orc_ctx* c, int comp) {
    if (comp >= 6 && comp <= 8) fold_one(c, c->J[comp - 6]);
    else if (comp >= 9 && comp < 9 + PIC_MAX_SPECIES) fold_one(c, c->rho[comp - 9]);
}

/* binomial_filter_currents, particle.cpp:520-574: per array jx, jy, jz a
 * 1-2-1 pass along x over the interior rows, then along y, then along z,
 * each pass reading the line as it was before that pass (the reference
 * copies each line into a buffer, ghosts included, before filtering). */
static void filter_array(orc_ctx* c, double* D) {
    const int m = c->m;
    const int nx = c->n[0], ny = c->n[1], nz = c->n[2];
    int nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
    double* buf = (double*)malloc((size_t)(nmax + 2) * sizeof(double));
    for (int k = m; k < m + nz; k++)
        for (int j = m; j < m + ny; j++) {
            for (int i = -1; i <= nx; i++) buf[i + 1] = D[JI(c, k, j, m + i)];
            for (int i = 0; i < nx; i++)
                D[JI(c, k, j, m + i)] = 0.25 * buf[i] + 0.5 * buf[i + 1] + 0.25 * buf[i + 2];
        }
    for (int k = m; k < m + nz; k++)
        for (int i = m; i < m + nx; i++) {
            for (int j = -1; j <= ny; j++) buf[j + 1] = D[JI(c, k, m + j, i)];
            for (int j = 0; j < ny; j++)
                D[JI(c, k, m + j, i)] = 0.25 * buf[j] + 0.5 * buf[j + 1] + 0.25 * buf[j + 2];
        }
    for (int j = m; j < m + ny; j++)
        for (int i = m; i < m + nx; i++) {
            for (int k = -1; k <= nz; k++) buf[k + 1] = D[JI(c, m + k, j, i)];
            for (int k = 0; k < nz; k++)
                D[JI(c, m + k, j, i)] = 0.25 * buf[k] + 0.5 * buf[k + 1] + 0.25 * buf[k + 2];
        }
    free(buf);
}

void orc_filter_currents(orc_ctx* c) {
    for (int f = 0; f < 3; f++) filter_array(c, c->J[f]);
}

void orc_fold_currents(orc_ctx* c) {
    for (int f = 0; f < 3; f++) fold_one(c, c->J[f]);
    for (int s = 0; s < PIC_MAX_SPECIES; s++)
        if (c->sp[s].declared) fold_one(c, c->rho[s]);
}

/* ------------------------------------------------------------------ */
/* Esirkepov deposit: currents_particle, particle.cpp:372-515          */
/* ------------------------------------------------------------------ */

static void deposit_one(const orc_ctx* c, double q, double wgt, double xo, double yo, double zo,
                        double xn, double yn, double zn, double* jx, double* jy, double* jz,
                        double* rho) {
    const int m = c->m, hm = c->hm;
    const double org = (double)(1 - m); /* nc_xst_new on one domain */
    const double x1 = xo - org, x2 = xn - org;
    const double y1 = yo - org, y2 = yn - org;
    const double z1 = zo - org, z2 = zn - org;
    double Sx0[5], Sx1[5], Sy0[5], Sy1[5], Sz0[5], Sz1[5];
    shape(m, x1, Sx0); shape(m, x2, Sx1);
    shape(m, y1, Sy0); shape(m, y2, Sy1);
    shape(m, z1, Sz0); shape(m, z2, Sz1);
    const int i1 = (int)(x1 + 0.5), i2 = (int)(x2 + 0.5);
    const int j1 = (int)(y1 + 0.5), j2 = (int)(y2 + 0.5);
    const int k1 = (int)(z1 + 0.5), k2 = (int)(z2 + 0.5);
    const int imin = (i1 < i2 ? i1 : i2) - hm, imax = (i1 > i2 ? i1 : i2) + hm;
    const int jmin = (j1 < j2 ? j1 : j2) - hm, jmax = (j1 > j2 ? j1 : j2) + hm;
    const int kmin = (k1 < k2 ? k1 : k2) - hm, kmax = (k1 > k2 ? k1 : k2) + hm;
    const double qw = q * wgt;
    const double cx = qw * c->p.dx / c->p.dt;
    const double cy = qw * c->p.dy / c->p.dt;
    const double cz = qw * c->p.dz / c->p.dt;
#define SV(S, ic, v) (((v) >= (ic) - hm && (v) <= (ic) + hm) ? (S)[(v) - (ic) + hm] : 0.0)
    for (int k = kmin; k <= kmax; k++) {
        double sz0 = SV(Sz0, k1, k), dsz = SV(Sz1, k2, k) - sz0;
        for (int j = jmin; j <= jmax; j++) {
            double sy0 = SV(Sy0, j1, j), dsy = SV(Sy1, j2, j) - sy0;
            double wyz = sy0 * sz0 + 0.5 * dsy * sz0 + 0.5 * sy0 * dsz + (1.0 / 3.0) * dsy * dsz;
            double acc = 0.0;
            for (int i = imin; i <= imax; i++) {
                acc += SV(Sx1, i2, i) - SV(Sx0, i1, i);
                jx[JI(c, k, j, i)] -= cx * acc * wyz;
            }
        }
    }
    for (int k = kmin; k <= kmax; k++) {
        double sz0 = SV(Sz0, k1, k), dsz = SV(Sz1, k2, k) - sz0;
        for (int i = imin; i <= imax; i++) {
            double sx0 = SV(Sx0, i1, i), dsx = SV(Sx1, i2, i) - sx0;
            double wxz = sx0 * sz0 + 0.5 * dsx * sz0 + 0.5 * sx0 * dsz + (1.0 / 3.0) * dsx * dsz;
            double acc = 0.0;
            for (int j = jmin; j <= jmax; j++) {
                acc += SV(Sy1, j2, j) - SV(Sy0, j1, j);
                jy[JI(c, k, j, i)] -= cy * acc * wxz;
            }
        }
    }
    for (int j = jmin; j <= jmax; j++) {
        double sy0 = SV(Sy0, j1, j), dsy = SV(Sy1, j2, j) - sy0;
        for (int i = imin; i <= imax; i++) {
            double sx0 = SV(Sx0, i1, i), dsx = SV(Sx1, i2, i) - sx0;
            double wxy = sx0 * sy0 + 0.5 * dsx * sy0 + 0.5 * sx0 * dsy + (1.0 / 3.0) * dsx * dsy;
            double acc = 0.0;
            for (int k = kmin; k <= kmax; k++) {
                acc += SV(Sz1, k2, k) - SV(Sz0, k1, k);
                jz[JI(c, k, j, i)] -= cz * acc * wxy;
            }
        }
    }
    for (int k = kmin; k <= kmax; k++) {
        double sz = SV(Sz1, k2, k);
        if (sz == 0.0) continue;
        for (int j = jmin; j <= jmax; j++) {
            double sy = SV(Sy1, j2, j);
            if (sy == 0.0) continue;
            for (int i = imin; i <= imax; i++) {
                double sx = SV(Sx1, i2, i);
                rho[JI(c, k, j, i)] += qw * sx * sy * sz;
            }
        }
    }
#undef SV
}

/* ------------------------------------------------------------------ */
/* push: update_particle_position_momentum, particle.cpp:113-316       */
/* ------------------------------------------------------------------ */

/* Returns 0 if pushed, 1 if the particle was outside the domain (counted,
 * left untouched: the reference would read outside its arrays). */
/* ------------------------------------------------------------------ */
/* initialize_particles (particle.cpp:50-111) with the box density of  */
/* plasma_density (constants.cpp:87-103), on one domain (xst = 0,      */
/* xend = n, simulation.cpp:31-34).  Writes at most cap particles;     */
/* returns the number written.                                         */
/* ------------------------------------------------------------------ */
static double box_density(const pic_params* p, const pic_lattice* L, double x, double y, double z) {
    const double xp = x * p->dx, yp = y * p->dy, zp = z * p->dz;
    if (xp > L->lo[0] * p->l0 && xp <= L->hi[0] * p->l0 && yp > L->lo[1] * p->l0 &&
        yp <= L->hi[1] * p->l0 && zp > L->lo[2] * p->l0 && zp <= L->hi[2] * p->l0)
        return L->density;
    return 0.0;
}

int64_t orc_lattice(const pic_params* p, const pic_lattice* L, int64_t cap, double* x, double* y,
                    double* z, double* ux, double* uy, double* uz, double* w, int64_t* id) {
    const int cpp = L->cell_per_particles;
    const int64_t num = (int64_t)p->nx * p->ny * p->nz * cpp;
    const int num_x = p->nx * cbrt(cpp), num_y = p->ny * cbrt(cpp), num_z = p->nz * cbrt(cpp);
    const double norm_x = ((double)p->nx - 0.0) / (num_x);
    const double norm_y = ((double)p->ny - 0.0) / (num_y);
    const double norm_z = ((double)p->nz - 0.0) / (num_z);
    int64_t q = 0;
    for (int k = 0; k < num_z; k++)
        for (int j = 0; j < num_y; j++)
            for (int i = 0; i < num_x; i++) {
                if (q >= num || q >= cap) break;
                const double px = 0.0 + (i)*norm_x + norm_x / 2.0;
                const double py = 0.0 + (j)*norm_y + norm_y / 2.0;
                const double pz = 0.0 + (k)*norm_z + norm_z / 2.0;
                const double wt = box_density(p, L, px, py, pz) / cpp;
                if (wt <= 0.0) continue;
                x[q] = px; y[q] = py; z[q] = pz;
                ux[q] = L->u[0]; uy[q] = L->u[1]; uz[q] = L->u[2];
                w[q] = wt;
                id[q] = L->id_offset + q;
                q++;
            }
    return q;
}

/* ------------------------------------------------------------------ */
/* spin precession, particle.cpp:318-368, with g_factor (603-637)      */
/* ------------------------------------------------------------------ */

int orc_spin_setup(orc_ctx* c, const double* g, int n, double e_schwinger, double m_e) {
    if (n < 1) return PIC_ERR_ARG;
    free(c->gtab);
    c->gtab = (double*)malloc((size_t)n * sizeof(double));
    memcpy(c->gtab, g, (size_t)n * sizeof(double));
    c->ngtab = n;
    c->e_schwinger = e_schwinger;
    c->spin_mass = m_e;
    c->spin_on = 1;
    return PIC_OK;
}

void orc_spin_set(orc_ctx* c, int s, const double* sx, const double* sy, const double* sz) {
    orc_species* S = &c->sp[s];
    size_t b = (size_t)S->n * sizeof(double);
    memcpy(S->sx, sx, b); memcpy(S->sy, sy, b); memcpy(S->sz, sz, b);
}

void orc_spin_get(const orc_ctx* c, int s, double* sx, double* sy, double* sz) {
    const orc_species* S = &c->sp[s];
    size_t b = (size_t)S->n * sizeof(double);
    memcpy(sx, S->sx, b); memcpy(sy, S->sy, b); memcpy(sz, S->sz, b);
}

int64_t orc_spin_errors(const orc_ctx* c, int s) { return c->sp[s].spin_errors; }

/* g_factor, particle.cpp:603-637.  Out of [0, 100] (or NaN) the reference
 * prints and exits: *bad is set.  From chi = 99.9 its index passes the end
 * of the table; here it stops at the last entry. */
double orc_g_factor(const orc_ctx* c, double x, int* bad) {
    int idx;
    if (x < 0.005 && x >= 0.0) return 2.003219;
    else if (x >= 0.005 && x < 0.01) idx = (int)((x - 0.005) / 0.0001);
    else if (x >= 0.01 && x < 1) idx = (int)((x - 0.01) / 0.01) + 50;
    else if (x >= 1 && x < 10) idx = (int)((x - 1.0) / 0.01) + 150;
    else if (x >= 10 && x < 20) idx = (int)((x - 10.0) / 0.01) + 1050;
    else if (x >= 20 && x <= 100) idx = (int)((x - 20.0) / 0.1) + 2050;
    else { *bad = 1; return 2.0; }
    if (idx > c->ngtab - 1) idx = c->ngtab - 1;
    return c->gtab[idx];
}

/* one precession step of particle p (particle.cpp:321-367); e*, b*: the
 * gathered fields times qm, as the Boris push holds them; u, gam: after
 * the push.  Returns 1 where g_factor has no value (spin unchanged). */
static int spin_one(const orc_ctx* c, orc_species* S, int64_t p, double qm, double exq, double eyq,
                    double ezq, double bxq, double byq, double bzq, double ux, double uy, double uz,
                    double gam) {
    const double ex = exq / qm, ey = eyq / qm, ez = ezq / qm;
    const double bx = bxq / qm, by = byq / qm, bz = bzq / qm;
    const double vx = ux / gam, vy = uy / gam, vz = uz / gam;
    const double px0 = vx * ex, py0 = vy * ey, pz0 = vz * ez;
    const double fLx = (ex + vy * bz - vz * by);
    const double fLy = (ey + vz * bx - vx * bz);
    const double fLz = (ez + vx * by - vy * bx);
    double chi = fLx * fLx + fLy * fLy + fLz * fLz - (px0 + py0 + pz0) * (px0 + py0 + pz0);
    chi = 0.0 < chi ? chi : 0.0; /* std::max(0.0, chi) */
    chi = sqrt(chi) / c->e_schwinger * gam;
    int bad = 0;
    const double g = orc_g_factor(c, chi, &bad);
    if (bad) return 1;
    const double alpha = -(g / 2.0 - 1.0) * (gam) / (1.0 + gam);
    const double beta = (g / 2.0 - 1.0) + 1.0 / gam;
    const double gm = g / 2.0 - gam / (1.0 + gam);
    const double VB = vx * bx + vy * by + vz * bz;
    const double ox = alpha * VB * vx + beta * bx + gm * (vy * ez - vz * ey);
    const double oy = alpha * VB * vy + beta * by + gm * (vz * ex - vx * ez);
    const double oz = alpha * VB * vz + beta * bz + gm * (vx * ey - vy * ex);
    const double qmdt2 = S->q / S->m * (c->p.dt / 2.0);
    const double tx = ox * qmdt2, ty = oy * qmdt2, tz = oz * qmdt2;
    const double tt = 2.0 / (1.0 + tx * tx + ty * ty + tz * tz);
    const double sx = S->sx[p], sy = S->sy[p], sz = S->sz[p];
    const double sxr = sx + (sy * tz - sz * ty);
    const double syr = sy + (sz * tx - sx * tz);
    const double szr = sz + (sx * ty - sy * tx);
    S->sx[p] = sx + (syr * tz - szr * ty) * tt;
    S->sy[p] = sy + (szr * tx - sxr * tz) * tt;
    S->sz[p] = sz + (sxr * ty - syr * tx) * tt;
    return 0;
}

/* push_one's return: bit 0 = out of bounds, bit 1 = spin step skipped */
static int push_one(const orc_ctx* c, orc_species* S, int64_t p, double* jx, double* jy,
                    double* jz, double* rho) {
    const pic_params* P = &c->p;
    const int m = c->m, hm = c->hm;
    const double N0 = c->n[0], N1 = c->n[1], N2 = c->n[2];
    double x = S->x[p], y = S->y[p], z = S->z[p];
    /* periodic wrap at the start of the push, particle.cpp:123-167 */
    if (x < 0.0 || x > N0 || y < 0.0 || y > N1 || z < 0.0 || z > N2) {
        if (P->periodic[0] || P->periodic[1] || P->periodic[2]) {
            if (P->periodic[0]) { if (x < 0.0) x += N0; if (x > N0) x -= N0; }
            if (P->periodic[1]) { if (y < 0.0) y += N1; if (y > N1) y -= N1; }
            if (P->periodic[2]) { if (z < 0.0) z += N2; if (z > N2) z -= N2; }
        }
    }
    if (!(x >= 0.0 && x <= N0 && y >= 0.0 && y <= N1 && z >= 0.0 && z <= N2)) return 1;

    const double org = (double)(1 - m);
    const double xc = x - org, yc = y - org, zc = z - org;
    double Wxp[5], Wyp[5], Wzp[5], Wxd[5], Wyd[5], Wzd[5];
    shape(m, xc, Wxp); const int ip = (int)(xc + 0.5);
    shape(m, yc, Wyp); const int jp = (int)(yc + 0.5);
    shape(m, zc, Wzp); const int kp = (int)(zc + 0.5);
    shape(m, xc - 0.5, Wxd); const int id = (int)(xc);
    shape(m, yc - 0.5, Wyd); const int jd = (int)(yc);
    shape(m, zc - 0.5, Wzd); const int kd = (int)(zc);

    const double *ex = c->F[EX], *ey = c->F[EY], *ez = c->F[EZ];
    const double *bx = c->F[BX], *by = c->F[BY], *bz = c->F[BZ];
    double exp_ = 0.0, eyp = 0.0, ezp = 0.0, bxp = 0.0, byp = 0.0, bzp = 0.0;
    /* gather, particle.cpp:207-259: stagger (x,y,z) dual/primal per component */
    for (int k = -hm; k <= hm; k++)
        for (int j = -hm; j <= hm; j++)
            for (int i = -hm; i <= hm; i++)
                exp_ += Wxd[i + hm] * Wyp[j + hm] * Wzp[k + hm] * ex[FI(c, kp + k, jp + j, id + i)];
    for (int k = -hm; k <= hm; k++)
        for (int j = -hm; j <= hm; j++)
            for (int i = -hm; i <= hm; i++)
                eyp += Wxp[i + hm] * Wyd[j + hm] * Wzp[k + hm] * ey[FI(c, kp + k, jd + j, ip + i)];
    for (int k = -hm; k <= hm; k++)
        for (int j = -hm; j <= hm; j++)
            for (int i = -hm; i <= hm; i++)
                ezp += Wxp[i + hm] * Wyp[j + hm] * Wzd[k + hm] * ez[FI(c, kd + k, jp + j, ip + i)];
    for (int k = -hm; k <= hm; k++)
        for (int j = -hm; j <= hm; j++)
            for (int i = -hm; i <= hm; i++)
                bxp += Wxp[i + hm] * Wyd[j + hm] * Wzd[k + hm] * bx[FI(c, kd + k, jd + j, ip + i)];
    for (int k = -hm; k <= hm; k++)
        for (int j = -hm; j <= hm; j++)
            for (int i = -hm; i <= hm; i++)
                byp += Wxd[i + hm] * Wyp[j + hm] * Wzd[k + hm] * by[FI(c, kd + k, jp + j, id + i)];
    for (int k = -hm; k <= hm; k++)
        for (int j = -hm; j <= hm; j++)
            for (int i = -hm; i <= hm; i++)
                bzp += Wxd[i + hm] * Wyd[j + hm] * Wzp[k + hm] * bz[FI(c, kp + k, jd + j, id + i)];

    /* Boris, particle.cpp:262-311 */
    const double pxo = S->ux[p], pyo = S->uy[p], pzo = S->uz[p];
    const double qm = S->q / (S->m) * (P->dt / 2.0);
    exp_ *= qm; eyp *= qm; ezp *= qm; bxp *= qm; byp *= qm; bzp *= qm;
    const double px1 = pxo + exp_, py1 = pyo + eyp, pz1 = pzo + ezp;
    const double g1 = sqrt(1.0 + (px1 * px1 + py1 * py1 + pz1 * pz1));
    const double tx = bxp / g1, ty = byp / g1, tz = bzp / g1;
    const double tt = 2.0 / (1.0 + tx * tx + ty * ty + tz * tz);
    const double sx = tx * tt, sy = ty * tt, sz = tz * tt;
    const double px2 = px1 + (py1 * tz - pz1 * ty);
    const double py2 = py1 + (pz1 * tx - px1 * tz);
    const double pz2 = pz1 + (px1 * ty - py1 * tx);
    const double px3 = px1 + (py2 * sz - pz2 * sy);
    const double py3 = py1 + (pz2 * sx - px2 * sz);
    const double pz3 = pz1 + (px2 * sy - py2 * sx);
    const double ux = px3 + exp_, uy = py3 + eyp, uz = pz3 + ezp;
    const double gam = sqrt(1.0 + (ux * ux + uy * uy + uz * uz));
    /* position advance, particle.cpp:313-315 */
    const double xn = x + ux / gam * P->dt / P->dx;
    const double yn = y + uy / gam * P->dt / P->dy;
    const double zn = z + uz / gam * P->dt / P->dz;
    S->xo[p] = x; S->yo[p] = y; S->zo[p] = z;
    S->x[p] = xn; S->y[p] = yn; S->z[p] = zn;
    S->ux[p] = ux; S->uy[p] = uy; S->uz[p] = uz;
    int spin_bad = 0;
    if (c->spin_on && S->m == c->spin_mass)
        spin_bad = spin_one(c, S, p, qm, exp_, eyp, ezp, bxp, byp, bzp, ux, uy, uz, gam);
    /* a non-finite momentum would put the deposit box outside the arrays */
    if (!(fabs(xn - x) < 1.0 && fabs(yn - y) < 1.0 && fabs(zn - z) < 1.0)) return 1 | (spin_bad << 1);
    deposit_one(c, S->q, S->w[p], x, y, z, xn, yn, zn, jx, jy, jz, rho);
    return spin_bad << 1;
}

/* currents_particle(p, s) for explicit old/new positions and weights,
 * particle.cpp:372-515, in the given order.  Returns the number skipped
 * (old position outside [0, N] or a move of a cell or more). */
int64_t orc_deposit(orc_ctx* c, int s, int64_t n, const double* xo, const double* yo,
                    const double* zo, const double* xn, const double* yn, const double* zn,
                    const double* w) {
    int64_t bad = 0;
    for (int64_t p = 0; p < n; p++) {
        const int inside = xo[p] >= 0.0 && xo[p] <= c->n[0] && yo[p] >= 0.0 && yo[p] <= c->n[1] &&
                           zo[p] >= 0.0 && zo[p] <= c->n[2];
        if (!inside || !(fabs(xn[p] - xo[p]) < 1.0 && fabs(yn[p] - yo[p]) < 1.0 && fabs(zn[p] - zo[p]) < 1.0)) {
            bad++;
            continue;
        }
        deposit_one(c, c->sp[s].q, w[p], xo[p], yo[p], zo[p], xn[p], yn[p], zn[p], c->J[0], c->J[1],
                    c->J[2], c->rho[s]);
    }
    return bad;
}

void orc_push(orc_ctx* c, int s) {
    orc_species* S = &c->sp[s];
    if (S->immobile || S->n == 0) return;
    if (c->nthreads == 1) {
        int64_t oob = 0, sbad = 0;
        for (int64_t p = 0; p < S->n; p++) {
            const int f = push_one(c, S, p, c->J[0], c->J[1], c->J[2], c->rho[s]);
            oob += f & 1;
            sbad += f >> 1;
        }
        c->oob += oob;
        S->spin_errors += sbad;
        return;
    }
#ifdef _OPENMP
    const int nt = c->nthreads;
    const int64_t js = c->jsize;
    if (!c->tbuf) c->tbuf = (double*)malloc((size_t)nt * 4 * (size_t)js * sizeof(double));
    int64_t oob = 0, sbad = 0;
#pragma omp parallel num_threads(nt) reduction(+ : oob, sbad)
    {
        int t = omp_get_thread_num();
        double* b = c->tbuf + (size_t)t * 4 * (size_t)js;
        memset(b, 0, 4 * (size_t)js * sizeof(double));
        int64_t lo = S->n * t / nt, hi = S->n * (t + 1) / nt;
        for (int64_t p = lo; p < hi; p++) {
            const int f = push_one(c, S, p, b, b + js, b + 2 * js, b + 3 * js);
            oob += f & 1;
            sbad += f >> 1;
        }
#pragma omp barrier
        /* reduce thread buffers in thread order, split over elements */
        int64_t e0 = js * t / nt, e1 = js * (t + 1) / nt;
        double* dst[4] = {c->J[0], c->J[1], c->J[2], c->rho[s]};
        for (int f = 0; f < 4; f++)
            for (int tt = 0; tt < nt; tt++) {
                const double* src = c->tbuf + (size_t)tt * 4 * (size_t)js + (size_t)f * js;
                for (int64_t e = e0; e < e1; e++) dst[f][e] += src[e];
            }
    }
    c->oob += oob;
    S->spin_errors += sbad;
#endif
}

/* ------------------------------------------------------------------ */
/* particle boundary: exchange_particle_data on one domain,            */
/* mpi_utils.cpp:62-230                                                */
/* ------------------------------------------------------------------ */

static void permute(orc_species* S, const int64_t* idx, int64_t n) {
    double* tmp = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    double* arrs[13] = {S->x, S->y, S->z, S->xo, S->yo, S->zo, S->ux, S->uy, S->uz, S->w,
                        S->sx, S->sy, S->sz};
    for (int a = 0; a < 13; a++) {
        for (int64_t q = 0; q < n; q++) tmp[q] = arrs[a][idx[q]];
        memcpy(arrs[a], tmp, (size_t)n * sizeof(double));
    }
    free(tmp);
    int64_t* ti = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    for (int64_t q = 0; q < n; q++) ti[q] = S->id[idx[q]];
    memcpy(S->id, ti, (size_t)n * sizeof(int64_t));
    free(ti);
}

void orc_particle_boundary(orc_ctx* c, int s, pic_counts* out) {
    orc_species* S = &c->sp[s];
    pic_counts cnt;
    memset(&cnt, 0, sizeof(cnt));
    if (S->immobile) {
        cnt.remaining = S->n;
        if (out) *out = cnt;
        return;
    }
    int64_t* keep = (int64_t*)malloc((size_t)(S->n > 0 ? S->n : 1) * sizeof(int64_t));
    int64_t* minus = (int64_t*)malloc((size_t)(S->n > 0 ? S->n : 1) * sizeof(int64_t));
    int64_t* plus = (int64_t*)malloc((size_t)(S->n > 0 ? S->n : 1) * sizeof(int64_t));
    for (int a = 0; a < 3; a++) {
        const double* v = a == 0 ? S->x : (a == 1 ? S->y : S->z);
        const double N = (double)c->n[a];
        const int per = c->p.periodic[a];
        int64_t nk = 0, nm = 0, np = 0;
        for (int64_t q = 0; q < S->n; q++) {
            double t = v[q];
            if (per && t < 0.0) minus[nm++] = q;       /* sent to x- (self)  */
            else if (per && t >= N) plus[np++] = q;    /* sent to x+ (self)  */
            else if (!per && (t < 0.0 || t >= N)) {    /* removed at the edge */
                if (t < 0.0) cnt.lost[2 * a]++; else cnt.lost[2 * a + 1]++;
            } else keep[nk++] = q;
        }
        cnt.wrapped[2 * a] += nm;
        cnt.wrapped[2 * a + 1] += np;
        /* survivors in order, then arrivals from x- (what went out at x+),
         * then arrivals from x+ (what went out at x-), mpi_utils.cpp:122-123 */
        for (int64_t q = 0; q < np; q++) keep[nk++] = plus[q];
        for (int64_t q = 0; q < nm; q++) keep[nk++] = minus[q];
        permute(S, keep, nk);
        S->n = nk;
    }
    free(keep); free(minus); free(plus);
    cnt.remaining = S->n;
    if (out) *out = cnt;
}

/* ------------------------------------------------------------------ */
/* drivers: main.cpp:75-79 and simulation.cpp:159-183                  */
/* ------------------------------------------------------------------ */

void orc_preloop(orc_ctx* c) {
    for (int s = 0; s < PIC_MAX_SPECIES; s++)
        if (c->sp[s].declared && !c->sp[s].immobile) orc_push(c, s);
    for (int s = 0; s < PIC_MAX_SPECIES; s++)
        if (c->sp[s].declared && !c->sp[s].immobile) orc_particle_boundary(c, s, NULL);
}

void orc_step(orc_ctx* c, int first_step, int nsteps) {
    for (int st = first_step; st < first_step + nsteps; st++) {
        orc_laser_boundary(c, st);
        orc_fields_first(c);
        orc_clear_currents(c);
        for (int s = 0; s < PIC_MAX_SPECIES; s++)
            if (c->sp[s].declared && !c->sp[s].immobile) orc_push(c, s);
        for (int s = 0; s < PIC_MAX_SPECIES; s++)
            if (c->sp[s].declared && !c->sp[s].immobile) orc_particle_boundary(c, s, NULL);
        orc_fold_currents(c);
        orc_fields_second(c);
    }
}

/* ------------------------------------------------------------------ */
/* diagnostics (new: the reference has none)                           */
/* ------------------------------------------------------------------ */

void orc_diag(orc_ctx* c, pic_diag_out* o) {
    const int m = c->m;
    const double dV = c->p.dx * c->p.dy * c->p.dz;
    double emax = o->spectrum_emax > 0.0 ? o->spectrum_emax : 1.0;
    memset(o, 0, sizeof(*o));
    o->spectrum_emax = emax;
    /* Gauss residual R = div E - (rho_prev + rho)/2 - rho_bg (time-centred:
     * E's half steps straddle the push, simulation.cpp:166-182); mode 0: no
     * rho_prev yet, 1: record R as the drift reference, 2: drift R - R_ref */
    const int mode = !c->rho_prev_valid ? 0 : (c->gauss_ref_valid ? 2 : 1);
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        const int margin = c->p.periodic[a] ? 0 : m + 2; /* inner region */
        lo[a] = m + margin;
        hi[a] = m + c->n[a] - 1 - margin;
    }
    double ee = 0.0, be = 0.0, g2 = 0.0, gmax = 0.0;
    double d2 = 0.0, dmax = 0.0, i2 = 0.0, imax = 0.0, ref2 = 0.0;
    double chg[PIC_MAX_SPECIES] = {0};
    for (int k = m; k < m + c->n[2]; k++)
        for (int j = m; j < m + c->n[1]; j++)
            for (int i = m; i < m + c->n[0]; i++) {
                int64_t f = FI(c, k, j, i), q = JI(c, k, j, i);
                ee += c->F[EX][f] * c->F[EX][f] + c->F[EY][f] * c->F[EY][f] + c->F[EZ][f] * c->F[EZ][f];
                be += c->F[BX][f] * c->F[BX][f] + c->F[BY][f] * c->F[BY][f] + c->F[BZ][f] * c->F[BZ][f];
                double div = (c->F[EX][f] - c->F[EX][FI(c, k, j, i - 1)]) / c->p.dx +
                             (c->F[EY][f] - c->F[EY][FI(c, k, j - 1, i)]) / c->p.dy +
                             (c->F[EZ][f] - c->F[EZ][FI(c, k - 1, j, i)]) / c->p.dz;
                double now = 0.0;
                for (int s = 0; s < PIC_MAX_SPECIES; s++) {
                    double r = c->rho[s][q];
                    chg[s] += r;
                    if (c->sp[s].declared && !c->sp[s].immobile) now += r;
                }
                double prev = mode ? c->rho_prev[q] : now;
                double res = div - 0.5 * (prev + now) - c->rho_bg;
                g2 += res * res;
                if (fabs(res) > gmax) gmax = fabs(res);
                if (mode == 1) {
                    c->gauss_ref[q] = res;
                    ref2 += res * res;
                } else if (mode == 2) {
                    double dr = res - c->gauss_ref[q];
                    d2 += dr * dr;
                    if (fabs(dr) > dmax) dmax = fabs(dr);
                    if (i >= lo[0] && i <= hi[0] && j >= lo[1] && j <= hi[1] && k >= lo[2] && k <= hi[2]) {
                        i2 += dr * dr;
                        if (fabs(dr) > imax) imax = fabs(dr);
                    }
                    ref2 += c->gauss_ref[q] * c->gauss_ref[q];
                }
            }
    if (mode == 1) c->gauss_ref_valid = 1;
    o->gauss_drift_l2 = sqrt(d2);
    o->gauss_drift_max = dmax;
    o->gauss_inner_drift_l2 = sqrt(i2);
    o->gauss_inner_drift_max = imax;
    o->gauss_ref_l2 = sqrt(ref2);
    o->gauss_state = mode;
    o->e_energy = 0.5 * ee * dV;
    o->b_energy = 0.5 * be * dV;
    o->field_energy = o->e_energy + o->b_energy;
    o->gauss_l2 = sqrt(g2);
    o->gauss_max = gmax;
    for (int s = 0; s < PIC_MAX_SPECIES; s++) {
        o->charge[s] = chg[s];
        const orc_species* S = &c->sp[s];
        double ke = 0.0;
        for (int64_t p = 0; p < S->n; p++) {
            double u2 = S->ux[p] * S->ux[p] + S->uy[p] * S->uy[p] + S->uz[p] * S->uz[p];
            double gm1 = u2 / (sqrt(1.0 + u2) + 1.0); /* gamma - 1 without cancellation */
            ke += S->m * S->w[p] * gm1;
            if (s == 0) {
                int b = (int)(gm1 / emax * PIC_SPECTRUM_BINS);
                if (b >= 0 && b < PIC_SPECTRUM_BINS) o->spectrum[b] += S->w[p];
            }
        }
        o->kinetic[s] = ke * dV;
    }
    o->out_of_bounds = c->oob;
}

/* ------------------------------------------------------------------ */
/* loader (same generator as pic3d_amd/csrc/pic_loader_impl.h)          */
/* ------------------------------------------------------------------ */

static int64_t* loader_prefix(const pic_params* p, const pic_loader* l) {
    int64_t* pre = (int64_t*)malloc((size_t)(p->nx + 1) * sizeof(int64_t));
    pre[0] = 0;
    for (int ix = 0; ix < p->nx; ix++)
        pre[ix + 1] = pre[ix] + pic_profile_count(l->profile, l->ppc, l->density, l->x0, l->x1,
                                                  l->scale_len, l->cutoff, ix);
    return pre;
}

int64_t orc_loader_count(const pic_params* p, const pic_loader* l) {
    int64_t* pre = loader_prefix(p, l);
    int64_t n = pre[p->nx] * (int64_t)p->ny * p->nz;
    free(pre);
    return n;
}

void orc_loader(const pic_params* p, const pic_loader* l, int64_t first, int64_t n, double* x,
                double* y, double* z, double* ux, double* uy, double* uz, double* w,
                int64_t* id) {
    int64_t* pre = loader_prefix(p, l);
    const double wt = l->density / l->ppc;
    for (int64_t q = 0; q < n; q++) {
        int64_t pidx = first + q;
        int32_t ix, iy, iz;
        pic_locate(pidx, pre, p->nx, p->ny, &ix, &iy, &iz);
        uint64_t pid = (uint64_t)(l->id_offset + pidx);
        pic_load_one(l->seed, pid, ix, iy, iz, l->u_th, &x[q], &y[q], &z[q], &ux[q], &uy[q], &uz[q]);
        w[q] = wt;
        id[q] = (int64_t)pid;
    }
    free(pre);
}
