/*
 * pic_oracle.h -- CPU oracle for the PIC3D hot loop.  TEST INFRASTRUCTURE
 * ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg, never by the product path.
 *
 * A clean-room FP64 restatement of the reference's per-step loop on one
 * domain, keeping the reference's floating-point operation order so it can
 * be pinned bit-for-bit against the reference itself (oracle/ref_harness).
 * Each function cites the reference file:line it restates.  Optional OpenMP
 * (nthreads > 1) splits particles over threads with private current buffers
 * reduced in thread order; nthreads == 1 is the parity mode.
 */
#ifndef PIC_ORACLE_H
#define PIC_ORACLE_H

#include "../include/pic3d_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(const pic_params* p, int nthreads);
void orc_destroy(orc_ctx* c);
int orc_species_set(orc_ctx* c, int s, double charge, double mass, int immobile);
void orc_set_background(orc_ctx* c, double rho_bg);
int orc_particles_set(orc_ctx* c, int s, int64_t n, const double* x, const double* y,
                      const double* z, const double* ux, const double* uy, const double* uz,
                      const double* w, const int64_t* id);
int64_t orc_num_particles(const orc_ctx* c, int s);
/* Copies species s in oracle order; any pointer may be NULL. */
void orc_particles_get(const orc_ctx* c, int s, double* x, double* y, double* z, double* ux,
                       double* uy, double* uz, double* w, int64_t* id, double* x_old,
                       double* y_old, double* z_old);

void orc_laser_boundary(orc_ctx* c, int step);  /* boundary.cpp:42-99      */
void orc_fields_first(orc_ctx* c);              /* field.cpp:56-86         */
void orc_fields_second(orc_ctx* c);             /* field.cpp:88-117        */
void orc_clear_currents(orc_ctx* c);            /* particle.cpp:576-592    */
void orc_push(orc_ctx* c, int s);               /* particle.cpp:113-316,372-515 */
void orc_particle_boundary(orc_ctx* c, int s, pic_counts* out); /* mpi_utils.cpp:62-230 */
void orc_fold_currents(orc_ctx* c);             /* mpi_utils.cpp:232-417   */
void orc_field_wrap(orc_ctx* c, int comp);      /* mpi_utils.cpp:9-60      */
void orc_preloop(orc_ctx* c);                   /* main.cpp:75-79          */
void orc_radiating_boundary(orc_ctx* c, int step, int is_left, int is_right); /* boundary.cpp:42-99 */
/* 1: place the x-max Silver-Mueller write at the last dual column n+m-1
 * (PIC_OPT_XMAX_BOUNDARY); 0 (default): column local_xcells+1 as the
 * reference (boundary.cpp:92-96). */
void orc_set_xmax_face(orc_ctx* c, int on);
void orc_fields_full(orc_ctx* c);               /* field.cpp:24-54         */
void orc_fold_one(orc_ctx* c, int comp);        /* mpi_utils.cpp:232-417, one array */
void orc_filter_currents(orc_ctx* c);           /* particle.cpp:520-574    */
int64_t orc_deposit(orc_ctx* c, int s, int64_t n, const double* xo, const double* yo,
                    const double* zo, const double* xn, const double* yn, const double* zn,
                    const double* w);           /* particle.cpp:372-515    */
void orc_step(orc_ctx* c, int first_step, int nsteps); /* simulation.cpp:159-183 */

void orc_field_get(const orc_ctx* c, int comp, double* out);
void orc_field_set(orc_ctx* c, int comp, const double* in);
/* FP64 diagnostics (pic_diag_out).  The first call after a current clear
 * since the state was last set records the Gauss residual's reference. */
void orc_diag(orc_ctx* c, pic_diag_out* out);
int64_t orc_out_of_bounds(const orc_ctx* c);

/* Individual shape functions (particle.cpp:8-48), for closed-form tests. */
/* initialize_particles (particle.cpp:50-111) with plasma_density's box */
int64_t orc_lattice(const pic_params* p, const pic_lattice* L, int64_t cap, double* x, double* y,
                    double* z, double* ux, double* uy, double* uz, double* w, int64_t* id);

/* spin precession, particle.cpp:318-368 (species of mass m_e) */
int orc_spin_setup(orc_ctx* c, const double* g, int n, double e_schwinger, double m_e);
void orc_spin_set(orc_ctx* c, int s, const double* sx, const double* sy, const double* sz);
void orc_spin_get(const orc_ctx* c, int s, double* sx, double* sy, double* sz);
int64_t orc_spin_errors(const orc_ctx* c, int s);
double orc_g_factor(const orc_ctx* c, double x, int* bad); /* particle.cpp:603-637 */

void orc_shape2(double x, double* w3);
void orc_shape4(double x, double* w5);
/* Gpulse (boundary.cpp:29-40) for the context's laser parameters. */
double orc_gpulse(const orc_ctx* c, double t, double x, double y, double z);
/* Host loader restated from pic3d_amd/csrc/pic_loader_impl.h. */
int64_t orc_loader_count(const pic_params* p, const pic_loader* l);
void orc_loader(const pic_params* p, const pic_loader* l, int64_t first, int64_t n, double* x,
                double* y, double* z, double* ux, double* uy, double* uz, double* w,
                int64_t* id);

#ifdef __cplusplus
}
#endif
#endif
