/*
 * pic3d_amd.h -- C ABI of the MI355X-native PIC3D hot loop.
 *
 * One context = one single-domain simulation resident on one GPU: the six
 * Yee field components, the current/charge accumulators and the particle
 * species, all in HBM.  Every entry point returns an int status
 * (PIC_OK == 0, negative on error; the message is in pic_last_error()).
 * No exception crosses this boundary.  A context owns one HIP stream; calls
 * on one context are not re-entrant.  All calls are asynchronous on that
 * stream except *_get, pic_diag, pic_num_particles, pic_particle_counts and
 * pic_sync, which synchronise.
 *
 * Coordinates follow the reference exactly (include/particle.h:9-29,
 * src/simulation.cpp:23-63 of vvoidpointer/PIC3D):
 *   - particle positions are in GLOBAL CELL UNITS, x in [0, nx];
 *   - field arrays are [z][y][x] with x fastest; E and B have extents
 *     [0, n+2*mem-1] per axis, J and rho have [-1, n+2*mem] (stored from 0,
 *     i.e. shifted by +1);
 *   - a particle at global x sits at buffer coordinate x - (1 - mem);
 *   - Ex, Jx, By, Bz live on the dual (i+1/2) x grid, etc. (standard Yee).
 *
 * Parity-critical functions cite the reference file:line they replace.
 */
#ifndef PIC3D_AMD_H
#define PIC3D_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIC_OK 0
#define PIC_ERR_ARG (-1)       /* bad argument or parameter set            */
#define PIC_ERR_HIP (-2)       /* a HIP runtime call failed                 */
#define PIC_ERR_NOMEM (-3)     /* device allocation failed                  */
#define PIC_ERR_STATE (-4)     /* call not valid in the context's state     */
#define PIC_ERR_CAPACITY (-5)  /* particle count exceeds species capacity   */

#define PIC_MAX_SPECIES 4

enum pic_precision {
    PIC_FP64 = 0, /* reference arithmetic: doubles everywhere                  */
    PIC_FP32 = 1  /* fp32 fields/momenta/deposit, int cell + fp32 offset       */
};

/* Field component ids for pic_field_get/pic_field_set. */
enum pic_field {
    PIC_EX = 0, PIC_EY, PIC_EZ, PIC_BX, PIC_BY, PIC_BZ, /* E/B extents       */
    PIC_JX, PIC_JY, PIC_JZ,                             /* J extents (+2)    */
    PIC_RHO0, PIC_RHO1, PIC_RHO2, PIC_RHO3              /* per-species rho   */
};

/*
 * The constants.cpp parameter set (src/constants.cpp:6-85,
 * include/constants.h:16-99) as a runtime struct for one single domain.
 * pic_params_default() fills the reference's shipped values.
 */
typedef struct pic_params {
    int32_t nx, ny, nz;        /* global cells per axis (global_nx..nz)         */
    int32_t mem;               /* 2: quadratic shapes, 4: quartic; = ghost width */
    int32_t periodic[3];       /* rank_period: 1 periodic, 0 open/absorbing      */
    int32_t precision;         /* enum pic_precision                             */
    double dx, dy, dz, dt;     /* grid spacing and time step (c/omega units)     */
    double pi;                 /* Pi as the reference truncates it (3.1415926536)*/
    double omega;              /* laser angular frequency (1 in the reference)   */
    double l0;                 /* laser wavelength 2*Pi*c/omega                  */
    double lx;                 /* Lx, used only for the x-max laser source       */
    /* laser at x-min, boundary.cpp:8-99 */
    int32_t laser_polarization;  /* 1 S, 2 P, 3 both; 0 no source              */
    int32_t laser_right;         /* is_right of radiating_boundary (ref: 0)    */
    double laser_amplitude;      /* a0                                          */
    double laser_pulse_duration; /* T of the sin(pi t/T) envelope               */
    double laser_start_time;
    double waist, xfocus, yfocus, zfocus, phase;
} pic_params;

/* Per-species boundary accounting from one pic_particle_boundary call
 * (src/mpi_utils.cpp:62-230 on a single domain).  Face order:
 * x-, x+, y-, y+, z-, z+.  "lost" = removed at an open face; "wrapped" =
 * crossed a periodic face (re-entered, wrapped at the next push). */
typedef struct pic_counts {
    int64_t lost[6];
    int64_t wrapped[6];
    int64_t remaining;
} pic_counts;

/* FP64 diagnostics: energies, charge, spectrum and the Gauss-law residual
 * (no counterpart in the reference, which has no diagnostics).
 *
 * Gauss law.  In the reference's step order (simulation.cpp:166-182) E's
 * two half steps straddle the push, the first with the previous push's J
 * and the second with the new one, and Esirkepov's deposit satisfies
 * rho^{n+1} - rho^n = -dt div J^{n+1/2} (particle.cpp:372-515).  So the
 * time-centred residual at every interior node
 *     R = div E - (rho_prev + rho) / 2 - rho_background
 * stays constant from step to step (to rounding) on periodic axes, where
 *   - div E uses backward differences on the E stagger,
 *   - rho is the sum of the mobile species' charge arrays now (rho^{n+1}),
 *   - rho_prev is the same sum as it stood when the currents were last
 *     cleared (rho^n; rho itself when no clear happened since the
 *     particles, fields, species or background were last set).
 * R changes at open faces: their ghost E is never updated and the charge
 * of particles that leave there is dropped.  It also changes at periodic
 * faces between steps 1 and 2, because the pre-loop push deposits without
 * a fold (main.cpp:75-79): step 1's rho_prev is that unfolded charge.
 * The first pic_diag call made after a current clear (since the state was
 * last set) records R as the reference R_ref; later calls report R - R_ref,
 * the drift.  The inner region leaves out the nodes within mem + 2 cells
 * of an open face. */
#define PIC_SPECTRUM_BINS 64
typedef struct pic_diag_out {
    double field_energy;       /* 0.5 * sum_interior (E^2 + B^2) * dV          */
    double e_energy, b_energy; /* the two halves of field_energy              */
    double kinetic[PIC_MAX_SPECIES]; /* sum_p m * w * (gamma - 1) * dV          */
    double charge[PIC_MAX_SPECIES];  /* sum of rho over interior nodes         */
    double gauss_l2;           /* ||R||_2 over the interior nodes              */
    double gauss_max;          /* max |R|                                      */
    double spectrum_emax;      /* upper edge of the spectrum, in m c^2 units  */
    double spectrum[PIC_SPECTRUM_BINS]; /* species-0 weight histogram of gamma-1 */
    int64_t out_of_bounds;     /* particles outside the domain at a push       */
    double gauss_drift_l2;     /* ||R - R_ref||_2 over the interior nodes      */
    double gauss_drift_max;    /* max |R - R_ref|                              */
    double gauss_inner_drift_l2;  /* the same over the inner region            */
    double gauss_inner_drift_max;
    double gauss_ref_l2;       /* ||R_ref||_2                                  */
    int32_t gauss_state;       /* 0: rho_prev unknown (R uses rho for it), no
                                  drift; 1: this call recorded R_ref (drift 0);
                                  2: drift against R_ref                       */
    int32_t reserved;
} pic_diag_out;

/* Seeded initial-state generator.  Identical streams on host and device:
 * a splitmix64 hash of (seed, particle id, draw index). */
enum pic_profile {
    PIC_PROFILE_UNIFORM = 0,     /* density everywhere                         */
    PIC_PROFILE_SLAB = 1,        /* density for x in [x0, x1)                  */
    PIC_PROFILE_RAMP_TARGET = 2  /* exp ramp to x0, flat target on [x0, x1)    */
};
typedef struct pic_loader {
    int32_t profile;     /* enum pic_profile                                  */
    int32_t ppc;         /* particles per cell at density                     */
    double density;      /* n in n_c at full density (weight = density/ppc)   */
    double x0, x1;       /* slab or target extent, global cells               */
    double scale_len;    /* ramp e-folding length, cells                       */
    double cutoff;       /* ramp: no particles below this density             */
    double u_th;         /* Maxwellian: each momentum component ~ N(0, u_th^2) */
    uint64_t seed;
    int64_t id_offset;   /* id of the first particle                           */
} pic_loader;

/* ---- parameter set and loader (host only, no GPU needed) ---- */

/* Reference defaults (src/constants.cpp:6-65) for an nx x ny x nz domain. */
void pic_params_default(pic_params* p, int nx, int ny, int nz);
/* Validate a parameter set; returns PIC_OK or PIC_ERR_ARG and, when msg is
 * not NULL, writes a reason into msg[0..msg_len). */
int pic_params_check(const pic_params* p, char* msg, int msg_len);
/* Number of particles the loader produces for this domain. */
int64_t pic_loader_count(const pic_params* p, const pic_loader* l);
/* Generate particles [first, first+n) of the loader on the host. */
int pic_loader_host(const pic_params* p, const pic_loader* l, int64_t first, int64_t n,
                    double* x, double* y, double* z, double* ux, double* uy,
                    double* uz, double* w, int64_t* id);
/* Library version string. */
const char* pic_version(void);

/* ---- context ---- */

typedef struct pic_ctx pic_ctx;

/* Allocates the fields (initialize_domain, src/simulation.cpp:132-151) on
 * device 'device' and zeroes them. */
int pic_ctx_create(const pic_params* p, int device, pic_ctx** out);
int pic_ctx_destroy(pic_ctx* ctx);
/* The parameter set the context was created with. */
int pic_ctx_params(const pic_ctx* ctx, pic_params* out);
const char* pic_last_error(const pic_ctx* ctx);
int pic_sync(pic_ctx* ctx);

/* Tuning / debugging options (no counterpart in the reference). */
enum pic_option {
    PIC_OPT_PUSH_KERNEL = 1,   /* 0 auto, 1 direct (global atomics), 2 tiled (LDS) */
    PIC_OPT_KEEP_OLD = 2,      /* 1: keep x_old/y_old/z_old per particle (FP64)    */
    PIC_OPT_SORT_INTERVAL = 3, /* re-sort particles by tile every N pushes (>= 1; default 1) */
    PIC_OPT_TIMING = 4,        /* 1: time every phase with HIP events              */
    PIC_OPT_FP32_ACCUM = 5,    /* FP32 only: 1 (default) int32 fixed-point deposit
                                  accumulators in LDS (each add rounded to 2^-19 of
                                  one particle's largest contribution, 2^-18 in
                                  the quadratic push's 8192-particle chunks); 0 FP64
                                  accumulators in LDS (smaller tiles, slower) */
    PIC_OPT_ADJ_PUSH = 6,      /* differentiated runs: 0 (default) the LDS-tiled
                                  adjoint push (HBM kernel after a direct forward);
                                  1 dual-number kernel (FP64; FP32 takes 2);
                                  2 separable HBM kernel */
    PIC_OPT_XMAX_BOUNDARY = 7, /* x-max Silver-Mueller write (x open):
                                  0 (default) the reference's: column
                                  local_xcells+1 from column local_xcells
                                  (boundary.cpp:92-96), three columns inside
                                  the domain at mem 4, and unstable once a
                                  wave reaches it; 1 the mirror of the x-min
                                  write: the last dual column n+mem-1 from
                                  E, Bx and J at the primal column n+mem-1
                                  and By, Bz at dual column n+mem-2 */
    PIC_OPT_CELL_ORDER = 8     /* FP64, and FP32 quartic with fixed-point
                                  accumulators: N >= 1 keeps each tile's
                                  particles in cell order (every N-th sort bins
                                  by cell, the tile sorts in between keep the
                                  order; default 4), and the tiled push gathers
                                  with the lanes of a wave in neighbouring cells
                                  and deposits with its lanes spread over cells
                                  (two kernels: gather, Boris and move, then the
                                  deposit from the per-slot records the first
                                  wrote: FP64 old positions, FP32 old cell and
                                  offsets);
                                  0: tile bins (chunks stored transposed) and
                                  one pass with scattered lanes.  Results are
                                  the same up to the order of the fixed-point
                                  sums, which is exact. */
};
int pic_set_option(pic_ctx* ctx, int option, int64_t value);

/* Phase categories reported by pic_timing_get. */
enum pic_timing_cat {
    PIC_T_FIELDS = 0,   /* Yee half steps + halo wrap + laser boundary          */
    PIC_T_CLEAR = 1,    /* current/charge clear                                 */
    PIC_T_PUSH = 2,     /* push kernel (gather + Boris + deposit)               */
    PIC_T_SORT = 3,     /* tile sort + work list                                 */
    PIC_T_BOUNDARY = 4, /* particle boundary                                     */
    PIC_T_FOLD = 5,     /* current ghost fold                                    */
    PIC_TIMING_CATS = 6
};
/* Accumulated milliseconds and launch counts per category since the last
 * reset (synchronises); reset != 0 clears the accumulators afterwards. */
int pic_timing_get(pic_ctx* ctx, double* ms, int64_t* count, int reset);

/* Declare species s (charge q, mass m in reference units; an immobile
 * species is never pushed nor deposited and acts as a uniform background
 * of charge density rho_background).  capacity = max particles kept. */
int pic_species_set(pic_ctx* ctx, int s, double charge, double mass, int immobile,
                    int64_t capacity);
int pic_set_background(pic_ctx* ctx, double rho_background);

/* Host SoA -> device (positions in global cell units, u = p/(m c)). */
int pic_particles_set(pic_ctx* ctx, int s, int64_t n, const double* x, const double* y,
                      const double* z, const double* ux, const double* uy,
                      const double* uz, const double* w, const int64_t* id);
/* Device -> host; x_old may be NULL.  *n in: array capacity, out: count.
 * Order is the device order (compare particle sets by id). */
int pic_particles_get(pic_ctx* ctx, int s, int64_t* n, double* x, double* y, double* z,
                      double* ux, double* uy, double* uz, double* w, int64_t* id,
                      double* x_old, double* y_old, double* z_old);
/* Run the seeded loader on the device (no host staging). */
int pic_particles_load(pic_ctx* ctx, int s, const pic_loader* l);

/* The reference's own initial loading: initialize_particles
 * (particle.cpp:50-111) with the shape of its density function
 * plasma_density (constants.cpp:87-103).  Particles sit at the centres of
 * a lattice of (int)(n * cbrt(cell_per_particles)) points per axis, in
 * k, j, i order; a point (x, y, z) (global cell units) gets weight
 * density / cell_per_particles when lo[a]*l0 < x_a * d_a <= hi[a]*l0 on
 * every axis (the reference: 5 and 15) and is skipped otherwise; at most
 * n_x n_y n_z cell_per_particles particles.  Every particle starts with
 * momentum u and (species with spin) spin s; ids are id_offset + index.
 * Runs on the device; FP64 positions are the reference's bit for bit. */
typedef struct pic_lattice {
    int32_t cell_per_particles;
    int32_t reserved;
    double density;
    double lo[3], hi[3];   /* density box per axis, in units of l0          */
    double u[3];           /* initial_px, initial_py, initial_pz            */
    double spin[3];        /* initial_sx, initial_sy, initial_sz            */
    int64_t id_offset;     /* my_rank * num_particles in the reference      */
} pic_lattice;
int64_t pic_lattice_count(const pic_params* p, const pic_lattice* l);
int pic_particles_lattice(pic_ctx* ctx, int s, const pic_lattice* l);
int pic_num_particles(pic_ctx* ctx, int s, int64_t* n);

/* ---- spin precession (particle.cpp:318-368) ---- */

/* Turn on the spin precession that update_particle_position_momentum
 * applies to electrons (particle.cpp:318-368, for particles whose mass
 * equals m_e) for every species of mass electron_mass, declared now or
 * later.  g_table: the n values of g.txt that load_g_data reads
 * (constants.cpp:114-136; n = 2849 in the reference), indexed as g_factor
 * does (particle.cpp:603-637); e_schwinger: E_schwinger (constants.cpp:83).
 * Spins start at 0 (pic_spin_set / pic_spin_fill give them values) and
 * follow their particles through sorts and boundaries.  Off by default:
 * the spin feeds back into nothing else.  Not available inside a
 * differentiated run. */
int pic_spin_setup(pic_ctx* ctx, const double* g_table, int n, double e_schwinger,
                   double electron_mass);
/* Spins of the n live particles of species s, in the order
 * pic_particles_get returns them (host -> device, device -> host). */
int pic_spin_set(pic_ctx* ctx, int s, int64_t n, const double* sx, const double* sy,
                 const double* sz);
int pic_spin_get(pic_ctx* ctx, int s, int64_t* n, double* sx, double* sy, double* sz);
/* The same spin for every particle of species s (the initial_sx..sz of
 * initialize_particles, particle.cpp:50-111). */
int pic_spin_fill(pic_ctx* ctx, int s, double sx, double sy, double sz);
/* Precession steps skipped so far because chi fell outside g.txt's range
 * [0, 100] or was NaN (the reference prints an error and exits there,
 * particle.cpp:632-636); the spin is left unchanged.  Synchronises. */
int pic_spin_errors(pic_ctx* ctx, int s, int64_t* n);

/* ---- per-phase hot path (one call = one reference call) ---- */

/* radiating_boundary(rank, 1, laser_right, nprocs, step), boundary.cpp:42-99 */
int pic_laser_boundary(pic_ctx* ctx, int step);
/* update_electromagnetic_fields_first, field.cpp:56-86 (+ halo wrap) */
int pic_fields_first(pic_ctx* ctx);
/* update_electromagnetic_fields_second, field.cpp:88-117 (+ halo wrap) */
int pic_fields_second(pic_ctx* ctx);
/* clear_currents, particle.cpp:576-592 */
int pic_clear_currents(pic_ctx* ctx);
/* update_particle_position_momentum(particles[s], n, rank, s) without the
 * spin part, particle.cpp:113-316 + currents_particle 372-515 */
int pic_push(pic_ctx* ctx, int s);
/* exchange_particle_data on one domain, mpi_utils.cpp:62-230; counts may be
 * NULL (then nothing synchronises). */
int pic_particle_boundary(pic_ctx* ctx, int s, pic_counts* counts);
/* exchange_currents on one domain, particle.cpp:594-601 and
 * mpi_utils.cpp:232-417 */
int pic_fold_currents(pic_ctx* ctx);
/* The pre-loop push of main.cpp:75-79: push + boundary of every species with
 * no current clear and no fold. */
int pic_preloop(pic_ctx* ctx);
/* Steps first_step .. first_step+nsteps-1 of run_main_time_loop
 * (simulation.cpp:159-183) without output. */
int pic_step(pic_ctx* ctx, int first_step, int nsteps);

/* ---- the rest of the reference's per-call surface ---- */

/* radiating_boundary(rank, is_left, is_right, nprocs, step),
 * boundary.cpp:42-99: is_left / is_right switch the x-min / x-max laser
 * sources; the Silver-Mueller writes happen whenever x is open. */
int pic_radiating_boundary(pic_ctx* ctx, int step, int is_left, int is_right);
/* update_electromagnetic_fields, field.cpp:24-54: E then B advanced a full
 * dt, each followed by its halo exchange. */
int pic_fields_full(pic_ctx* ctx);
/* exchange_boundary_data(data), mpi_utils.cpp:9-60, for one E/B
 * component (PIC_EX..PIC_BZ). */
int pic_field_exchange(pic_ctx* ctx, int comp);
/* exchange_boundary_current_fixed(data), mpi_utils.cpp:232-417, for one
 * J or rho array (PIC_JX..PIC_JZ, PIC_RHO0..PIC_RHO3). */
int pic_current_exchange(pic_ctx* ctx, int comp);
/* binomial_filter_currents, particle.cpp:520-574: one 1-2-1 pass per axis
 * (X, Y, Z) over the interior of jx, jy, jz. */
int pic_filter_currents(pic_ctx* ctx);
/* currents_particle(p, s), particle.cpp:372-515, for n particles given on
 * the host by old and new positions and weight (charge from species s):
 * Esirkepov current and charge deposit into J and rho[s].  FP64 contexts
 * only.  Particles whose old position is outside [0, n] on an axis, or
 * that moved a cell or more, are skipped and counted in *skipped (may be
 * NULL). Synchronises. */
int pic_deposit(pic_ctx* ctx, int s, int64_t n, const double* x_old, const double* y_old,
                const double* z_old, const double* x, const double* y, const double* z,
                const double* w, int64_t* skipped);
/* No counterpart in the reference.  pic_deposit for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same deposit, with the
 * caller's arrays staged in a fixed per-GPU device array instead of a
 * buffer allocated per call.  It calls no hipMalloc/hipFree.  At most
 * PIC_NOALLOC_MAX_PARTICLES particles per call (more: PIC_ERR_CAPACITY).
 * Every call on one GPU shares that GPU's staging array: calls that
 * overlap on one GPU race.  Synchronises. */
#define PIC_NOALLOC_MAX_PARTICLES 65536
int pic_deposit_noalloc(pic_ctx* ctx, int s, int64_t n, const double* x_old,
                        const double* y_old, const double* z_old, const double* x,
                        const double* y, const double* z, const double* w, int64_t* skipped);

/* ---- differentiation form (no counterpart in the reference) ----
 *
 * Gradients of the final kinetic energy of one species,
 *   K = sum_p m w_p (gamma_p - 1) dx dy dz   (pic_diag_out.kinetic[s]),
 * after nsteps steps of the time loop, with respect to
 *   - the x-min laser envelope value of every step, e_n (the value
 *     generate_laser_pulse(t_n, 0) returns, boundary.cpp:8-27; the source
 *     is Gpulse with e_n in its place), and
 *   - the initial weight of every particle slot (the density profile).
 * One mobile species (others immobile), in the context's precision: an
 * FP32 context runs the FP32 forward and FP32 adjoint kernels (fields,
 * momenta, offsets and their adjoints in fp32; the objective, the envelope
 * gradient and the weight gradients accumulated in FP64).  A tiled run
 * re-sorts the species by tile every PIC_OPT_SORT_INTERVAL pushes with a
 * sort that keeps removed slots (particles leaving through open faces are
 * marked dead, so the slot count never changes) and records each sort's
 * permutation; the recompute replays the recorded permutations, so it
 * puts every particle in the slot the forward did, and the adjoint carries
 * the particle adjoints back through them.  The forward keeps a
 * full-state checkpoint every K steps; the backward recomputes each
 * K-step segment from its checkpoint, saving the per-step push inputs,
 * and runs the adjoint kernels of that segment in reverse.  Memory: one
 * checkpoint per K steps (fields, currents, particles), K push-input
 * states and one int32 permutation per sort.  In FP32 the deposit's
 * fixed-point unit follows the largest weight: pic_diff_forward with new
 * weights sizes it for them. */

/* Snapshot the current state of the context as the initial state of a
 * differentiated run of species s. */
int pic_diff_begin(pic_ctx* ctx, int s);
/* Number of particle slots (entries of the weight arrays below). */
int pic_diff_slots(pic_ctx* ctx, int64_t* slots);
/* Initial positions and weights of the slots (NaN weight: a dead slot);
 * any pointer may be NULL.  Synchronises. */
int pic_diff_initial(pic_ctx* ctx, double* x, double* y, double* z, double* w);
/* From the initial state (with 'weights' in place of the initial weights
 * when not NULL; dead slots stay dead) run steps first_step ..
 * first_step+nsteps-1 with the x-min envelope 'envelope' (nsteps values;
 * NULL = the reference's sin envelope), checkpointing every K steps, and
 * return K in *objective.  Synchronises. */
int pic_diff_forward(pic_ctx* ctx, int first_step, int nsteps, int K, const double* envelope,
                     const double* weights, double* objective);
/* Reverse of the last pic_diff_forward: grad_objective * dK/de_n into
 * grad_envelope (nsteps, may be NULL) and grad_objective * dK/dw into
 * grad_weights (slots, may be NULL).  Leaves the context at the run's
 * initial state.  Synchronises. */
int pic_diff_backward(pic_ctx* ctx, double grad_objective, double* grad_envelope,
                      double* grad_weights);
/* Release the differentiated run's buffers. */
int pic_diff_end(pic_ctx* ctx);

/* ---- readback / diagnostics ---- */

/* Full ghosted array of component comp, as doubles in the reference layout. */
int pic_field_get(pic_ctx* ctx, int comp, double* host);
int pic_field_set(pic_ctx* ctx, int comp, const double* host);
/* Number of doubles pic_field_get writes for comp. */
int64_t pic_field_size(const pic_params* p, int comp);
int pic_diag(pic_ctx* ctx, pic_diag_out* out);
/* Cumulative boundary counts of species s since the last pic_particles_set. */
int pic_particle_counts(pic_ctx* ctx, int s, pic_counts* out);

/* ---- output staging (io_utils.cpp:11-347 write these; libpic3d_io holds
 * the HDF5 side, include/pic3d_io.h) ---- */

/* Page-locked host memory (hipHostMalloc) for staging buffers that the
 * device transfers into directly; free with pic_pinned_free. */
int pic_pinned_alloc(size_t bytes, void** host);
int pic_pinned_free(void* host);
/* Physical cells of component comp as doubles, [nz][ny][nx] -- the buffer
 * io_utils.cpp:84-97 (fields) and 108-119 (charge densities) fill before
 * each H5Dwrite.  Packed on the device (FP32 widened), then one transfer
 * into 'host' (direct DMA when host is pinned).  Synchronises. */
int pic_field_interior(pic_ctx* ctx, int comp, double* host);
/* Particle slots of species s (live and removed-but-not-compacted). */
int pic_particle_slots(pic_ctx* ctx, int s, int64_t* slots);
/* The live particles among slots [first_slot, first_slot + nslots) as the
 * reference's 144-byte particle records (include/particle.h:9-29, the
 * records io_utils.cpp:218-236 writes), in slot order, into 'records'
 * (room for nslots); *nrec = records written.  id is the low 32 bits of
 * the id; x..z in global cell units; x_old..z_old as kept with
 * PIC_OPT_KEEP_OLD (FP64), else the position; px..pz = u; sx..sz the spin
 * (0 for a species without precession, pic_spin_setup); gamma = sqrt(1 + u.u) (particle.cpp:309); charge
 * and mass of the species; weight; rank 0.  Synchronises. */
int pic_particle_records(pic_ctx* ctx, int s, int64_t first_slot, int64_t nslots, void* records,
                         int64_t* nrec);

/* ---- timing helpers for measurement (HIP events on the context stream) ---- */
int pic_timer_start(pic_ctx* ctx);
/* Milliseconds since pic_timer_start, after synchronising the stream. */
int pic_timer_stop(pic_ctx* ctx, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* PIC3D_AMD_H */
