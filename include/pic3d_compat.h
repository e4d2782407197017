/*
 * pic3d_compat.h -- the reference's C++ surface (vvoidpointer/PIC3D
 * include/simulation.h, field.h, particle.h, boundary.h, mpi_utils.h and the
 * parameter externs of constants.h), implemented by libpic3d_compat.so on
 * top of the C ABI in pic3d_amd.h.
 *
 * Drop-in use: build the reference's main.cpp, src/constants.cpp and
 * src/io_utils.cpp unchanged and link them against libpic3d_compat.so in
 * place of src/field.cpp, src/particle.cpp, src/boundary.cpp,
 * src/mpi_utils.cpp and src/simulation.cpp.  The declarations below repeat
 * the reference's signatures exactly so both headers can be used in one
 * program; struct particle has the reference's 144-byte layout.
 *
 * One process drives one GPU with one domain: rank_len must be {1,1,1}.
 * MPI_Comm arguments are accepted and unused.  The device is the home of
 * the fields and, inside run_main_time_loop, of the particles; the host
 * arrays (ex .. bz, jx .. jz, rho_particle) are mirrors brought up to date
 * after every per-call function that changes them and before every output
 * call of run_main_time_loop.  Environment: PIC3D_DEVICE (default 0),
 * PIC3D_PRECISION ("fp64" default, or "fp32").
 */
#ifndef PIC3D_COMPAT_H
#define PIC3D_COMPAT_H

#include <mpi.h>
#include <stdint.h>

#include <vector>

#if defined(__has_include)
#if __has_include(<hdf5.h>)
#include <hdf5.h>
#define PIC3D_COMPAT_HAVE_HDF5 1
#endif
#endif
#ifndef PIC3D_COMPAT_HAVE_HDF5
typedef int64_t hid_t; /* HDF5 >= 1.10 defines hid_t as int64_t */
#endif

#include "pic3d_amd.h"

/* include/particle.h:9-29 */
struct particle {
    int id;
    double x;
    double y;
    double z;
    double x_old;
    double y_old;
    double z_old;
    double px;
    double py;
    double pz;
    double sx;
    double sy;
    double sz;
    double gamma;
    double charge;
    double mass;
    double weight;
    int rank;
};

/* ---- include/constants.h:16-99: defined by the caller's constants.cpp ---- */
extern const double Pi, c1, q_e, m_e, omega, epsilon_0, l0, t0, Lx, Ly, Lz;
extern const int resx, resy, resz, rest;
extern const int rank_len[3];
extern const int rank_period[3];
extern int rank_reorder;
extern int rank_coords[3];
extern int rank_xminus, rank_xplus, rank_yminus, rank_yplus, rank_zminus, rank_zplus;
extern const int local_xcells, local_ycells, local_zcells;
extern const int global_nx, global_ny, global_nz;
extern const int cell_per_particles;
extern const double density, dx, dy, dz, dt, dt_half, simulation_time;
extern const int timesteps, output_interval, num_steps;
extern bool is_periodic;
extern const int mem;
extern const double laser_amplitude, laser_pulse_duration, laser_start_time;
extern const int laser_polarization;
extern const double waist, xfocus, phase, yfocus, zfocus;
extern int nc_xst, nc_xend, nc_xst_new, nc_xend_new, nc_xst_out, nc_xend_out;
extern double xst, xend;
extern int nc_yst, nc_yend, nc_yst_new, nc_yend_new, nc_yst_out, nc_yend_out;
extern double yst, yend;
extern int nc_zst, nc_zend, nc_zst_new, nc_zend_new, nc_zst_out, nc_zend_out;
extern double zst, zend;
extern double E_schwinger;
extern const double alpha0, alpha1, alpha2, alpha3, beta0, beta2;
extern const int num_kind;
extern const int g_data_size;
extern double g_data[];

/* ---- include/field.h:11-31 ---- */
extern double*** rho_local;
extern double**** rho_particle;
extern double ***ex, ***phi, ***phi_new, ***ey, ***ez, ***bx, ***by, ***bz, ***jx, ***jy, ***jz;
extern double ***current_recv_xminus, ***current_recv_xplus, ***current_recv_yminus,
    ***current_recv_yplus, ***current_recv_zminus, ***current_recv_zplus;
extern double max_error;
void update_electromagnetic_fields(int my_rank, int num_procs, MPI_Comm my_comm);
void update_electromagnetic_fields_first(int my_rank, int num_procs, MPI_Comm my_comm);
void update_electromagnetic_fields_second(int my_rank, int num_procs, MPI_Comm my_comm);

/* ---- include/particle.h:31-47 ---- */
void initialize_particles(std::vector<particle>& particles, int& num_particles,
                          double density_func(double, double, double), double charge,
                          double mass, double initial_px, double initial_py, double initial_pz,
                          double initial_sx, double initial_sy, double initial_sz, int my_rank);
void update_particle_position_momentum(std::vector<particle>& particles, int num_particles,
                                       int my_rank, int kind);
void currents_particle(const particle& particles, int kind);
void binomial_filter_currents();
void shape_function_2(double x, double* Interpolation);
void shape_function_4(double x, double* Interpolation);
void clear_currents();
void exchange_currents(MPI_Comm my_comm);
double g_factor(double x);

/* ---- include/boundary.h:4-6 ---- */
double generate_laser_pulse(double time, double position);
double Gpulse(double time, double x, double y, double z);
void radiating_boundary(int my_rank, int is_left, int is_right, int num_procs, int step);

/* ---- include/mpi_utils.h:13-16 ---- */
void exchange_boundary_data(double*** data, MPI_Comm my_comm);
void exchange_particle_data(std::vector<particle>& particles, int& num_particles, int my_rank,
                            MPI_Comm my_comm);
void exchange_boundary_current_fixed(double*** data, MPI_Comm my_comm);

/* ---- include/simulation.h:10-23 ---- */
void initialize_simulation(int my_rank, int num_procs);
void initialize_domain(int my_rank);
void run_main_time_loop(std::vector<std::vector<particle>>& particles, int* num_particles,
                        int my_rank, int num_procs, hid_t plist_id, MPI_Comm my_comm);
double** dmatrix(long nrl, long nrh, long ncl, long nch);
void free_dmatrix(double** m, long nrl, long ncl);
void free_dmatrix(double** m, long nrl);
double*** dtensor3(long nzl, long nzh, long nyl, long nyh, long nxl, long nxh);
double**** dtensor4(long n4l, long n4h, long n3l, long n3h, long n2l, long n2h, long n1l,
                    long n1h);
void free_dtensor3(double*** m, long nzl, long nyl, long nxl);
void free_dtensor4(double**** m, long n4l, long n3l, long n2l, long n1l);
void free_memory();

/* ---- called by run_main_time_loop (include/io_utils.h).  Defined by the
 * caller's io_utils.cpp, or by libpic3d_compat_io.so: the reference's files
 * written with serial HDF5 through libpic3d_io (fields staged from the
 * device, particles from the vectors) ---- */
void output_data_hdf5(int step, int my_rank, int num_procs, MPI_Comm my_comm, hid_t plist_id);
void output_particles_hdf5(int step, int my_rank, int num_procs, MPI_Comm my_comm,
                           const std::vector<std::vector<particle>>& particles,
                           const int* num_particles);

/* ---- additions of this library (no counterpart in the reference) ---- */
namespace pic3d_compat {
/* The context behind the globals (created by initialize_domain). */
pic_ctx* context();
/* Copy the device fields into the host mirrors / the mirrors to the device. */
void pull_fields();
void push_fields();
}  // namespace pic3d_compat

#endif /* PIC3D_COMPAT_H */
