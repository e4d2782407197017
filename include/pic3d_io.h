/* pic3d_io.h -- HDF5 output of a pic_ctx in the reference's file format
 * (io_utils.cpp:11-347), written by one process (serial HDF5) for the
 * whole single-GPU domain.
 *
 * Files, as the reference names and lays them out:
 *   <dir>/output_step_%04d.h5     datasets /rho /rho_electron /rho_proton
 *                                 /ex /ey /ez /bx /by /bz /jx /jy /jz,
 *                                 float64 [nz][ny][nx] over the physical
 *                                 cells (io_utils.cpp:27, 53-56, 133-144)
 *   <dir>/particles_step_%04d.h5  dataset /particles: the 18-member
 *                                 compound of include/particle.h (144-byte
 *                                 records, io_utils.cpp:218-236), species
 *                                 in order; int attributes timestep,
 *                                 total_particles, num_processes (= 1)
 *                                 (io_utils.cpp:310-329).  No dataset when
 *                                 there are no particles (io_utils.cpp:200).
 * /rho is the sum of /rho_electron and /rho_proton.  (The reference writes
 * its rho_local array there, which no reference code computes.)
 *
 * Device data is packed on the GPU and transferred into page-locked
 * staging buffers (pic_field_interior, pic_particle_records); the transfer
 * of the next dataset or particle chunk overlaps the HDF5 write of the
 * current one.  Every function returns 0 or a negative pic_err code;
 * pic_io_last_error() describes the last HDF5-side failure. */
#ifndef PIC3D_IO_H
#define PIC3D_IO_H

#include <stdint.h>

#include "pic3d_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* output_data_hdf5 (io_utils.cpp:11-153) for a context: the field file of
 * 'step' under 'dir' (created when missing). */
int pic_output_fields_h5(pic_ctx* ctx, int step, const char* dir);

/* output_particles_hdf5 (io_utils.cpp:157-347) for a context: the live
 * particles of every mobile species, staged chunk_slots slots at a time
 * (<= 0: 16M); *total (may be NULL) = particles written. */
int pic_output_particles_h5(pic_ctx* ctx, int step, const char* dir, int64_t chunk_slots,
                            int64_t* total);

/* The particle file from host records (144-byte include/particle.h
 * records), as io_utils.cpp:157-347 writes a process's vectors. */
int pic_h5_write_records(const char* path, const void* records, int64_t n, int step,
                         int num_processes);

/* The field file from host arrays: 'arrays' holds 12 pointers to float64
 * [nz][ny][nx] data in the order rho, rho_electron, rho_proton, ex, ey, ez,
 * bx, by, bz, jx, jy, jz. */
int pic_h5_write_fields(const char* path, int nx, int ny, int nz, const double* const* arrays);

/* Message of the last failure in this library (thread-local). */
const char* pic_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PIC3D_IO_H */
