// pic3d_compat_io.cpp -- the reference's io_utils.h functions
// (io_utils.cpp:11-347) for programs on the C++ drop-in
// (libpic3d_compat.so), written with serial HDF5 through libpic3d_io.
//
// output_data_hdf5 stages the fields straight from the device through the
// drop-in's context (the host mirrors are not read); output_particles_hdf5
// writes the caller's particle vectors, species in order, as the reference
// does.  Files: output/output_step_%04d.h5 and output/particles_step_%04d.h5
// under the working directory (io_utils.cpp:14-25, 160-173).  A failed
// write ends the process with a message, as H5_SAFE_CALL's MPI_Abort does
// (io_utils.h:9-14).
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pic3d_compat.h"
#include "pic3d_io.h"

namespace {
[[noreturn]] void h5_fail(int my_rank, const char* what, const char* why) {
    std::fprintf(stderr, "Rank %d: HDF5 error in %s: %s\n", my_rank, what, why);
    std::exit(1);
}
}  // namespace

void output_data_hdf5(int step, int my_rank, int num_procs, MPI_Comm my_comm, hid_t plist_id) {
    (void)num_procs; (void)my_comm; (void)plist_id;
    if (pic_output_fields_h5(pic3d_compat::context(), step, "output") != PIC_OK)
        h5_fail(my_rank, "output_data_hdf5", pic_io_last_error());
}

void output_particles_hdf5(int step, int my_rank, int num_procs, MPI_Comm my_comm,
                           const std::vector<std::vector<particle>>& particles,
                           const int* num_particles) {
    (void)my_comm;
    if (mkdir("output", 0755) != 0 && errno != EEXIST)  // io_utils.cpp:160-168
        h5_fail(my_rank, "output_particles_hdf5", "cannot create output/");
    std::vector<particle> all;  // every kind in order (io_utils.cpp:257-271)
    for (int s = 0; s < num_kind && s < (int)particles.size(); s++)
        for (int i = 0; i < num_particles[s]; i++) all.push_back(particles[s][i]);
    char path[256];
    std::snprintf(path, sizeof(path), "output/particles_step_%04d.h5", step);
    if (pic_h5_write_records(path, all.data(), (int64_t)all.size(), step, num_procs) != PIC_OK)
        h5_fail(my_rank, "output_particles_hdf5", pic_io_last_error());
    if (my_rank == 0 && !all.empty())  // io_utils.cpp:341-345
        std::printf("Particle data written to %s (total particles: %d)\n", path, (int)all.size());
}
