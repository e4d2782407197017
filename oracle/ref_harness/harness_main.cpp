// harness_main.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference's own hot-path functions (src/field.cpp,
// src/particle.cpp, src/boundary.cpp, src/mpi_utils.cpp,
// src/simulation.cpp:initialize_domain, all compiled unmodified) on one MPI
// rank with a 1x1x1 Cartesian topology, on particles read from a file, and
// writes the resulting fields and particles.  Used only to produce golden
// vectors that pin the CPU oracle.
//
//   pic_ref_harness IN OUT CMD...
//   CMD: pre                 pre-loop push + exchange (main.cpp:75-79)
//        steps:A:N           steps A..A+N-1 of run_main_time_loop
//                            (simulation.cpp:159-183) without output
//        laser:S first second clear push:K bnd:K fold   single phases
//        rad:S:L:R           radiating_boundary(rank, L, R, 1, S)
//        full                update_electromagnetic_fields
//        filter              binomial_filter_currents
//        xfield:C            exchange_boundary_data(ex|ey|ez|bx|by|bz [C])
//        xcur:C              exchange_boundary_current_fixed(jx|jy|jz|rho0|rho1 [C])
//        dep:K               currents_particle(p, K) for every particle of kind K
//        gfac:PATH           g_factor(x) for the doubles x in PATH, into PATH.out
//        init:PX:PY:PZ:SX:SY:SZ  initialize_particles(kind 0, plasma_density, ...)
//                            in place of the kind-0 input particles
//
// With PIC_REF_SPIN=1 in the environment the input carries sx sy sz (n
// doubles each) after the ids of each species, and the output sx sy sz
// gamma after the ids.
//
// MPI_PROC_NULL is -1 under MPICH but the reference tests for OpenMPI's -2
// (boundary.cpp:87,92; mpi_utils.cpp:16-50,80-222).  On an open x axis the
// harness therefore presents -2 to radiating_boundary (so the laser and
// Silver-Mueller boundary run as under OpenMPI) and MPI_PROC_NULL
// everywhere else (where both values lead to the same no-op or removal).
#include <mpi.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "boundary.h"
#include "constants.h"
#include "field.h"
#include "mpi_utils.h"
#include "particle.h"
#include "simulation.h"

static MPI_Comm g_comm;
static int g_rank = 0;
static int g_mobile[2] = {1, 1};

static void zero3(double*** a, int lo, int hiz, int hiy, int hix) {
    for (int k = lo; k <= hiz; k++)
        for (int j = lo; j <= hiy; j++)
            for (int i = lo; i <= hix; i++) a[k][j][i] = 0.0;
}

static void laser(int step, int left = 1, int right = 0) {
    int xm = rank_xminus, xp = rank_xplus;
    if (!rank_period[0]) rank_xminus = rank_xplus = -2;
    radiating_boundary(g_rank, left, right, 1, step);
    rank_xminus = xm;
    rank_xplus = xp;
}

static void push_all(std::vector<std::vector<particle>>& P, int* n) {
    for (int s = 0; s < 2; s++)
        if (g_mobile[s]) update_particle_position_momentum(P[s], n[s], g_rank, s);
}
static void bnd_all(std::vector<std::vector<particle>>& P, int* n) {
    for (int s = 0; s < 2; s++)
        if (g_mobile[s]) exchange_particle_data(P[s], n[s], g_rank, g_comm);
}

static void write_field(FILE* f, double*** a, int lo, int hiz, int hiy, int hix) {
    for (int k = lo; k <= hiz; k++)
        for (int j = lo; j <= hiy; j++) fwrite(&a[k][j][lo], sizeof(double), hix - lo + 1, f);
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s IN OUT CMD...\n", argv[0]);
        return 2;
    }
    MPI_Init(&argc, &argv);
    int dims[3] = {1, 1, 1};
    int periods[3] = {rank_period[0], rank_period[1], rank_period[2]};
    MPI_Cart_create(MPI_COMM_WORLD, 3, dims, periods, 0, &g_comm);
    MPI_Comm_rank(g_comm, &g_rank);
    MPI_Cart_coords(g_comm, g_rank, 3, rank_coords);
    MPI_Cart_shift(g_comm, 0, +1, &rank_xminus, &rank_xplus);
    MPI_Cart_shift(g_comm, 1, +1, &rank_yminus, &rank_yplus);
    MPI_Cart_shift(g_comm, 2, +1, &rank_zminus, &rank_zplus);
    initialize_domain(g_rank);
    load_g_data();

    // initialize_domain allocates with new[] and no value-initialisation.
    zero3(ex, nc_xst_out, nc_zend_out, nc_yend_out, nc_xend_out);
    zero3(ey, nc_xst_out, nc_zend_out, nc_yend_out, nc_xend_out);
    zero3(ez, nc_xst_out, nc_zend_out, nc_yend_out, nc_xend_out);
    zero3(bx, nc_xst_out, nc_zend_out, nc_yend_out, nc_xend_out);
    zero3(by, nc_xst_out, nc_zend_out, nc_yend_out, nc_xend_out);
    zero3(bz, nc_xst_out, nc_zend_out, nc_yend_out, nc_xend_out);
    clear_currents();

    // input: per species int32 mobile, double q, double m, int64 n, then
    // x y z ux uy uz w (n doubles each) and id (n int64)
    const char* spin_env = std::getenv("PIC_REF_SPIN");
    const bool spin_io = spin_env && std::atoi(spin_env) == 1;
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 3;
    std::vector<std::vector<particle>> P(2);
    int n[2] = {0, 0};
    for (int s = 0; s < 2; s++) {
        int32_t mobile;
        double q, m;
        int64_t cnt;
        if (std::fread(&mobile, 4, 1, in) != 1 || std::fread(&q, 8, 1, in) != 1 ||
            std::fread(&m, 8, 1, in) != 1 || std::fread(&cnt, 8, 1, in) != 1)
            return 4;
        g_mobile[s] = mobile;
        std::vector<double> a[7];
        for (int f = 0; f < 7; f++) {
            a[f].resize(cnt);
            if (cnt && std::fread(a[f].data(), 8, cnt, in) != (size_t)cnt) return 4;
        }
        std::vector<int64_t> id(cnt);
        if (cnt && std::fread(id.data(), 8, cnt, in) != (size_t)cnt) return 4;
        std::vector<double> sp[3];
        for (int f = 0; f < 3 && spin_io; f++) {
            sp[f].resize(cnt);
            if (cnt && std::fread(sp[f].data(), 8, cnt, in) != (size_t)cnt) return 4;
        }
        P[s].resize(cnt);
        for (int64_t p = 0; p < cnt; p++) {
            particle& q_ = P[s][p];
            std::memset(&q_, 0, sizeof(q_));
            q_.id = (int)id[p];
            q_.x = q_.x_old = a[0][p];
            q_.y = q_.y_old = a[1][p];
            q_.z = q_.z_old = a[2][p];
            q_.px = a[3][p];
            q_.py = a[4][p];
            q_.pz = a[5][p];
            q_.weight = a[6][p];
            q_.charge = q;
            q_.mass = m;
            q_.rank = g_rank;
            if (spin_io) {
                q_.sx = sp[0][p];
                q_.sy = sp[1][p];
                q_.sz = sp[2][p];
            }
        }
        n[s] = (int)cnt;
    }
    std::fclose(in);

    for (int c = 3; c < argc; c++) {
        std::string cmd = argv[c];
        if (cmd.rfind("gfac:", 0) != 0) continue;
        // g_factor on its own (particle.cpp:603-637): values in range only
        const std::string path = cmd.substr(5);
        FILE* gi = std::fopen(path.c_str(), "rb");
        if (!gi) return 7;
        std::vector<double> xs;
        double v;
        while (std::fread(&v, 8, 1, gi) == 1) xs.push_back(v);
        std::fclose(gi);
        FILE* go = std::fopen((path + ".out").c_str(), "wb");
        if (!go) return 7;
        for (double x : xs) {
            const double g = g_factor(x);
            std::fwrite(&g, 8, 1, go);
        }
        std::fclose(go);
    }

    for (int c = 3; c < argc; c++) {
        std::string cmd = argv[c];
        if (cmd == "pre") {
            push_all(P, n);
            bnd_all(P, n);
        } else if (cmd.rfind("steps:", 0) == 0) {
            int a = 0, cnt = 0;
            std::sscanf(cmd.c_str(), "steps:%d:%d", &a, &cnt);
            for (int st = a; st < a + cnt; st++) {
                laser(st);
                update_electromagnetic_fields_first(g_rank, 1, g_comm);
                clear_currents();
                push_all(P, n);
                bnd_all(P, n);
                exchange_currents(g_comm);
                update_electromagnetic_fields_second(g_rank, 1, g_comm);
            }
        } else if (cmd.rfind("laser:", 0) == 0) {
            laser(std::atoi(cmd.c_str() + 6));
        } else if (cmd == "first") {
            update_electromagnetic_fields_first(g_rank, 1, g_comm);
        } else if (cmd == "second") {
            update_electromagnetic_fields_second(g_rank, 1, g_comm);
        } else if (cmd == "clear") {
            clear_currents();
        } else if (cmd.rfind("push:", 0) == 0) {
            int s = std::atoi(cmd.c_str() + 5);
            update_particle_position_momentum(P[s], n[s], g_rank, s);
        } else if (cmd.rfind("bnd:", 0) == 0) {
            int s = std::atoi(cmd.c_str() + 4);
            exchange_particle_data(P[s], n[s], g_rank, g_comm);
        } else if (cmd == "fold") {
            exchange_currents(g_comm);
        } else if (cmd.rfind("rad:", 0) == 0) {
            int st = 0, l = 0, r = 0;
            std::sscanf(cmd.c_str(), "rad:%d:%d:%d", &st, &l, &r);
            laser(st, l, r);
        } else if (cmd == "full") {
            update_electromagnetic_fields(g_rank, 1, g_comm);
        } else if (cmd == "filter") {
            binomial_filter_currents();
        } else if (cmd.rfind("xfield:", 0) == 0) {
            double*** a[6] = {ex, ey, ez, bx, by, bz};
            exchange_boundary_data(a[std::atoi(cmd.c_str() + 7)], g_comm);
        } else if (cmd.rfind("xcur:", 0) == 0) {
            double*** a[5] = {jx, jy, jz, rho_particle[0], rho_particle[1]};
            exchange_boundary_current_fixed(a[std::atoi(cmd.c_str() + 5)], g_comm);
        } else if (cmd.rfind("dep:", 0) == 0) {
            int s = std::atoi(cmd.c_str() + 4);
            for (int p = 0; p < n[s]; p++) currents_particle(P[s][p], s);
        } else if (cmd.rfind("gfac:", 0) == 0) {
            // done above
        } else if (cmd.rfind("init:", 0) == 0) {
            double v[6] = {0, 0, 0, 0, 0, 0};
            std::sscanf(cmd.c_str(), "init:%lf:%lf:%lf:%lf:%lf:%lf", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5]);
            const double q = P[0].empty() ? -1.0 : P[0][0].charge, m = P[0].empty() ? 1.0 : P[0][0].mass;
            initialize_particles(P[0], n[0], plasma_density, q, m, v[0], v[1], v[2], v[3], v[4], v[5], g_rank);
        } else {
            std::fprintf(stderr, "harness: unknown command %s\n", cmd.c_str());
            return 5;
        }
    }

    // output: 6 E/B arrays (E/B extents), jx jy jz rho0 rho1 (J extents),
    // then per species int64 n and x y z x_old y_old z_old px py pz weight
    // (n doubles each) and id (n int64), in vector order.
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 6;
    double*** eb[6] = {ex, ey, ez, bx, by, bz};
    for (int f = 0; f < 6; f++) write_field(out, eb[f], nc_xst_out, nc_zend_out, nc_yend_out, nc_xend_out);
    double*** jr[5] = {jx, jy, jz, rho_particle[0], rho_particle[1]};
    for (int f = 0; f < 5; f++)
        write_field(out, jr[f], nc_xst_out - 1, nc_zend_out + 1, nc_yend_out + 1, nc_xend_out + 1);
    for (int s = 0; s < 2; s++) {
        int64_t cnt = n[s];
        std::fwrite(&cnt, 8, 1, out);
        for (int f = 0; f < 10; f++)
            for (int64_t p = 0; p < cnt; p++) {
                const particle& q_ = P[s][p];
                double v[10] = {q_.x, q_.y, q_.z, q_.x_old, q_.y_old, q_.z_old,
                                q_.px, q_.py, q_.pz, q_.weight};
                std::fwrite(&v[f], 8, 1, out);
            }
        for (int64_t p = 0; p < cnt; p++) {
            int64_t id = P[s][p].id;
            std::fwrite(&id, 8, 1, out);
        }
        for (int f = 0; f < 4 && spin_io; f++)
            for (int64_t p = 0; p < cnt; p++) {
                const particle& q_ = P[s][p];
                const double v[4] = {q_.sx, q_.sy, q_.sz, q_.gamma};
                std::fwrite(&v[f], 8, 1, out);
            }
    }
    std::fclose(out);
    MPI_Finalize();
    return 0;
}
