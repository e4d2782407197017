// pic3d_compat.cpp -- the reference's C++ surface on top of the C ABI.
//
// Each function keeps the reference's name, signature and meaning (file:line
// cited per function) and does its work on the GPU through libpic3d_amd:
// fields live on the device, the host arrays are mirrors.  Fatal errors end
// the process with a message, as the reference's exit()/MPI_Abort paths do.
#include "pic3d_compat.h"

#include <algorithm>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

// ---- field.h globals (field.cpp:8-22) ----
double max_error = 1.0;
double*** rho_local = nullptr;
double**** rho_particle = nullptr;
double ***ex = nullptr, ***phi = nullptr, ***phi_new = nullptr, ***ey = nullptr, ***ez = nullptr;
double ***bx = nullptr, ***by = nullptr, ***bz = nullptr;
double ***jx = nullptr, ***jy = nullptr, ***jz = nullptr;
double ***current_recv_xminus = nullptr, ***current_recv_xplus = nullptr,
       ***current_recv_yminus = nullptr, ***current_recv_yplus = nullptr,
       ***current_recv_zminus = nullptr, ***current_recv_zplus = nullptr;

namespace {

pic_ctx* g_ctx = nullptr;
pic_params g_p;
int g_nspecies = 0;          // one past the highest species declared on the device
bool g_declared[PIC_MAX_SPECIES];
double g_q[PIC_MAX_SPECIES], g_m[PIC_MAX_SPECIES];
int64_t g_cap[PIC_MAX_SPECIES];
std::map<const void*, int> g_kind_of;  // particle vector -> kind (from the push calls)
bool g_spin_ready = false;            // pic_spin_setup done with the caller's g_data

[[noreturn]] void fail(const char* what, int rc) {
    std::fprintf(stderr, "pic3d_compat: %s failed (%d): %s\n", what, rc,
                 g_ctx ? pic_last_error(g_ctx) : "no context");
    std::exit(1);
}
void chk(int rc, const char* what) {
    if (rc != PIC_OK) fail(what, rc);
}
pic_ctx* ctx() {
    if (!g_ctx) {
        std::fprintf(stderr, "pic3d_compat: initialize_domain() has not been called\n");
        std::exit(1);
    }
    return g_ctx;
}

// First element of a mirror (the arrays are one contiguous block in the
// device's [z][y][x] order, so a mirror transfer is one copy).
double* field_base(double*** a) { return &a[nc_zst_out][nc_yst_out][nc_xst_out]; }
double* current_base(double*** a) { return &a[nc_zst_out - 1][nc_yst_out - 1][nc_xst_out - 1]; }

void pull(int comp) {
    double*** a = nullptr;
    bool cur = comp >= PIC_JX;
    switch (comp) {
        case PIC_EX: a = ex; break;
        case PIC_EY: a = ey; break;
        case PIC_EZ: a = ez; break;
        case PIC_BX: a = bx; break;
        case PIC_BY: a = by; break;
        case PIC_BZ: a = bz; break;
        case PIC_JX: a = jx; break;
        case PIC_JY: a = jy; break;
        case PIC_JZ: a = jz; break;
        default:
            if (comp - PIC_RHO0 < num_kind) a = rho_particle[comp - PIC_RHO0];
    }
    if (!a) return;
    chk(pic_field_get(ctx(), comp, cur ? current_base(a) : field_base(a)), "pic_field_get");
}
void pull_currents() {
    for (int c = PIC_JX; c <= PIC_JZ; c++) pull(c);
    for (int s = 0; s < num_kind && s < PIC_MAX_SPECIES; s++) pull(PIC_RHO0 + s);
}
void pull_eb() {
    for (int c = PIC_EX; c <= PIC_BZ; c++) pull(c);
}

// Which device component a host mirror is.
int comp_of(double*** a) {
    double*** eb[6] = {ex, ey, ez, bx, by, bz};
    for (int c = 0; c < 6; c++)
        if (a == eb[c]) return PIC_EX + c;
    if (a == jx) return PIC_JX;
    if (a == jy) return PIC_JY;
    if (a == jz) return PIC_JZ;
    for (int s = 0; s < num_kind && s < PIC_MAX_SPECIES; s++)
        if (a == rho_particle[s]) return PIC_RHO0 + s;
    return -1;
}

// Declare (or grow) species 'kind' for n particles of charge q and mass m.
void ensure_species(int kind, double q, double m, int64_t n) {
    if (kind < 0 || kind >= PIC_MAX_SPECIES) {
        std::fprintf(stderr, "pic3d_compat: particle kind %d outside [0, %d)\n", kind, PIC_MAX_SPECIES);
        std::exit(1);
    }
    const bool same = g_declared[kind] && g_q[kind] == q && g_m[kind] == m;
    if (same && g_cap[kind] >= n) return;
    int64_t cap = same ? g_cap[kind] : 0;
    while (cap < n) cap = cap ? cap * 2 : std::max<int64_t>(n, 1024);
    // charge arrays of a re-declared species keep their contents
    chk(pic_species_set(ctx(), kind, q, m, 0, cap), "pic_species_set");
    g_q[kind] = q;
    g_m[kind] = m;
    g_cap[kind] = cap;
    g_declared[kind] = true;
    if (kind >= g_nspecies) g_nspecies = kind + 1;
}

// The reference precesses the spin of every particle of mass m_e
// (particle.cpp:318-368) with the g_data table main() loads from g.txt
// (constants.cpp:114-136) and E_schwinger: hand them to the device once,
// at the first push (after main's load_g_data).
void ensure_spin() {
    if (g_spin_ready) return;
    chk(pic_spin_setup(ctx(), g_data, g_data_size, E_schwinger, m_e), "pic_spin_setup");
    g_spin_ready = true;
}

// g_factor outside its table stops the reference (particle.cpp:632-636);
// the device skips such a precession step and counts it
void check_spin(int kind) {
    if (!g_spin_ready || !g_declared[kind] || g_m[kind] != m_e) return;
    int64_t bad = 0;
    chk(pic_spin_errors(ctx(), kind, &bad), "pic_spin_errors");
    if (bad) {
        std::printf("Error: x value out of g_data range! (%lld precession steps)\n", (long long)bad);
        std::exit(1);
    }
}

// Upload a particle vector as species 'kind' (device id = vector index).
void upload(int kind, const std::vector<particle>& P, int64_t n) {
    ensure_spin();
    if (n == 0) {  // an empty vector: nothing on the device either
        if (g_declared[kind]) {
            double z = 0.0;
            int64_t i0 = 0;
            chk(pic_particles_set(ctx(), kind, 0, &z, &z, &z, &z, &z, &z, &z, &i0), "pic_particles_set");
        }
        return;
    }
    const double q = P[0].charge, m = P[0].mass;
    for (int64_t i = 1; i < n; i++)
        if (P[i].charge != q || P[i].mass != m) {
            std::fprintf(stderr, "pic3d_compat: kind %d mixes charges or masses; one species per "
                                 "kind is required\n", kind);
            std::exit(1);
        }
    ensure_species(kind, q, m, n);
    std::vector<double> a[7];
    for (auto& v : a) v.resize(n);
    std::vector<int64_t> id(n);
    for (int64_t i = 0; i < n; i++) {
        a[0][i] = P[i].x; a[1][i] = P[i].y; a[2][i] = P[i].z;
        a[3][i] = P[i].px; a[4][i] = P[i].py; a[5][i] = P[i].pz;
        a[6][i] = P[i].weight;
        id[i] = i;
    }
    chk(pic_particles_set(ctx(), kind, n, a[0].data(), a[1].data(), a[2].data(), a[3].data(),
                          a[4].data(), a[5].data(), a[6].data(), id.data()),
        "pic_particles_set");
    if (m == m_e) {
        for (int64_t i = 0; i < n; i++) { a[0][i] = P[i].sx; a[1][i] = P[i].sy; a[2][i] = P[i].sz; }
        chk(pic_spin_set(ctx(), kind, n, a[0].data(), a[1].data(), a[2].data()), "pic_spin_set");
    }
}

struct Down {
    std::vector<double> x, y, z, ux, uy, uz, w, xo, yo, zo, sx, sy, sz;
    std::vector<int64_t> id;
    int64_t n = 0;
    bool spin = false;
};
Down download(int kind) {
    Down d;
    int64_t n = 0;
    chk(pic_num_particles(ctx(), kind, &n), "pic_num_particles");
    d.n = n;
    for (auto* v : {&d.x, &d.y, &d.z, &d.ux, &d.uy, &d.uz, &d.w, &d.xo, &d.yo, &d.zo}) v->resize(std::max<int64_t>(n, 1));
    d.id.resize(std::max<int64_t>(n, 1));
    int64_t cap = std::max<int64_t>(n, 1);
    chk(pic_particles_get(ctx(), kind, &cap, d.x.data(), d.y.data(), d.z.data(), d.ux.data(),
                          d.uy.data(), d.uz.data(), d.w.data(), d.id.data(), d.xo.data(),
                          d.yo.data(), d.zo.data()),
        "pic_particles_get");
    d.n = cap;
    if (g_declared[kind] && g_m[kind] == m_e && g_spin_ready) {
        for (auto* v : {&d.sx, &d.sy, &d.sz}) v->resize(std::max<int64_t>(d.n, 1));
        int64_t ns = std::max<int64_t>(d.n, 1);
        chk(pic_spin_get(ctx(), kind, &ns, d.sx.data(), d.sy.data(), d.sz.data()), "pic_spin_get");
        d.spin = true;
    }
    return d;
}

// Write device state j of 'd' into host particle 'q' (the reference's
// update_particle_position_momentum writes x_old, x, p and gamma,
// particle.cpp:297-314, and sets rank, :120).
void store(particle& q, const Down& d, int64_t j, int rank) {
    q.x = d.x[j]; q.y = d.y[j]; q.z = d.z[j];
    q.x_old = d.xo[j]; q.y_old = d.yo[j]; q.z_old = d.zo[j];
    q.px = d.ux[j]; q.py = d.uy[j]; q.pz = d.uz[j];
    q.gamma = std::sqrt(1.0 + (q.px * q.px + q.py * q.py + q.pz * q.pz));
    if (d.spin) { q.sx = d.sx[j]; q.sy = d.sy[j]; q.sz = d.sz[j]; }
    q.rank = rank;
}

int kind_of(const std::vector<particle>& P) {
    auto it = g_kind_of.find(&P);
    if (it != g_kind_of.end()) return it->second;
    // not pushed yet: match charge and mass against the declared species
    if (!P.empty())
        for (int s = 0; s < g_nspecies; s++)
            if (g_declared[s] && g_q[s] == P[0].charge && g_m[s] == P[0].mass) return s;
    return -1;
}

pic_params params_from_constants() {
    pic_params p;
    pic_params_default(&p, global_nx, global_ny, global_nz);
    p.mem = mem;
    for (int a = 0; a < 3; a++) p.periodic[a] = rank_period[a] ? 1 : 0;
    const char* prec = std::getenv("PIC3D_PRECISION");
    p.precision = (prec && std::strcmp(prec, "fp32") == 0) ? PIC_FP32 : PIC_FP64;
    p.dx = dx; p.dy = dy; p.dz = dz; p.dt = dt;
    p.pi = Pi; p.omega = omega; p.l0 = l0; p.lx = Lx;
    p.laser_polarization = laser_polarization;
    p.laser_right = 0;  // run_main_time_loop calls radiating_boundary(rank, 1, 0, ...)
    p.laser_amplitude = laser_amplitude;
    p.laser_pulse_duration = laser_pulse_duration;
    p.laser_start_time = laser_start_time;
    p.waist = waist; p.xfocus = xfocus; p.yfocus = yfocus; p.zfocus = zfocus; p.phase = phase;
    return p;
}

}  // namespace

namespace pic3d_compat {
pic_ctx* context() { return g_ctx; }
void pull_fields() {
    pull_eb();
    pull_currents();
}
void push_fields() {
    double*** eb[6] = {ex, ey, ez, bx, by, bz};
    for (int c = 0; c < 6; c++) chk(pic_field_set(ctx(), PIC_EX + c, field_base(eb[c])), "pic_field_set");
    double*** j3[3] = {jx, jy, jz};
    for (int c = 0; c < 3; c++) chk(pic_field_set(ctx(), PIC_JX + c, current_base(j3[c])), "pic_field_set");
    for (int s = 0; s < g_nspecies && s < num_kind; s++)
        if (g_declared[s])
            chk(pic_field_set(ctx(), PIC_RHO0 + s, current_base(rho_particle[s])), "pic_field_set");
}
}  // namespace pic3d_compat

// =====================================================================
// simulation.h
// =====================================================================

// simulation.cpp:11-21
void initialize_simulation(int my_rank, int num_procs) {
    (void)num_procs;
    if (my_rank == 0) {
        std::printf("Starting PIC simulation:\n");
        std::printf("  Total timesteps: %d\n", timesteps);
        std::printf("  Output interval: every %d steps\n", output_interval);
        std::printf("  Domain: Lx=%.2f, dx=%.6f, dt=%.6f\n", Lx, dx, dt);
        std::printf("  Expected output files: %d\n", (timesteps / output_interval) + 1);
    }
}

// simulation.cpp:23-152: the domain indices of one rank, and the field
// arrays (here: the device context plus host mirrors).
void initialize_domain(int my_rank) {
    (void)my_rank;
    if (rank_len[0] * rank_len[1] * rank_len[2] != 1) {
        std::fprintf(stderr, "pic3d_compat: rank_len = {%d,%d,%d}; this library runs one domain "
                             "on one GPU (rank_len {1,1,1})\n", rank_len[0], rank_len[1], rank_len[2]);
        std::exit(1);
    }
    const int cells[3] = {local_xcells, local_ycells, local_zcells};
    int* st[3] = {&nc_xst, &nc_yst, &nc_zst};
    int* en[3] = {&nc_xend, &nc_yend, &nc_zend};
    int* stn[3] = {&nc_xst_new, &nc_yst_new, &nc_zst_new};
    int* so[3] = {&nc_xst_out, &nc_yst_out, &nc_zst_out};
    int* eo[3] = {&nc_xend_out, &nc_yend_out, &nc_zend_out};
    double* lo[3] = {&xst, &yst, &zst};
    double* hi[3] = {&xend, &yend, &zend};
    for (int a = 0; a < 3; a++) {
        *st[a] = rank_coords[a] * cells[a] + 1;
        *en[a] = *st[a] + cells[a] - 1;
        *lo[a] = (double)(*st[a] - 1);
        *hi[a] = (double)(*en[a]);
        *stn[a] = *st[a] - mem;
        *st[a] = *st[a] - *stn[a];
        *en[a] = *en[a] - *stn[a];
        *so[a] = *st[a] - mem;
        *eo[a] = *en[a] + mem;
    }
    // OpenMPI's MPI_PROC_NULL on open faces, the value the reference tests
    // for (boundary.cpp:87,92; mpi_utils.cpp:80-222); 0 (this rank) on
    // periodic faces.
    rank_xminus = rank_xplus = rank_period[0] ? 0 : -2;
    rank_yminus = rank_yplus = rank_period[1] ? 0 : -2;
    rank_zminus = rank_zplus = rank_period[2] ? 0 : -2;

    g_p = params_from_constants();
    char msg[256];
    if (pic_params_check(&g_p, msg, sizeof(msg)) != PIC_OK) {
        std::fprintf(stderr, "pic3d_compat: parameter set rejected: %s\n", msg);
        std::exit(1);
    }
    const char* dev = std::getenv("PIC3D_DEVICE");
    pic_ctx* c = nullptr;
    int rc = pic_ctx_create(&g_p, dev ? std::atoi(dev) : 0, &c);
    if (rc != PIC_OK) {
        std::fprintf(stderr, "pic3d_compat: pic_ctx_create failed (%d)\n", rc);
        std::exit(1);
    }
    g_ctx = c;
    if (g_p.precision == PIC_FP64)  // x_old of the particle vectors (FP32: kept on the host)
        chk(pic_set_option(g_ctx, PIC_OPT_KEEP_OLD, 1), "pic_set_option");

    ex = dtensor3(nc_zst_out, nc_zend_out, nc_yst_out, nc_yend_out, nc_xst_out, nc_xend_out);
    ey = dtensor3(nc_zst_out, nc_zend_out, nc_yst_out, nc_yend_out, nc_xst_out, nc_xend_out);
    ez = dtensor3(nc_zst_out, nc_zend_out, nc_yst_out, nc_yend_out, nc_xst_out, nc_xend_out);
    bx = dtensor3(nc_zst_out, nc_zend_out, nc_yst_out, nc_yend_out, nc_xst_out, nc_xend_out);
    by = dtensor3(nc_zst_out, nc_zend_out, nc_yst_out, nc_yend_out, nc_xst_out, nc_xend_out);
    bz = dtensor3(nc_zst_out, nc_zend_out, nc_yst_out, nc_yend_out, nc_xst_out, nc_xend_out);
    const long z0 = nc_zst_out - 1, z1 = nc_zend_out + 1, y0 = nc_yst_out - 1, y1 = nc_yend_out + 1,
               x0 = nc_xst_out - 1, x1 = nc_xend_out + 1;
    jx = dtensor3(z0, z1, y0, y1, x0, x1);
    jy = dtensor3(z0, z1, y0, y1, x0, x1);
    jz = dtensor3(z0, z1, y0, y1, x0, x1);
    rho_local = dtensor3(z0, z1, y0, y1, x0, x1);  // written by no hot-path call; zero
    rho_particle = dtensor4(0, num_kind - 1, z0, z1, y0, y1, x0, x1);
    pic3d_compat::pull_fields();
}

// Offset-indexed arrays (simulation.cpp:218-491): one contiguous data
// block per array, row pointers shifted so that m[lo] is the first row.
double** dmatrix(long nrl, long nrh, long ncl, long nch) {
    const long nr = nrh - nrl + 1, nc = nch - ncl + 1;
    double** m = new (std::nothrow) double*[nr];
    double* d = new (std::nothrow) double[(size_t)(nr * nc)]();
    if (!m || !d) throw std::runtime_error("dmatrix: allocation failed");
    for (long r = 0; r < nr; r++) m[r] = d + r * nc - ncl;
    return m - nrl;
}
void free_dmatrix(double** m, long nrl, long ncl) {
    if (!m) return;
    delete[] (m[nrl] + ncl);
    delete[] (m + nrl);
}
void free_dmatrix(double** m, long nrl) { free_dmatrix(m, nrl, 0); }

double*** dtensor3(long nzl, long nzh, long nyl, long nyh, long nxl, long nxh) {
    const long nz = nzh - nzl + 1, ny = nyh - nyl + 1, nx = nxh - nxl + 1;
    double*** m = new (std::nothrow) double**[nz];
    double** rows = new (std::nothrow) double*[(size_t)(nz * ny)];
    double* d = new (std::nothrow) double[(size_t)(nz * ny * nx)]();
    if (!m || !rows || !d) throw std::runtime_error("dtensor3: allocation failed");
    for (long k = 0; k < nz; k++) {
        for (long j = 0; j < ny; j++) rows[k * ny + j] = d + (k * ny + j) * nx - nxl;
        m[k] = rows + k * ny - nyl;
    }
    return m - nzl;
}
void free_dtensor3(double*** m, long nzl, long nyl, long nxl) {
    if (!m) return;
    double** rows = m[nzl] + nyl;
    delete[] (rows[0] + nxl);
    delete[] rows;
    delete[] (m + nzl);
}

double**** dtensor4(long n4l, long n4h, long n3l, long n3h, long n2l, long n2h, long n1l, long n1h) {
    const long n4 = n4h - n4l + 1;
    double**** m = new (std::nothrow) double***[n4];
    if (!m) throw std::runtime_error("dtensor4: allocation failed");
    for (long q = 0; q < n4; q++) m[q] = dtensor3(n3l, n3h, n2l, n2h, n1l, n1h);
    return m - n4l;
}
void free_dtensor4(double**** m, long n4l, long n3l, long n2l, long n1l) {
    if (!m) return;
    for (long q = n4l; q < n4l + num_kind; q++) free_dtensor3(m[q], n3l, n2l, n1l);
    delete[] (m + n4l);
}

// simulation.cpp:493-588
void free_memory() {
    const long z0 = nc_zst_out - 1, y0 = nc_yst_out - 1, x0 = nc_xst_out - 1;
    free_dtensor3(rho_local, z0, y0, x0); rho_local = nullptr;
    free_dtensor4(rho_particle, 0, z0, y0, x0); rho_particle = nullptr;
    double*** eb[6] = {ex, ey, ez, bx, by, bz};
    for (auto* a : eb) free_dtensor3(a, nc_zst_out, nc_yst_out, nc_xst_out);
    ex = ey = ez = bx = by = bz = nullptr;
    free_dtensor3(jx, z0, y0, x0); free_dtensor3(jy, z0, y0, x0); free_dtensor3(jz, z0, y0, x0);
    jx = jy = jz = nullptr;
    if (g_ctx) pic_ctx_destroy(g_ctx);
    g_ctx = nullptr;
    g_nspecies = 0;
    for (bool& d : g_declared) d = false;
    g_kind_of.clear();
}

// simulation.cpp:154-205.  Particles are uploaded once and stay on the
// device; steps run in batches between output steps; before each output
// call the mirrors and the particle vectors are brought up to date.  The
// vectors list particles in device order (grouped by tile), not in the
// reference's append order.
void run_main_time_loop(std::vector<std::vector<particle>>& particles, int* num_particles,
                        int my_rank, int num_procs, hid_t plist_id, MPI_Comm my_comm) {
    const int nk = (int)particles.size();
    // static attributes by device id (= index at upload)
    std::vector<std::vector<particle>> orig(nk);
    for (int k = 0; k < nk; k++) {
        orig[k].assign(particles[k].begin(), particles[k].begin() + num_particles[k]);
        upload(k, particles[k], num_particles[k]);
    }
    auto refresh = [&]() {
        pic3d_compat::pull_fields();
        for (int k = 0; k < nk; k++) {
            if (k >= PIC_MAX_SPECIES || !g_declared[k]) continue;  // never had particles
            Down d = download(k);
            std::vector<particle>& P = particles[k];
            P.resize(d.n);
            if (g_p.precision != PIC_FP64) {  // no x_old in the FP32 form
                d.xo = d.x; d.yo = d.y; d.zo = d.z;
            }
            for (int64_t j = 0; j < d.n; j++) {
                P[j] = orig[k][d.id[j]];
                store(P[j], d, j, my_rank);
            }
            num_particles[k] = (int)d.n;
        }
    };
    int step = 0;
    while (step < timesteps) {
        int next = timesteps;
        if (output_interval > 0) next = std::min(timesteps, (step / output_interval + 1) * output_interval);
        chk(pic_step(ctx(), step + 1, next - step), "pic_step");
        for (int k = 0; k < nk && k < PIC_MAX_SPECIES; k++) check_spin(k);
        step = next;
        if (output_interval > 0 && step % output_interval == 0) {
            refresh();
            output_data_hdf5(step, my_rank, num_procs, my_comm, plist_id);
            output_particles_hdf5(step, my_rank, num_procs, my_comm, particles, num_particles);
        }
    }
    if (timesteps > 0) refresh();  // with no step the vectors are as given
    if (my_rank == 0) {
        std::printf("PIC simulation completed! Total %d timesteps executed.\n", timesteps);
        std::printf("Final output files: output/output_step_0000.h5 to output/output_step_%04d.h5\n",
                    output_interval > 0 ? (timesteps / output_interval) * output_interval : 0);
    }
}

// =====================================================================
// field.h: field.cpp:24-117
// =====================================================================

void update_electromagnetic_fields(int, int, MPI_Comm) {
    chk(pic_fields_full(ctx()), "pic_fields_full");
    pull_eb();
}
void update_electromagnetic_fields_first(int, int, MPI_Comm) {
    chk(pic_fields_first(ctx()), "pic_fields_first");
    pull_eb();
}
void update_electromagnetic_fields_second(int, int, MPI_Comm) {
    chk(pic_fields_second(ctx()), "pic_fields_second");
    pull_eb();
}

// =====================================================================
// particle.h
// =====================================================================

// particle.cpp:8-19
void shape_function_2(double x, double* W) {
    const int ic = (int)(x + 0.5);
    const double d = x - (double)ic;
    W[0] = 0.5 * (0.5 - d) * (0.5 - d);
    W[1] = 0.75 - d * d;
    W[2] = 0.5 * (0.5 + d) * (0.5 + d);
}

// particle.cpp:21-48: quartic B-spline weights at offsets -2..+2 of the
// nearest node, evaluated in the reference's operation order.
void shape_function_4(double x, double* W) {
    const int ic = (int)(x + 0.5);
    const double t = x - (ic - 2);
    double u = t - 2.5, u2 = u * u;
    W[0] = u2 * u2 / 24.0;
    u = t - 1.0; u2 = u * u;
    W[1] = -u2 * u2 / 6.0 + u2 * u * alpha3 - u2 * alpha2 + u * alpha1 + alpha0;
    u = t - 2.0; u2 = u * u;
    W[2] = u2 * u2 / 4.0 - u2 * beta2 + beta0;
    u = t - 3.0; u2 = u * u;
    W[3] = -u2 * u2 / 6.0 - u2 * u * alpha3 - u2 * alpha2 - u * alpha1 + alpha0;
    u = t - 1.5; u2 = u * u;
    W[4] = u2 * u2 / 24.0;
}

// particle.cpp:50-111: particles at the centres of a cbrt(cpp)^3 lattice
// per cell, weight density(x,y,z)/cpp, zero-density points skipped.
void initialize_particles(std::vector<particle>& particles, int& num_particles,
                          double density_func(double, double, double), double charge, double mass,
                          double initial_px, double initial_py, double initial_pz, double initial_sx,
                          double initial_sy, double initial_sz, int my_rank) {
    num_particles = local_xcells * local_ycells * local_zcells * cell_per_particles;
    particles.resize(num_particles);
    const double per_axis = std::cbrt((double)cell_per_particles);
    const int nx = (int)(local_xcells * per_axis), ny = (int)(local_ycells * per_axis),
              nz = (int)(local_zcells * per_axis);
    const double hx = (xend - xst) / nx, hy = (yend - yst) / ny, hz = (zend - zst) / nz;
    int n = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                if (n >= num_particles) break;
                const double px = xst + i * hx + hx / 2.0;
                const double py = yst + j * hy + hy / 2.0;
                const double pz = zst + k * hz + hz / 2.0;
                const double w = density_func(px, py, pz) / cell_per_particles;
                if (w <= 0.0) continue;
                particle& q = particles[n];
                q.id = my_rank * num_particles + n;
                q.x = px; q.y = py; q.z = pz;
                q.px = initial_px; q.py = initial_py; q.pz = initial_pz;
                q.sx = initial_sx; q.sy = initial_sy; q.sz = initial_sz;
                q.gamma = std::sqrt(1 + (q.px * q.px + q.py * q.py + q.pz * q.pz));
                q.charge = charge;
                q.mass = mass;
                q.weight = w;
                q.rank = my_rank;
                n++;
            }
    particles.resize(n);
    num_particles = n;
}

// particle.cpp:113-370: wrap, gather, Boris push, Esirkepov deposit and
// (particles of mass m_e) spin precession of every particle of the vector
// on the GPU.
void update_particle_position_momentum(std::vector<particle>& particles, int num_particles,
                                       int my_rank, int kind) {
    g_kind_of[&particles] = kind;
    if (num_particles <= 0) return;
    upload(kind, particles, num_particles);
    chk(pic_push(ctx(), kind), "pic_push");
    check_spin(kind);
    Down d = download(kind);
    if (g_p.precision != PIC_FP64) {
        // the FP32 form keeps no x_old: it is the start position after the
        // periodic wrap of particle.cpp:120-167
        const int N[3] = {global_nx, global_ny, global_nz};
        for (int64_t j = 0; j < d.n; j++) {
            const particle& q = particles[d.id[j]];
            double o[3] = {q.x, q.y, q.z};
            for (int a = 0; a < 3; a++)
                if (rank_period[a]) {
                    if (o[a] < 0) o[a] += N[a];
                    else if (o[a] > N[a]) o[a] -= N[a];
                }
            d.xo[j] = o[0]; d.yo[j] = o[1]; d.zo[j] = o[2];
        }
    }
    for (int64_t j = 0; j < d.n; j++) store(particles[d.id[j]], d, j, my_rank);
    for (int i = 0; i < num_particles; i++) particles[i].rank = my_rank;
    pull_currents();
}

// particle.cpp:372-515 for one particle (x_old -> x); its charge is the
// species charge of 'kind'.
void currents_particle(const particle& p, int kind) {
    ensure_species(kind, p.charge, p.mass, 1);
    int64_t skipped = 0;
    chk(pic_deposit(ctx(), kind, 1, &p.x_old, &p.y_old, &p.z_old, &p.x, &p.y, &p.z, &p.weight, &skipped),
        "pic_deposit");
    pull(PIC_JX); pull(PIC_JY); pull(PIC_JZ); pull(PIC_RHO0 + kind);
}

// particle.cpp:520-574
void binomial_filter_currents() {
    chk(pic_filter_currents(ctx()), "pic_filter_currents");
    pull(PIC_JX); pull(PIC_JY); pull(PIC_JZ);
}

// particle.cpp:576-592
void clear_currents() {
    chk(pic_clear_currents(ctx()), "pic_clear_currents");
    pull_currents();
}

// particle.cpp:594-601
void exchange_currents(MPI_Comm) {
    chk(pic_fold_currents(ctx()), "pic_fold_currents");
    pull_currents();
}

// particle.cpp:603-637: table lookup in g.txt's values by ranges of x.
double g_factor(double x) {
    int idx;
    if (x >= 0.0 && x < 0.005) return 2.003219;
    if (x >= 0.005 && x < 0.01) idx = (int)((x - 0.005) / 0.0001);
    else if (x >= 0.01 && x < 1) idx = (int)((x - 0.01) / 0.01) + 50;
    else if (x >= 1 && x < 10) idx = (int)((x - 1.0) / 0.01) + 150;
    else if (x >= 10 && x < 20) idx = (int)((x - 10.0) / 0.01) + 1050;
    else if (x >= 20 && x <= 100) idx = (int)((x - 20.0) / 0.1) + 2050;
    else {
        std::printf("Error: x value out of g_data range! x = %.2f\n", x);
        std::exit(1);
    }
    // from x = 99.9 the reference's index passes the end of the table
    return g_data[std::min(idx, g_data_size - 1)];
}

// =====================================================================
// boundary.h: boundary.cpp:8-99
// =====================================================================

// sin envelope on [0, T] in amplitude (boundary.cpp:8-27)
double generate_laser_pulse(double time, double position) {
    const double t = time - position - laser_start_time;
    const double env = (t >= 0.0 && t <= 1.0 * laser_pulse_duration) ? std::sin(Pi * t / laser_pulse_duration) : 0.0;
    return laser_amplitude * env;
}

// Gaussian beam (boundary.cpp:29-40)
double Gpulse(double time, double x, double y, double z) {
    const double k = 2.0 * Pi / l0;
    const double zr = 0.5 * k * waist * waist;
    const double xe = x - xfocus;
    const double gouy = std::atan(xe / zr);
    const double w = waist * std::sqrt(1 + std::pow(xe / zr, 2));
    const double ph = (xe / zr) * (y * y + z * z) / w / w - gouy + k * xe - omega * time;
    return -1.0 * generate_laser_pulse(time, x) * waist / w * std::exp(-(y * y + z * z) / (w * w)) *
           std::cos(ph + phase);
}

void radiating_boundary(int, int is_left, int is_right, int, int step) {
    chk(pic_radiating_boundary(ctx(), step, is_left, is_right), "pic_radiating_boundary");
    pull(PIC_BY); pull(PIC_BZ);
}

// =====================================================================
// mpi_utils.h
// =====================================================================

// mpi_utils.cpp:9-60 for one E/B array
void exchange_boundary_data(double*** data, MPI_Comm) {
    const int c = comp_of(data);
    if (c < PIC_EX || c > PIC_BZ) {
        std::fprintf(stderr, "pic3d_compat: exchange_boundary_data: not a field array\n");
        std::exit(1);
    }
    chk(pic_field_exchange(ctx(), c), "pic_field_exchange");
    pull(c);
}

// mpi_utils.cpp:232-417 for one J or rho array
void exchange_boundary_current_fixed(double*** data, MPI_Comm) {
    const int c = comp_of(data);
    if (c < PIC_JX) {
        std::fprintf(stderr, "pic3d_compat: exchange_boundary_current_fixed: not a current array\n");
        std::exit(1);
    }
    chk(pic_current_exchange(ctx(), c), "pic_current_exchange");
    pull(c);
}

// mpi_utils.cpp:62-230 on one domain.  The device decides which particles
// leave through open faces; the vector is then rebuilt in the reference's
// order: per phase X, Y, Z, particles that stay keep their order and those
// crossing a periodic face are appended, first the ones that crossed the
// high face (received back from the low neighbour, this rank), then the
// low face.  Positions stay unwrapped; the next push wraps them.
void exchange_particle_data(std::vector<particle>& particles, int& num_particles, int my_rank,
                            MPI_Comm) {
    const int kind = kind_of(particles);
    if (kind < 0 || num_particles <= 0) {
        if (num_particles <= 0) return;
        std::fprintf(stderr, "pic3d_compat: exchange_particle_data on a vector that was never pushed\n");
        std::exit(1);
    }
    upload(kind, particles, num_particles);
    pic_counts cnt;
    chk(pic_particle_boundary(ctx(), kind, &cnt), "pic_particle_boundary");
    Down d = download(kind);
    std::vector<char> alive(num_particles, 0);
    for (int64_t j = 0; j < d.n; j++) alive[d.id[j]] = 1;
    std::vector<int> order;
    order.reserve(num_particles);
    for (int i = 0; i < num_particles; i++)
        if (alive[i]) order.push_back(i);
    const double lo[3] = {xst, yst, zst}, hi[3] = {xend, yend, zend};
    for (int a = 0; a < 3; a++) {
        if (!rank_period[a]) continue;  // open faces: the device removed the leavers
        std::vector<int> stay, up, down;
        for (int i : order) {
            const double v = a == 0 ? particles[i].x : (a == 1 ? particles[i].y : particles[i].z);
            if (v < lo[a]) down.push_back(i);
            else if (v >= hi[a]) up.push_back(i);
            else stay.push_back(i);
        }
        stay.insert(stay.end(), up.begin(), up.end());
        stay.insert(stay.end(), down.begin(), down.end());
        order.swap(stay);
    }
    std::vector<particle> out;
    out.reserve(order.size());
    for (int i : order) {
        out.push_back(particles[i]);
        out.back().rank = my_rank;
    }
    particles.swap(out);
    num_particles = (int)particles.size();
    if ((int64_t)num_particles != cnt.remaining) {
        std::fprintf(stderr, "pic3d_compat: particle count mismatch after exchange (%d vs %lld)\n",
                     num_particles, (long long)cnt.remaining);
        std::exit(1);
    }
}
