// k_lattice.hip -- the reference's initial loading on the device:
// initialize_particles (particle.cpp:50-111) with the box density of
// plasma_density (constants.cpp:87-103), FP64 or FP32 layout.
//
// The density test is a product of one interval per axis, and the lattice
// coordinate grows with its index, so the kept points are a box of lattice
// indices [lo, lo + cnt) per axis; the host finds it with the reference's
// own expressions and each thread writes one kept point.
#include <cmath>

#include "pic_internal.h"

namespace pic {

namespace {

struct LatticeGeom {
    int num[3];       // lattice points per axis
    double norm[3];   // lattice spacing, cells
    int lo[3], cnt[3];  // kept index range per axis
    int64_t total;    // particles written
};

// position of lattice index i on an axis of n cells (particle.cpp:75-77;
// xst = 0 and xend = n on one domain, simulation.cpp:31-34)
__host__ __device__ inline double lattice_pos(int i, double norm) { return 0.0 + i * norm + norm / 2.0; }

LatticeGeom lattice_geom(const pic_params* p, const pic_lattice* L) {
    LatticeGeom G{};
    const int n[3] = {p->nx, p->ny, p->nz};
    const double d[3] = {p->dx, p->dy, p->dz};
    const int cpp = L->cell_per_particles;
    for (int a = 0; a < 3; a++) {
        G.num[a] = (int)(n[a] * std::cbrt(cpp));  // particle.cpp:59-61
        G.norm[a] = ((double)n[a] - 0.0) / (G.num[a]);
        // constants.cpp:95: lo*l0 < x*d <= hi*l0
        int first = -1, last = -1;
        for (int i = 0; i < G.num[a]; i++) {
            const double xp = lattice_pos(i, G.norm[a]) * d[a];
            if (xp > L->lo[a] * p->l0 && xp <= L->hi[a] * p->l0) {
                if (first < 0) first = i;
                last = i;
            }
        }
        G.lo[a] = first < 0 ? 0 : first;
        G.cnt[a] = first < 0 ? 0 : last - first + 1;
    }
    // the reference writes at most n_x n_y n_z cpp particles (particle.cpp:58, 71-72)
    const int64_t cap = (int64_t)n[0] * n[1] * n[2] * cpp;
    const int64_t kept = L->density / cpp > 0.0 ? (int64_t)G.cnt[0] * G.cnt[1] * G.cnt[2] : 0;
    G.total = kept < cap ? kept : cap;
    return G;
}

struct LatticeOut64 {
    double *x, *y, *z, *ux, *uy, *uz, *w;
    int64_t* id;
    double *sx, *sy, *sz;
};
struct LatticeOut32 {
    int32_t* cell;
    float *fx, *fy, *fz, *ux, *uy, *uz, *w;
    int64_t* id;
    float *sx, *sy, *sz;
    int n[3];
};

__device__ __forceinline__ void lattice_point(int64_t q, const LatticeGeom& G, double& x, double& y, double& z) {
    const int i = G.lo[0] + (int)(q % G.cnt[0]);
    const int j = G.lo[1] + (int)((q / G.cnt[0]) % G.cnt[1]);
    const int k = G.lo[2] + (int)(q / ((int64_t)G.cnt[0] * G.cnt[1]));
    x = lattice_pos(i, G.norm[0]);
    y = lattice_pos(j, G.norm[1]);
    z = lattice_pos(k, G.norm[2]);
}

__global__ __launch_bounds__(256) void k_lattice64(LatticeOut64 O, LatticeGeom G, double w, double u0, double u1,
                                                   double u2, double s0, double s1, double s2, int64_t id0) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < G.total;
         q += (int64_t)gridDim.x * blockDim.x) {
        double x, y, z;
        lattice_point(q, G, x, y, z);
        O.x[q] = x; O.y[q] = y; O.z[q] = z;
        O.ux[q] = u0; O.uy[q] = u1; O.uz[q] = u2;
        O.w[q] = w;
        O.id[q] = id0 + q;
        if (O.sx) { O.sx[q] = s0; O.sy[q] = s1; O.sz[q] = s2; }
    }
}

__global__ __launch_bounds__(256) void k_lattice32(LatticeOut32 O, LatticeGeom G, double w, double u0, double u1,
                                                   double u2, double s0, double s1, double s2, int64_t id0) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < G.total;
         q += (int64_t)gridDim.x * blockDim.x) {
        double v[3];
        lattice_point(q, G, v[0], v[1], v[2]);
        int c[3];
        float f[3];
        for (int a = 0; a < 3; a++) to_cell_offset(v[a], O.n[a], c[a], f[a]);
        O.cell[q] = (c[2] * O.n[1] + c[1]) * O.n[0] + c[0];
        O.fx[q] = f[0]; O.fy[q] = f[1]; O.fz[q] = f[2];
        O.ux[q] = (float)u0; O.uy[q] = (float)u1; O.uz[q] = (float)u2;
        O.w[q] = (float)w;
        O.id[q] = id0 + q;
        if (O.sx) { O.sx[q] = (float)s0; O.sy[q] = (float)s1; O.sz[q] = (float)s2; }
    }
}

}  // namespace

int launch_lattice(pic_ctx* c, int s, const pic_lattice* L, int64_t* n_out) {
    Species& S = c->sp[s];
    const LatticeGeom G = lattice_geom(&c->p, L);
    if (G.total > S.cap)
        return set_err(c, PIC_ERR_CAPACITY, "the lattice makes %lld particles, capacity %lld", (long long)G.total,
                       (long long)S.cap);
    const double w = L->density / L->cell_per_particles;  // particle.cpp:79
    const int b = S.cur;
    const int blocks = (int)std::min<int64_t>(blocks_for(std::max<int64_t>(G.total, 1), 256), 256 * 64);
    if (c->fp32) {
        LatticeOut32 O{S.cell[b], S.fx[b], S.fy[b], S.fz[b], S.vx[b], S.vy[b], S.vz[b], S.fw[b], S.id[b],
                       (float*)S.spin_v[0][b], (float*)S.spin_v[1][b], (float*)S.spin_v[2][b],
                       {c->g.n[0], c->g.n[1], c->g.n[2]}};
        k_lattice32<<<blocks, 256, 0, c->stream>>>(O, G, w, L->u[0], L->u[1], L->u[2], L->spin[0], L->spin[1],
                                                   L->spin[2], L->id_offset);
    } else {
        LatticeOut64 O{S.x[b], S.y[b], S.z[b], S.ux[b], S.uy[b], S.uz[b], S.w[b], S.id[b],
                       (double*)S.spin_v[0][b], (double*)S.spin_v[1][b], (double*)S.spin_v[2][b]};
        k_lattice64<<<blocks, 256, 0, c->stream>>>(O, G, w, L->u[0], L->u[1], L->u[2], L->spin[0], L->spin[1],
                                                   L->spin[2], L->id_offset);
    }
    PIC_LAUNCH_CHECK(c);
    S.wmax = std::fabs(w);
    *n_out = G.total;
    return PIC_OK;
}

}  // namespace pic

extern "C" int64_t pic_lattice_count(const pic_params* p, const pic_lattice* l) {
    if (!p || !l || l->cell_per_particles < 1) return -1;
    return pic::lattice_geom(p, l).total;
}
