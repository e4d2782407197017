// k_field.hip -- Yee half steps, field halo wrap, laser/Silver-Mueller
// boundary, current clear and current ghost fold.
//
// All of these are HBM-bound streaming kernels: one thread per cell (or per
// ghost element), x fastest so a wavefront reads 64 consecutive doubles.
// The FP64 instantiations keep the reference's operation order (this file
// is compiled with -ffp-contract=off) so single-step fields are bit-exact.
#include "laser.h"
#include "pic_internal.h"

namespace pic {

template <typename T> __device__ __forceinline__ T* F(void* p) { return reinterpret_cast<T*>(p); }

// ---------------------------------------------------------------------
// Yee: update_electromagnetic_fields_first/_second, field.cpp:56-117.
// E += dt/2 (curl B - J) with backward differences of B.
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_e_half(T* __restrict__ ex, T* __restrict__ ey,
                                                T* __restrict__ ez, const T* __restrict__ bx,
                                                const T* __restrict__ by, const T* __restrict__ bz,
                                                const T* __restrict__ jx, const T* __restrict__ jy,
                                                const T* __restrict__ jz, Grid g, T dth, T dx,
                                                T dy, T dz) {
    const int i = g.m + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.m + blockIdx.y;
    const int k = g.m + blockIdx.z;
    if (i >= g.m + g.n[0]) return;
    const int64_t o = g.fi(k, j, i);
    const int64_t sy = g.fe[0], sz = (int64_t)g.fe[0] * g.fe[1];
    const int64_t q = g.ji(k, j, i);
    const T bzo = bz[o], byo = by[o], bxo = bx[o];
    ex[o] += dth * ((bzo - bz[o - sy]) / dy - (byo - by[o - sz]) / dz - jx[q]);
    ey[o] += dth * ((bxo - bx[o - sz]) / dz - (bzo - bz[o - 1]) / dx - jy[q]);
    ez[o] += dth * ((byo - by[o - 1]) / dx - (bxo - bx[o - sy]) / dy - jz[q]);
}

// B -= dt/2 curl E with forward differences of E.
template <typename T>
__global__ __launch_bounds__(256) void k_b_half(const T* __restrict__ ex, const T* __restrict__ ey,
                                                const T* __restrict__ ez, T* __restrict__ bx,
                                                T* __restrict__ by, T* __restrict__ bz, Grid g,
                                                T dth, T dx, T dy, T dz) {
    const int i = g.m + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.m + blockIdx.y;
    const int k = g.m + blockIdx.z;
    if (i >= g.m + g.n[0]) return;
    const int64_t o = g.fi(k, j, i);
    const int64_t sy = g.fe[0], sz = (int64_t)g.fe[0] * g.fe[1];
    const T exo = ex[o], eyo = ey[o], ezo = ez[o];
    bx[o] -= dth * ((ez[o + sy] - ezo) / dy - (ey[o + sz] - eyo) / dz);
    by[o] -= dth * ((ex[o + sz] - exo) / dz - (ez[o + 1] - ezo) / dx);
    bz[o] -= dth * ((ey[o + 1] - eyo) / dx - (ex[o + sy] - exo) / dy);
}

// ---------------------------------------------------------------------
// Halo wrap of three components along one periodic axis: the single-domain
// form of exchange_boundary_data (mpi_utils.cpp:9-60).  Left ghosts
// [0, m-1] <- [N, N+m-1]; right ghosts [N+m, N+2m-1] <- [m, 2m-1]; over the
// full extent of the other two axes.  Axes are applied X, then Y, then Z.
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_wrap(T* __restrict__ a0, T* __restrict__ a1,
                                              T* __restrict__ a2, Grid g, int axis, int na) {
    const int m = g.m, N = g.n[axis];
    // element (u, v, gi) where gi in [0, 2m) selects the ghost layer
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int e0, e1;  // extents of the two other axes (faster one first)
    if (axis == 0) { e0 = g.fe[1]; e1 = g.fe[2]; }
    else if (axis == 1) { e0 = g.fe[0]; e1 = g.fe[2]; }
    else { e0 = g.fe[0]; e1 = g.fe[1]; }
    const int64_t total = (int64_t)2 * m * e0 * e1;
    if (t >= total) return;
    int gi, u, v;
    if (axis == 0) {  // ghost index fastest: x columns are contiguous
        gi = (int)(t % (2 * m)); int64_t r = t / (2 * m); u = (int)(r % e0); v = (int)(r / e0);
    } else {          // the x (or x,y) plane is contiguous
        u = (int)(t % e0); int64_t r = t / e0; gi = (int)(r % (2 * m)); v = (int)(r / (2 * m));
    }
    const int dst = gi < m ? gi : m + N + (gi - m);
    const int src = gi < m ? N + gi : m + (gi - m);
    int64_t od, os;
    if (axis == 0) { od = g.fi(v, u, dst); os = g.fi(v, u, src); }
    else if (axis == 1) { od = g.fi(v, dst, u); os = g.fi(v, src, u); }
    else { od = g.fi(dst, v, u); os = g.fi(src, v, u); }
    a0[od] = a0[os];
    if (na > 1) {
        a1[od] = a1[os];
        a2[od] = a2[os];
    }
}

// ---------------------------------------------------------------------
// One pass of the binomial current filter along one axis (filter_field,
// particle.cpp:520-567): interior nodes get 0.25 a[i-1] + 0.5 a[i] +
// 0.25 a[i+1] of the array as it was before the pass (src is that copy);
// the other two axes run over the interior only.
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_filter(const T* __restrict__ src, T* __restrict__ dst,
                                                Grid g, int axis) {
    const int i = g.m + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = g.m + blockIdx.y;
    const int k = g.m + blockIdx.z;
    if (i >= g.m + g.n[0]) return;
    const int64_t o = g.ji(k, j, i);
    const int64_t s = axis == 0 ? 1 : (axis == 1 ? (int64_t)g.je[0] : (int64_t)g.je[0] * g.je[1]);
    dst[o] = (T)0.25 * src[o - s] + (T)0.5 * src[o] + (T)0.25 * src[o + s];
}

// ---------------------------------------------------------------------
// Laser source + Silver-Mueller x boundary: radiating_boundary
// (boundary.cpp:42-99), Gpulse (29-40), generate_laser_pulse (8-27).
// is_left / is_right switch the x-min / x-max sources; both Silver-Mueller
// writes happen whenever x is open (the reference's rank_x* == -2).
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_laser(T* __restrict__ ey, T* __restrict__ ez,
                                               const T* __restrict__ bx, T* __restrict__ by,
                                               T* __restrict__ bz, const T* __restrict__ jy,
                                               const T* __restrict__ jz, Grid g, LaserArgs L) {
    const int j = g.m + blockIdx.x * blockDim.x + threadIdx.x;  // y
    const int k = g.m + blockIdx.y;                              // z
    if (j >= g.m + g.n[1]) return;
    const int m = g.m;
    const double dt = L.dt;
    const double alpha = 1 / (1 + dt / L.dx);
    const double beta = 1 - dt / L.dx;
    const double yy = ((double)j + 0.0) * L.dy - L.yfocus;
    const double zz = ((double)k + 0.0) * L.dz - L.zfocus;
    double Sl = 0.0, Pl = 0.0, Sr = 0.0, Pr = 0.0;
    if (L.left) {
        if (L.pol == 1 || L.pol == 3) Sl = d_gpulse(L, L.time, 0.0, yy, zz);
        if (L.pol == 2 || L.pol == 3) Pl = d_gpulse(L, L.time, 0.0, yy, zz);
    }
    if (L.right) {
        if (L.pol == 1 || L.pol == 3) Sr = d_gpulse(L, L.time, L.lx, yy, zz);
        if (L.pol == 2 || L.pol == 3) Pr = d_gpulse(L, L.time, L.lx, yy, zz);
    }
    const double dtdz = dt / L.dz, dtdy = dt / L.dy;
    {
        const int64_t o = g.fi(k, j, m);
        const double bzv = alpha * (4 * Sl - 2 * (double)ey[o] -
                                    dtdz * ((double)bx[o] - (double)bx[g.fi(k - 1, j, m)]) -
                                    beta * (double)bz[o] - dt * (double)jy[g.ji(k, j, m)]);
        const double byv = alpha * (-4 * Pl + 2 * (double)ez[o] +
                                    dtdy * ((double)bx[o] - (double)bx[g.fi(k, j - 1, m)]) -
                                    beta * (double)by[o] + dt * (double)jz[g.ji(k, j, m)]);
        bz[g.fi(k, j, m - 1)] = (T)bzv;
        by[g.fi(k, j, m - 1)] = (T)byv;
    }
    {
        // x-max: column local_xcells + 1 from column local_xcells
        // (boundary.cpp:92-96; an interior column when m > 2), or the face
        // form (L.xw, L.xre, L.xrb)
        const int64_t o = g.fi(k, j, L.xre), ob = g.fi(k, j, L.xrb);
        const double bzv = alpha * (-4 * Sr + 2 * (double)ey[o] -
                                    dtdz * ((double)bx[o] - (double)bx[g.fi(k - 1, j, L.xre)]) -
                                    beta * (double)bz[ob] + dt * (double)jy[g.ji(k, j, L.xre)]);
        const double byv = alpha * (4 * Pr - 2 * (double)ez[o] +
                                    dtdy * ((double)bx[o] - (double)bx[g.fi(k, j - 1, L.xre)]) -
                                    beta * (double)by[ob] - dt * (double)jz[g.ji(k, j, L.xre)]);
        bz[g.fi(k, j, L.xw)] = (T)bzv;
        by[g.fi(k, j, L.xw)] = (T)byv;
    }
}

// ---------------------------------------------------------------------
// Current ghost fold on one domain: exchange_boundary_current_fixed
// (mpi_utils.cpp:232-417).  Accumulate phase per periodic axis (X, Y, Z):
// interior [m, 2m] += ghost [N+m, N+2m]; interior [N-1, N+m-1] += ghost
// [-1, m-1].  Sync phase per periodic axis: ghosts <- those interiors.
// The slabs span the full J extent of the other axes.
// ---------------------------------------------------------------------
struct FoldPtrs { void* a[3 + PIC_MAX_SPECIES]; int na; };

template <typename T>
__global__ __launch_bounds__(256) void k_fold(FoldPtrs P, Grid g, int axis, int phase) {
    const int m = g.m, N = g.n[axis];
    const int st = m, end = m + N - 1;
    int e0, e1;
    if (axis == 0) { e0 = g.je[1]; e1 = g.je[2]; }
    else if (axis == 1) { e0 = g.je[0]; e1 = g.je[2]; }
    else { e0 = g.je[0]; e1 = g.je[1]; }
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)(m + 1) * e0 * e1;
    if (t >= total) return;
    int l, u, v;
    if (axis == 0) { l = (int)(t % (m + 1)); int64_t r = t / (m + 1); u = (int)(r % e0) - 1; v = (int)(r / e0) - 1; }
    else { u = (int)(t % e0) - 1; int64_t r = t / e0; l = (int)(r % (m + 1)); v = (int)(r / (m + 1)) - 1; }
    auto idx = [&](int s) -> int64_t {
        if (axis == 0) return g.ji(v, u, s);
        if (axis == 1) return g.ji(v, s, u);
        return g.ji(s, v, u);
    };
    const int64_t i_lo_int = idx(st + l), i_hi_gh = idx(end + 1 + l);
    const int64_t i_hi_int = idx(end - m + l), i_lo_gh = idx(-1 + l);
    for (int a = 0; a < P.na; a++) {
        T* D = reinterpret_cast<T*>(P.a[a]);
        if (phase == 0) {
            const T from_right = D[i_hi_gh], from_left = D[i_lo_gh];
            D[i_lo_int] += from_right;
            D[i_hi_int] += from_left;
        } else {
            D[i_lo_gh] = D[i_hi_int];
            D[i_hi_gh] = D[i_lo_int];
        }
    }
}

// zero fill (used instead of hipMemsetAsync so the step can be captured in
// a HIP graph)
__global__ __launch_bounds__(256) void k_fill0(uint4* __restrict__ p, int64_t n16) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; t < n16; t += stride) p[t] = make_uint4(0, 0, 0, 0);
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------

int launch_fill(pic_ctx* c, void* p, int64_t bytes) {
    if (bytes <= 0) return PIC_OK;
    // allocations are 256-byte aligned and sized to multiples of 16 bytes
    const int64_t n16 = (bytes + 15) / 16;
    int blocks = (int)std::min<int64_t>((n16 + 255) / 256, 8192);
    k_fill0<<<blocks, 256, 0, c->stream>>>(reinterpret_cast<uint4*>(p), n16);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

// Charge clear that keeps the previous charge for the Gauss diagnostics:
// prev = the sum of the mobile species' rho (species order), then every
// rho array is zeroed.  Grid-stride over the whole J layout.
struct RhoPtrs { void* a[PIC_MAX_SPECIES]; };

template <typename T>
__global__ __launch_bounds__(256) void k_clear_rho(RhoPtrs R, int mobile, double* __restrict__ prev,
                                                   int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        double sum = 0.0;
#pragma unroll
        for (int s = 0; s < PIC_MAX_SPECIES; s++) {
            T* a = reinterpret_cast<T*>(R.a[s]);
            if (!a) continue;
            if (mobile & (1 << s)) sum += (double)a[t];
            a[t] = (T)0;
        }
        prev[t] = sum;
    }
}

// clear_currents (particle.cpp:576-592): J and every rho array to zero,
// keeping the mobile charge in rho_prev first.
int launch_clear(pic_ctx* c) {
    const int64_t bytes = c->g.jsize * (c->fp32 ? 4 : 8);
    for (int f = 0; f < 3; f++) {
        int r = launch_fill(c, c->J[f], bytes);
        if (r) return r;
    }
    RhoPtrs R{};
    for (int s = 0; s < PIC_MAX_SPECIES; s++) R.a[s] = c->rho[s];
    const int64_t n = c->g.jsize;
    const int blocks = (int)std::min<int64_t>(blocks_for(n, 256), 8192);
    if (c->fp32) k_clear_rho<float><<<blocks, 256, 0, c->stream>>>(R, mobile_mask(c), c->rho_prev, n);
    else k_clear_rho<double><<<blocks, 256, 0, c->stream>>>(R, mobile_mask(c), c->rho_prev, n);
    PIC_LAUNCH_CHECK(c);
    c->rho_prev_valid = 1;
    return PIC_OK;
}

template <typename T> static int e_half_t(pic_ctx* c, double coef) {
    const Grid& g = c->g;
    dim3 grid(blocks_for(g.n[0], 256), g.n[1], g.n[2]);
    k_e_half<T><<<grid, 256, 0, c->stream>>>(
        (T*)c->F[0], (T*)c->F[1], (T*)c->F[2], (const T*)c->F[3], (const T*)c->F[4],
        (const T*)c->F[5], (const T*)c->J[0], (const T*)c->J[1], (const T*)c->J[2], g,
        (T)coef, (T)c->p.dx, (T)c->p.dy, (T)c->p.dz);
    PIC_LAUNCH_CHECK(c);
    return launch_wrap(c, 0);
}

template <typename T> static int b_half_t(pic_ctx* c, double coef) {
    const Grid& g = c->g;
    dim3 grid(blocks_for(g.n[0], 256), g.n[1], g.n[2]);
    k_b_half<T><<<grid, 256, 0, c->stream>>>(
        (const T*)c->F[0], (const T*)c->F[1], (const T*)c->F[2], (T*)c->F[3], (T*)c->F[4],
        (T*)c->F[5], g, (T)coef, (T)c->p.dx, (T)c->p.dy, (T)c->p.dz);
    PIC_LAUNCH_CHECK(c);
    return launch_wrap(c, 3);
}

// dt_half = dt / 2 (constants.cpp:42); the full step of
// update_electromagnetic_fields (field.cpp:24-54) uses dt.
int launch_e_update(pic_ctx* c, double coef) { return c->fp32 ? e_half_t<float>(c, coef) : e_half_t<double>(c, coef); }
int launch_b_update(pic_ctx* c, double coef) { return c->fp32 ? b_half_t<float>(c, coef) : b_half_t<double>(c, coef); }
int launch_e_half(pic_ctx* c) { return launch_e_update(c, c->p.dt / 2.0); }
int launch_b_half(pic_ctx* c) { return launch_b_update(c, c->p.dt / 2.0); }

template <typename T> static int wrap_t(pic_ctx* c, int first, int count) {
    const Grid& g = c->g;
    for (int a = 0; a < 3; a++) {
        if (!c->p.periodic[a]) continue;
        int e0 = a == 0 ? g.fe[1] : g.fe[0];
        int e1 = a == 2 ? g.fe[1] : g.fe[2];
        int64_t total = (int64_t)2 * g.m * e0 * e1;
        k_wrap<T><<<blocks_for(total, 256), 256, 0, c->stream>>>(
            (T*)c->F[first], (T*)c->F[first + (count > 1 ? 1 : 0)], (T*)c->F[first + (count > 1 ? 2 : 0)],
            g, a, count);
        PIC_LAUNCH_CHECK(c);
    }
    return PIC_OK;
}

int launch_wrap(pic_ctx* c, int first) { return c->fp32 ? wrap_t<float>(c, first, 3) : wrap_t<double>(c, first, 3); }
int launch_wrap_one(pic_ctx* c, int comp) { return c->fp32 ? wrap_t<float>(c, comp, 1) : wrap_t<double>(c, comp, 1); }

LaserArgs laser_args(const pic_ctx* c, int step, int left, int right) {
    const pic_params& P = c->p;
    LaserArgs L;
    L.pi = P.pi; L.l0 = P.l0; L.waist = P.waist; L.xfocus = P.xfocus; L.yfocus = P.yfocus;
    L.zfocus = P.zfocus; L.omega = P.omega; L.phase = P.phase; L.amp = P.laser_amplitude;
    L.dur = P.laser_pulse_duration; L.t_start = P.laser_start_time; L.lx = P.lx;
    L.dt = P.dt; L.dx = P.dx; L.dy = P.dy; L.dz = P.dz; L.time = step * P.dt;
    L.pol = P.laser_polarization; L.left = left; L.right = right;
    L.env_on = c->env_on; L.env_val = c->env_val;
    const int N = c->g.n[0], m = c->g.m;
    if (c->xmax_face) { L.xw = N + m - 1; L.xre = N + m - 1; L.xrb = N + m - 2; }
    else { L.xw = N + 1; L.xre = N; L.xrb = N; }
    return L;
}

template <typename T> static int laser_t(pic_ctx* c, int step, int left, int right) {
    const LaserArgs L = laser_args(c, step, left, right);
    dim3 grid(blocks_for(c->g.n[1], 256), c->g.n[2]);
    k_laser<T><<<grid, 256, 0, c->stream>>>((T*)c->F[1], (T*)c->F[2], (const T*)c->F[3],
                                            (T*)c->F[4], (T*)c->F[5], (const T*)c->J[1],
                                            (const T*)c->J[2], c->g, L);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

int launch_laser(pic_ctx* c, int step, int left, int right) {
    if (c->p.periodic[0]) return PIC_OK;  // x neighbours exist: no source
    return c->fp32 ? laser_t<float>(c, step, left, right) : laser_t<double>(c, step, left, right);
}

template <typename T> static int fold_t(pic_ctx* c, void* only) {
    FoldPtrs P{};
    P.na = 0;
    if (only) {
        P.a[P.na++] = only;
    } else {
        for (int f = 0; f < 3; f++) P.a[P.na++] = c->J[f];
        for (int s = 0; s < PIC_MAX_SPECIES; s++)
            if (c->sp[s].declared && c->rho[s]) P.a[P.na++] = c->rho[s];
    }
    const Grid& g = c->g;
    for (int phase = 0; phase < 2; phase++)
        for (int a = 0; a < 3; a++) {
            if (!c->p.periodic[a]) continue;
            int e0 = a == 0 ? g.je[1] : g.je[0];
            int e1 = a == 2 ? g.je[1] : g.je[2];
            int64_t total = (int64_t)(g.m + 1) * e0 * e1;
            k_fold<T><<<blocks_for(total, 256), 256, 0, c->stream>>>(P, g, a, phase);
            PIC_LAUNCH_CHECK(c);
        }
    return PIC_OK;
}

int launch_fold(pic_ctx* c) { return c->fp32 ? fold_t<float>(c, nullptr) : fold_t<double>(c, nullptr); }
int launch_fold_one(pic_ctx* c, void* arr) { return c->fp32 ? fold_t<float>(c, arr) : fold_t<double>(c, arr); }

template <typename T> static int filter_t(pic_ctx* c, void* tmp) {
    const Grid& g = c->g;
    dim3 grid(blocks_for(g.n[0], 256), g.n[1], g.n[2]);
    for (int f = 0; f < 3; f++)
        for (int a = 0; a < 3; a++) {
            PIC_HIP(c, hipMemcpyAsync(tmp, c->J[f], g.jsize * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
            k_filter<T><<<grid, 256, 0, c->stream>>>((const T*)tmp, (T*)c->J[f], g, a);
            PIC_LAUNCH_CHECK(c);
        }
    return PIC_OK;
}
int launch_filter(pic_ctx* c, void* tmp) { return c->fp32 ? filter_t<float>(c, tmp) : filter_t<double>(c, tmp); }

}  // namespace pic
