// k_push64.hip -- FP64 particle push: periodic wrap, staggered field gather,
// relativistic Boris rotation, position advance and Esirkepov
// charge-conserving current deposition (update_particle_position_momentum,
// particle.cpp:113-316, and currents_particle, particle.cpp:372-515).
//
// Compiled with -ffp-contract=off: every expression keeps the reference's
// association order, so a particle's new momentum and position are
// bit-identical to the reference given the same fields.  Only the order in
// which particles add into J (atomics) differs.
//
// Two kernels:
//  * k_push64_direct: one thread per particle over slots, gather from HBM
//    (through L2), deposit with global FP64 atomics.  Works on any order.
//  * k_push64_tiled: one workgroup per (tile, chunk) of the tile-sorted
//    particle list; currents are accumulated in LDS (ds_add_f64) over the
//    tile plus deposit halo and flushed to HBM once per workgroup.
#include "pic_internal.h"
#include "shapes.h"

namespace pic {

// Dead slots (removed at an open boundary, not yet compacted) carry a NaN
// weight and a negative id.
struct P64 {
    double *x, *y, *z, *ux, *uy, *uz;
    const double* w;
    double *xo, *yo, *zo;  // may be null
    int64_t* cnt;
};

struct Fields64 {
    const double *ex, *ey, *ez, *bx, *by, *bz;
    double *jx, *jy, *jz, *rho;
};

struct PushConsts {
    double q, qm, dt, dx, dy, dz, cx, cy, cz;  // cx = q*dx/dt (times w per particle)
    double N[3];
    int per[3];
};

// Periodic wrap at the start of the push (particle.cpp:123-167) followed by
// the in-domain check.  Returns false for a particle outside the domain.
__device__ __forceinline__ bool wrap_check(double& x, double& y, double& z, const PushConsts& K) {
    if (x < 0.0 || x > K.N[0] || y < 0.0 || y > K.N[1] || z < 0.0 || z > K.N[2]) {
        if (K.per[0] || K.per[1] || K.per[2]) {
            if (K.per[0]) { if (x < 0.0) x += K.N[0]; if (x > K.N[0]) x -= K.N[0]; }
            if (K.per[1]) { if (y < 0.0) y += K.N[1]; if (y > K.N[1]) y -= K.N[1]; }
            if (K.per[2]) { if (z < 0.0) z += K.N[2]; if (z > K.N[2]) z -= K.N[2]; }
        }
    }
    return x >= 0.0 && x <= K.N[0] && y >= 0.0 && y <= K.N[1] && z >= 0.0 && z <= K.N[2];
}

// Gather + Boris + advance for one particle.  Field reads go through the
// functor FR(comp, k, j, i) so the same code serves HBM and LDS tiles.
template <int M, typename FR>
__device__ __forceinline__ void gather_boris(double x, double y, double z, double& ux, double& uy,
                                             double& uz, double& xn, double& yn, double& zn,
                                             const PushConsts& K, int m, FR fr) {
    constexpr int H = M / 2;
    const double org = (double)(1 - m);
    const double xc = x - org, yc = y - org, zc = z - org;
    double Wxp[M + 1], Wyp[M + 1], Wzp[M + 1], Wxd[M + 1], Wyd[M + 1], Wzd[M + 1];
    const int ip = shape_at<M>(xc, Wxp), jp = shape_at<M>(yc, Wyp), kp = shape_at<M>(zc, Wzp);
    shape_at<M>(xc - 0.5, Wxd);
    shape_at<M>(yc - 0.5, Wyd);
    shape_at<M>(zc - 0.5, Wzd);
    const int id = (int)xc, jd = (int)yc, kd = (int)zc;
    double e0 = 0.0, e1 = 0.0, e2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
    // particle.cpp:207-259: sum_k sum_j sum_i ((Wx*Wy)*Wz)*F
#pragma unroll
    for (int k = 0; k <= M; k++)
#pragma unroll
        for (int j = 0; j <= M; j++)
#pragma unroll
            for (int i = 0; i <= M; i++) e0 += Wxd[i] * Wyp[j] * Wzp[k] * fr(0, kp + k - H, jp + j - H, id + i - H);
#pragma unroll
    for (int k = 0; k <= M; k++)
#pragma unroll
        for (int j = 0; j <= M; j++)
#pragma unroll
            for (int i = 0; i <= M; i++) e1 += Wxp[i] * Wyd[j] * Wzp[k] * fr(1, kp + k - H, jd + j - H, ip + i - H);
#pragma unroll
    for (int k = 0; k <= M; k++)
#pragma unroll
        for (int j = 0; j <= M; j++)
#pragma unroll
            for (int i = 0; i <= M; i++) e2 += Wxp[i] * Wyp[j] * Wzd[k] * fr(2, kd + k - H, jp + j - H, ip + i - H);
#pragma unroll
    for (int k = 0; k <= M; k++)
#pragma unroll
        for (int j = 0; j <= M; j++)
#pragma unroll
            for (int i = 0; i <= M; i++) b0 += Wxp[i] * Wyd[j] * Wzd[k] * fr(3, kd + k - H, jd + j - H, ip + i - H);
#pragma unroll
    for (int k = 0; k <= M; k++)
#pragma unroll
        for (int j = 0; j <= M; j++)
#pragma unroll
            for (int i = 0; i <= M; i++) b1 += Wxd[i] * Wyp[j] * Wzd[k] * fr(4, kd + k - H, jp + j - H, id + i - H);
#pragma unroll
    for (int k = 0; k <= M; k++)
#pragma unroll
        for (int j = 0; j <= M; j++)
#pragma unroll
            for (int i = 0; i <= M; i++) b2 += Wxd[i] * Wyd[j] * Wzp[k] * fr(5, kp + k - H, jd + j - H, id + i - H);
    // Boris, particle.cpp:271-311
    e0 *= K.qm; e1 *= K.qm; e2 *= K.qm; b0 *= K.qm; b1 *= K.qm; b2 *= K.qm;
    const double px1 = ux + e0, py1 = uy + e1, pz1 = uz + e2;
    const double g1 = sqrt(1.0 + (px1 * px1 + py1 * py1 + pz1 * pz1));
    const double tx = b0 / g1, ty = b1 / g1, tz = b2 / g1;
    const double tt = 2.0 / (1.0 + tx * tx + ty * ty + tz * tz);
    const double sx = tx * tt, sy = ty * tt, sz = tz * tt;
    const double px2 = px1 + (py1 * tz - pz1 * ty);
    const double py2 = py1 + (pz1 * tx - px1 * tz);
    const double pz2 = pz1 + (px1 * ty - py1 * tx);
    const double px3 = px1 + (py2 * sz - pz2 * sy);
    const double py3 = py1 + (pz2 * sx - px2 * sz);
    const double pz3 = pz1 + (px2 * sy - py2 * sx);
    ux = px3 + e0;
    uy = py3 + e1;
    uz = pz3 + e2;
    const double gam = sqrt(1.0 + (ux * ux + uy * uy + uz * uz));
    // particle.cpp:313-315
    xn = x + ux / gam * K.dt / K.dx;
    yn = y + uy / gam * K.dt / K.dy;
    zn = z + uz / gam * K.dt / K.dz;
}

// Esirkepov deposit of one particle moving (xo,yo,zo) -> (xn,yn,zn)
// (particle.cpp:381-514).  ADD(comp, k, j, i, v) adds v (comp 0..2 = J,
// 3 = rho).
//
// The reference walks a box [min(i1,i2)-H, max(i1,i2)+H] per axis (M+1 or
// M+2 points).  Here the box is always M+2 points with the old/new shape
// weights placed at compile-time offsets (selected by d0/d1 in {0,1}), so
// the loops unroll into registers; points past the reference's box are
// predicated off.  Values and the order of the cumulative sums are the
// reference's, so each deposited term is bit-identical.
template <int M>
__device__ __forceinline__ void place(const double* S, int d, double* A) {
#pragma unroll
    for (int t = 0; t < M + 2; t++) {
        const double a0 = (t <= M) ? S[t] : 0.0;
        const double a1 = (t >= 1) ? S[t - 1] : 0.0;
        A[t] = d ? a1 : a0;
    }
}

template <int M, typename ADD>
__device__ __noinline__ void esirkepov_generic(double qw, double xo, double yo, double zo,
                                               double xn, double yn, double zn, double org,
                                               const PushConsts& K, ADD add) {
    constexpr int H = M / 2;
    double Sx0[M + 1], Sx1[M + 1], Sy0[M + 1], Sy1[M + 1], Sz0[M + 1], Sz1[M + 1];
    const int i1 = shape_at<M>(xo - org, Sx0), i2 = shape_at<M>(xn - org, Sx1);
    const int j1 = shape_at<M>(yo - org, Sy0), j2 = shape_at<M>(yn - org, Sy1);
    const int k1 = shape_at<M>(zo - org, Sz0), k2 = shape_at<M>(zn - org, Sz1);
    const int imin = min(i1, i2) - H, imax = max(i1, i2) + H;
    const int jmin = min(j1, j2) - H, jmax = max(j1, j2) + H;
    const int kmin = min(k1, k2) - H, kmax = max(k1, k2) + H;
    const double cx = qw * K.dx / K.dt, cy = qw * K.dy / K.dt, cz = qw * K.dz / K.dt;
    auto sv = [](const double* S, int ic, int v) -> double {
        return (v >= ic - H && v <= ic + H) ? S[v - ic + H] : 0.0;
    };
    for (int k = kmin; k <= kmax; k++) {
        const double sz0 = sv(Sz0, k1, k), dsz = sv(Sz1, k2, k) - sz0;
        for (int j = jmin; j <= jmax; j++) {
            const double sy0 = sv(Sy0, j1, j), dsy = sv(Sy1, j2, j) - sy0;
            const double wyz = sy0 * sz0 + 0.5 * dsy * sz0 + 0.5 * sy0 * dsz + (1.0 / 3.0) * dsy * dsz;
            double acc = 0.0;
            for (int i = imin; i <= imax; i++) {
                acc += sv(Sx1, i2, i) - sv(Sx0, i1, i);
                add(0, k, j, i, -(cx * acc * wyz));
            }
        }
    }
    for (int k = kmin; k <= kmax; k++) {
        const double sz0 = sv(Sz0, k1, k), dsz = sv(Sz1, k2, k) - sz0;
        for (int i = imin; i <= imax; i++) {
            const double sx0 = sv(Sx0, i1, i), dsx = sv(Sx1, i2, i) - sx0;
            const double wxz = sx0 * sz0 + 0.5 * dsx * sz0 + 0.5 * sx0 * dsz + (1.0 / 3.0) * dsx * dsz;
            double acc = 0.0;
            for (int j = jmin; j <= jmax; j++) {
                acc += sv(Sy1, j2, j) - sv(Sy0, j1, j);
                add(1, k, j, i, -(cy * acc * wxz));
            }
        }
    }
    for (int j = jmin; j <= jmax; j++) {
        const double sy0 = sv(Sy0, j1, j), dsy = sv(Sy1, j2, j) - sy0;
        for (int i = imin; i <= imax; i++) {
            const double sx0 = sv(Sx0, i1, i), dsx = sv(Sx1, i2, i) - sx0;
            const double wxy = sx0 * sy0 + 0.5 * dsx * sy0 + 0.5 * sx0 * dsy + (1.0 / 3.0) * dsx * dsy;
            double acc = 0.0;
            for (int k = kmin; k <= kmax; k++) {
                acc += sv(Sz1, k2, k) - sv(Sz0, k1, k);
                add(2, k, j, i, -(cz * acc * wxy));
            }
        }
    }
    for (int k = kmin; k <= kmax; k++) {
        const double sz = sv(Sz1, k2, k);
        if (sz == 0.0) continue;
        for (int j = jmin; j <= jmax; j++) {
            const double sy = sv(Sy1, j2, j);
            if (sy == 0.0) continue;
            for (int i = imin; i <= imax; i++) add(3, k, j, i, qw * sv(Sx1, i2, i) * sy * sz);
        }
    }
}

template <int M, typename ADD>
__device__ __forceinline__ void esirkepov(double qw, double xo, double yo, double zo, double xn,
                                          double yn, double zn, const PushConsts& K, int m,
                                          ADD add) {
    constexpr int H = M / 2, L = M + 2;
    const double org = (double)(1 - m);
    double Sx0[M + 1], Sx1[M + 1], Sy0[M + 1], Sy1[M + 1], Sz0[M + 1], Sz1[M + 1];
    const int i1 = shape_at<M>(xo - org, Sx0), i2 = shape_at<M>(xn - org, Sx1);
    const int j1 = shape_at<M>(yo - org, Sy0), j2 = shape_at<M>(yn - org, Sy1);
    const int k1 = shape_at<M>(zo - org, Sz0), k2 = shape_at<M>(zn - org, Sz1);
    if (abs(i2 - i1) > 1 || abs(j2 - j1) > 1 || abs(k2 - k1) > 1) {
        // a move of more than one node in a step (dt beyond the Courant limit)
        esirkepov_generic<M>(qw, xo, yo, zo, xn, yn, zn, org, K, add);
        return;
    }
    const int bx = min(i1, i2) - H, by = min(j1, j2) - H, bz = min(k1, k2) - H;
    const int Lx = abs(i2 - i1) + M + 1, Ly = abs(j2 - j1) + M + 1, Lz = abs(k2 - k1) + M + 1;
    double X0[L], X1[L], Y0[L], Y1[L], Z0[L], Z1[L];
    place<M>(Sx0, i1 - H - bx, X0); place<M>(Sx1, i2 - H - bx, X1);
    place<M>(Sy0, j1 - H - by, Y0); place<M>(Sy1, j2 - H - by, Y1);
    place<M>(Sz0, k1 - H - bz, Z0); place<M>(Sz1, k2 - H - bz, Z1);
    const double cx = qw * K.dx / K.dt, cy = qw * K.dy / K.dt, cz = qw * K.dz / K.dt;
    // Jx: cumulative sum along x, weights W_yz (particle.cpp:427-450)
#pragma unroll
    for (int k = 0; k < L; k++) {
        const double sz0 = Z0[k], dsz = Z1[k] - sz0;
#pragma unroll
        for (int j = 0; j < L; j++) {
            const double sy0 = Y0[j], dsy = Y1[j] - sy0;
            const double wyz = sy0 * sz0 + 0.5 * dsy * sz0 + 0.5 * sy0 * dsz + (1.0 / 3.0) * dsy * dsz;
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < L; i++) {
                acc += X1[i] - X0[i];
                if (i < Lx && j < Ly && k < Lz) add(0, bz + k, by + j, bx + i, -(cx * acc * wyz));
            }
        }
    }
    // Jy (particle.cpp:452-475)
#pragma unroll
    for (int k = 0; k < L; k++) {
        const double sz0 = Z0[k], dsz = Z1[k] - sz0;
#pragma unroll
        for (int i = 0; i < L; i++) {
            const double sx0 = X0[i], dsx = X1[i] - sx0;
            const double wxz = sx0 * sz0 + 0.5 * dsx * sz0 + 0.5 * sx0 * dsz + (1.0 / 3.0) * dsx * dsz;
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < L; j++) {
                acc += Y1[j] - Y0[j];
                if (i < Lx && j < Ly && k < Lz) add(1, bz + k, by + j, bx + i, -(cy * acc * wxz));
            }
        }
    }
    // Jz (particle.cpp:477-500)
#pragma unroll
    for (int j = 0; j < L; j++) {
        const double sy0 = Y0[j], dsy = Y1[j] - sy0;
#pragma unroll
        for (int i = 0; i < L; i++) {
            const double sx0 = X0[i], dsx = X1[i] - sx0;
            const double wxy = sx0 * sy0 + 0.5 * dsx * sy0 + 0.5 * sx0 * dsy + (1.0 / 3.0) * dsx * dsy;
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < L; k++) {
                acc += Z1[k] - Z0[k];
                if (i < Lx && j < Ly && k < Lz) add(2, bz + k, by + j, bx + i, -(cz * acc * wxy));
            }
        }
    }
    // rho at the new position (particle.cpp:502-514): nodes i2-H .. i2+H
#pragma unroll
    for (int k = 0; k <= M; k++) {
        const double sz = Sz1[k];
        if (sz == 0.0) continue;
#pragma unroll
        for (int j = 0; j <= M; j++) {
            const double sy = Sy1[j];
            if (sy == 0.0) continue;
#pragma unroll
            for (int i = 0; i <= M; i++) add(3, k2 - H + k, j2 - H + j, i2 - H + i, qw * Sx1[i] * sy * sz);
        }
    }
}

// ---------------------------------------------------------------------
// direct kernel: any particle order, global atomics
// ---------------------------------------------------------------------
template <int M>
__global__ __launch_bounds__(256) void k_push64_direct(P64 P, Fields64 F, Grid g, PushConsts K) {
    const int64_t n = P.cnt[CNT_SLOTS];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t oob = 0;
    const double* fp[6] = {F.ex, F.ey, F.ez, F.bx, F.by, F.bz};
    auto fr = [&](int c, int k, int j, int i) -> double { return fp[c][g.fi(k, j, i)]; };
    double* jp[4] = {F.jx, F.jy, F.jz, F.rho};
    auto add = [&](int c, int k, int j, int i, double v) { unsafeAtomicAdd(&jp[c][g.ji(k, j, i)], v); };
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        const double wp = P.w[p];
        if (!(wp == wp)) continue;
        double x = P.x[p], y = P.y[p], z = P.z[p];
        if (!wrap_check(x, y, z, K)) { oob++; continue; }
        double ux = P.ux[p], uy = P.uy[p], uz = P.uz[p], xn, yn, zn;
        gather_boris<M>(x, y, z, ux, uy, uz, xn, yn, zn, K, g.m, fr);
        P.x[p] = xn; P.y[p] = yn; P.z[p] = zn;
        P.ux[p] = ux; P.uy[p] = uy; P.uz[p] = uz;
        if (P.xo) { P.xo[p] = x; P.yo[p] = y; P.zo[p] = z; }
        // a non-finite momentum would put the deposit box outside the arrays
        if (!(fabs(xn - x) < 1.0 && fabs(yn - y) < 1.0 && fabs(zn - z) < 1.0)) { oob++; continue; }
        esirkepov<M>(K.q * wp, x, y, z, xn, yn, zn, K, g.m, add);
    }
    if (oob) atomicAdd((unsigned long long*)&P.cnt[CNT_OOB], (unsigned long long)oob);
}

// ---------------------------------------------------------------------
// tiled kernel: particles sorted by tile; J and rho accumulated in LDS
// ---------------------------------------------------------------------
// LDS accumulator box for a tile whose first cell is (T0x, T0y, T0z) with
// edges (tx, ty, tz): buffer indices [T0 + m - 2 - H, T0 + t - 1 + m + 1 + H],
// i.e. t + m + 3 points per axis (particles move < 1 cell per step).
template <int M, int TX, int TY, int TZ>
struct AccBox {
    static constexpr int H = M / 2;
    static constexpr int BX = TX + M + 3, BY = TY + M + 3, BZ = TZ + M + 3;
    static constexpr int VOL = BX * BY * BZ;
};

template <int M, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) void k_push64_tiled(P64 P, Fields64 F, Grid g, PushConsts K,
                                                      Tiling tl, const int64_t* __restrict__ tile_off,
                                                      const int64_t* __restrict__ work,
                                                      const int64_t* __restrict__ nwork, int chunk) {
    using B = AccBox<M, TX, TY, TZ>;
    extern __shared__ double acc[];  // 4 * VOL
    const int64_t wi = blockIdx.x;
    if (wi >= *nwork) return;
    const int64_t item = work[wi];            // tile * 2^20 + chunk index
    const int tile = (int)(item >> 20);
    const int64_t cidx = item & ((1 << 20) - 1);
    const int64_t t_beg = tile_off[tile], t_end = tile_off[tile + 1];
    const int64_t beg = t_beg + cidx * chunk;
    const int64_t end = min(t_end, beg + chunk);
    const int tzi = tile / (tl.nt[0] * tl.nt[1]);
    const int tyi = (tile / tl.nt[0]) % tl.nt[1];
    const int txi = tile % tl.nt[0];
    // buffer-index origin of the accumulator box
    const int ox = txi * TX + g.m - 2 - B::H;
    const int oy = tyi * TY + g.m - 2 - B::H;
    const int oz = tzi * TZ + g.m - 2 - B::H;
    for (int e = threadIdx.x; e < 4 * B::VOL; e += blockDim.x) acc[e] = 0.0;
    __syncthreads();
    const double* fp[6] = {F.ex, F.ey, F.ez, F.bx, F.by, F.bz};
    auto fr = [&](int c, int k, int j, int i) -> double { return fp[c][g.fi(k, j, i)]; };
    int64_t oob = 0, spill = 0;
    double* jp[4] = {F.jx, F.jy, F.jz, F.rho};
    auto add = [&](int c, int k, int j, int i, double v) {
        const int lx = i - ox, ly = j - oy, lz = k - oz;
        if ((unsigned)lx < (unsigned)B::BX && (unsigned)ly < (unsigned)B::BY &&
            (unsigned)lz < (unsigned)B::BZ) {
            atomicAdd(&acc[c * B::VOL + (lz * B::BY + ly) * B::BX + lx], v);
        } else {  // outside the box (only for a particle outside its tile)
            unsafeAtomicAdd(&jp[c][g.ji(k, j, i)], v);
            spill++;
        }
    };
    for (int64_t p = beg + threadIdx.x; p < end; p += blockDim.x) {
        const double wp = P.w[p];
        if (!(wp == wp)) continue;
        double x = P.x[p], y = P.y[p], z = P.z[p];
        if (!wrap_check(x, y, z, K)) { oob++; continue; }
        double ux = P.ux[p], uy = P.uy[p], uz = P.uz[p], xn, yn, zn;
        gather_boris<M>(x, y, z, ux, uy, uz, xn, yn, zn, K, g.m, fr);
        P.x[p] = xn; P.y[p] = yn; P.z[p] = zn;
        P.ux[p] = ux; P.uy[p] = uy; P.uz[p] = uz;
        if (P.xo) { P.xo[p] = x; P.yo[p] = y; P.zo[p] = z; }
        // a non-finite momentum would put the deposit box outside the arrays
        if (!(fabs(xn - x) < 1.0 && fabs(yn - y) < 1.0 && fabs(zn - z) < 1.0)) { oob++; continue; }
        esirkepov<M>(K.q * wp, x, y, z, xn, yn, zn, K, g.m, add);
    }
    __syncthreads();
    // flush: one global atomic per nonzero accumulator point
    for (int e = threadIdx.x; e < 4 * B::VOL; e += blockDim.x) {
        const double v = acc[e];
        if (v == 0.0) continue;
        const int c = e / B::VOL, r = e % B::VOL;
        const int lx = r % B::BX, ly = (r / B::BX) % B::BY, lz = r / (B::BX * B::BY);
        const int gx = ox + lx, gy = oy + ly, gz = oz + lz;
        // a partial last tile's box reaches past the arrays; those points
        // are never deposited into, but keep the flush in bounds regardless
        if (gx > g.je[0] - 2 || gy > g.je[1] - 2 || gz > g.je[2] - 2) continue;
        unsafeAtomicAdd(&jp[c][g.ji(gz, gy, gx)], v);
    }
    if (oob) atomicAdd((unsigned long long*)&P.cnt[CNT_OOB], (unsigned long long)oob);
    (void)spill;
}

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------
static PushConsts make_consts(const pic_ctx* c, const Species& S) {
    PushConsts K;
    K.q = S.q;
    K.qm = S.q / (S.m) * (c->p.dt / 2.0);
    K.dt = c->p.dt; K.dx = c->p.dx; K.dy = c->p.dy; K.dz = c->p.dz;
    K.cx = K.cy = K.cz = 0.0;
    for (int a = 0; a < 3; a++) { K.N[a] = c->g.n[a]; K.per[a] = c->p.periodic[a]; }
    return K;
}

int launch_push64(pic_ctx* c, int s) {
    Species& S = c->sp[s];
    const int cur = S.cur;
    P64 P{S.x[cur], S.y[cur], S.z[cur], S.ux[cur], S.uy[cur], S.uz[cur], S.w[cur],
          c->keep_old ? S.xo[cur] : nullptr, c->keep_old ? S.yo[cur] : nullptr,
          c->keep_old ? S.zo[cur] : nullptr, S.d_cnt};
    Fields64 Fd{(const double*)c->F[0], (const double*)c->F[1], (const double*)c->F[2],
                (const double*)c->F[3], (const double*)c->F[4], (const double*)c->F[5],
                (double*)c->J[0], (double*)c->J[1], (double*)c->J[2], (double*)c->rho[s]};
    PushConsts K = make_consts(c, S);
    const int kernel = c->push_kernel == 0 ? 2 : c->push_kernel;
    if (kernel == 1) {
        PhaseTimer t(c, PIC_T_PUSH);
        int blocks = (int)std::min<int64_t>(blocks_for(std::max<int64_t>(S.cap, 1), 256), 256 * 64);
        if (c->g.m == 2) k_push64_direct<2><<<blocks, 256, 0, c->stream>>>(P, Fd, c->g, K);
        else k_push64_direct<4><<<blocks, 256, 0, c->stream>>>(P, Fd, c->g, K);
        PIC_LAUNCH_CHECK(c);
        return PIC_OK;
    }
    // tiled: particles grouped by tile (a stale grouping only costs speed:
    // deposits outside the tile's accumulator box go straight to HBM)
    {
        PhaseTimer t(c, PIC_T_SORT);
        if (!S.sorted || S.pushes_since_sort >= c->sort_interval) {
            int r = launch_sort(c, s);
            if (r) return r;
        }
        int r = build_work(c, s, PIC_PUSH_CHUNK);
        if (r) return r;
    }
    const int chunk = PIC_PUSH_CHUNK;
    PhaseTimer t(c, PIC_T_PUSH);
    S.pushes_since_sort++;
    const int64_t max_items = (int64_t)c->tl.ntiles + S.cap / chunk + 1;
    int64_t* work = c->work;
    int64_t* nwork = c->nwork;
    if (c->g.m == 2) {
        using B = AccBox<2, 8, 8, 8>;
        const size_t lds = 4 * B::VOL * sizeof(double);
        PIC_HIP(c, hipFuncSetAttribute((const void*)k_push64_tiled<2, 8, 8, 8>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        k_push64_tiled<2, 8, 8, 8><<<(unsigned)max_items, 256, lds, c->stream>>>(
            P, Fd, c->g, K, c->tl, S.tile_off, work, nwork, chunk);
    } else {
        using B = AccBox<4, 8, 8, 4>;
        const size_t lds = 4 * B::VOL * sizeof(double);
        PIC_HIP(c, hipFuncSetAttribute((const void*)k_push64_tiled<4, 8, 8, 4>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        k_push64_tiled<4, 8, 8, 4><<<(unsigned)max_items, 256, lds, c->stream>>>(
            P, Fd, c->g, K, c->tl, S.tile_off, work, nwork, chunk);
    }
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

}  // namespace pic
