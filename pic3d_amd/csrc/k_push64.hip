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
#include <type_traits>

#include "pic_internal.h"
#include "atomics.h"
#include "push64.h"
#include "shapes.h"
#include "wave.h"

namespace pic {

// Rolled loops over a small per-particle weight array read element 0 of a
// register copy and shift the copy down by one each iteration: only
// compile-time indices, so the arrays stay in registers (no scratch).
template <int L>
__device__ __forceinline__ void shift_down(double* A) {
#pragma unroll
    for (int t = 0; t < L - 1; t++) A[t] = A[t + 1];
}
template <int L>
__device__ __forceinline__ void copy_arr(const double* S, double* D) {
#pragma unroll
    for (int t = 0; t < L; t++) D[t] = S[t];
}

// Gather + Boris + advance for one particle.  Field reads go through the
// functor FR(comp, k, j, i).  Each component's sum runs k -> j -> i with
// terms ((Wx*Wy)*Wz)*F, the reference's order (particle.cpp:207-259); the
// k loop is rolled to keep the kernel inside the instruction cache.
// The primal weights at the old position and their centre nodes are
// returned in (Wxp, Wyp, Wzp, ip, jp, kp): the deposit's old-position
// weights are the same values (the same function of the same argument).
//
// LOWREG: the k loop recomputes its WX[i] * WY[j] products instead of
// holding the 25 (quartic) k-invariant ones across it (an empty asm
// statement makes the x weights opaque at each iteration, so they are not
// hoisted).  The same products, the same values: 50 fewer live registers,
// which lets the gather-only kernel run three waves per SIMD.
template <int M, typename FR, bool LOWREG = false>
__device__ __forceinline__ void gather_boris(double x, double y, double z, double& ux, double& uy,
                                             double& uz, double& xn, double& yn, double& zn,
                                             const PushConsts& K, int m, FR fr, double* Wxp, double* Wyp,
                                             double* Wzp, int& ip, int& jp, int& kp, double* ebq,
                                             double& gout) {
    constexpr int H = M / 2;
    const double org = (double)(1 - m);
    const double xc = x - org, yc = y - org, zc = z - org;
    double Wxd[M + 1], Wyd[M + 1], Wzd[M + 1];
    ip = shape_at<M>(xc, Wxp);
    jp = shape_at<M>(yc, Wyp);
    kp = shape_at<M>(zc, Wzp);
    shape_at<M>(xc - 0.5, Wxd);
    shape_at<M>(yc - 0.5, Wyd);
    shape_at<M>(zc - 0.5, Wzd);
    const int id = (int)xc, jd = (int)yc, kd = (int)zc;
    auto comp = [&](int c, const double* WX, const double* WY, const double* WZ, int b0, int b1,
                    int b2) -> double {
        double e = 0.0;
        double wzr[M + 1], wx[M + 1];
        copy_arr<M + 1>(WZ, wzr);
        copy_arr<M + 1>(WX, wx);
#pragma unroll 1
        for (int k = 0; k <= M; k++) {
            const double wz = wzr[0];
            if constexpr (LOWREG) {
#pragma unroll
                for (int t = 0; t <= M; t++) asm volatile("" : "+v"(wx[t]));
            }
#pragma unroll
            for (int j = 0; j <= M; j++)
#pragma unroll
                for (int i = 0; i <= M; i++) e += wx[i] * WY[j] * wz * fr(c, b2 + k - H, b1 + j - H, b0 + i - H);
            shift_down<M + 1>(wzr);
        }
        return e;
    };
    // two components with the same x and y weights in one k loop: their
    // WX[i] * WY[j] products are computed once per k (each sum keeps its
    // own order and terms)
    auto pair = [&](int c0, int c1, const double* WX, const double* WY, const double* WZ0, const double* WZ1,
                    int b0, int b1, int z0, int z1, double& e0, double& e1) {
        double s0 = 0.0, s1 = 0.0;
        double wz0[M + 1], wz1[M + 1], wx[M + 1];
        copy_arr<M + 1>(WZ0, wz0);
        copy_arr<M + 1>(WZ1, wz1);
        copy_arr<M + 1>(WX, wx);
#pragma unroll 1
        for (int k = 0; k <= M; k++) {
            const double za = wz0[0], zb = wz1[0];
#pragma unroll
            for (int t = 0; t <= M; t++) asm volatile("" : "+v"(wx[t]));
#pragma unroll
            for (int j = 0; j <= M; j++)
#pragma unroll
                for (int i = 0; i <= M; i++) {
                    const double xy = wx[i] * WY[j];
                    s0 += xy * za * fr(c0, z0 + k - H, b1 + j - H, b0 + i - H);
                    s1 += xy * zb * fr(c1, z1 + k - H, b1 + j - H, b0 + i - H);
                }
            shift_down<M + 1>(wz0);
            shift_down<M + 1>(wz1);
        }
        e0 = s0;
        e1 = s1;
    };
    double e0, e1, e2, b0, b1, b2;
    if constexpr (LOWREG) {
        pair(0, 4, Wxd, Wyp, Wzp, Wzd, id, jp, kp, kd, e0, b1);  // Ex, By
        pair(1, 3, Wxp, Wyd, Wzp, Wzd, ip, jd, kp, kd, e1, b0);  // Ey, Bx
        e2 = comp(2, Wxp, Wyp, Wzd, ip, jp, kd);
        b2 = comp(5, Wxd, Wyd, Wzp, id, jd, kp);
    } else {
        e0 = comp(0, Wxd, Wyp, Wzp, id, jp, kp);
        e1 = comp(1, Wxp, Wyd, Wzp, ip, jd, kp);
        e2 = comp(2, Wxp, Wyp, Wzd, ip, jp, kd);
        b0 = comp(3, Wxp, Wyd, Wzd, ip, jd, kd);
        b1 = comp(4, Wxd, Wyp, Wzd, id, jp, kd);
        b2 = comp(5, Wxd, Wyd, Wzp, id, jd, kp);
    }
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
    ebq[0] = e0; ebq[1] = e1; ebq[2] = e2; ebq[3] = b0; ebq[4] = b1; ebq[5] = b2;
    gout = gam;
    // particle.cpp:313-315
    xn = x + ux / gam * K.dt / K.dx;
    yn = y + uy / gam * K.dt / K.dy;
    zn = z + uz / gam * K.dt / K.dz;
}

// Spin precession of slot p after its push (particle.cpp:318-368); the
// species' mass is the electron mass (checked on the host).  Returns 1
// where g_factor has no value (the reference exits there).
__device__ __forceinline__ int spin_step(const P64& P, int64_t p, const double* ebq, double ux, double uy,
                                         double uz, double gam, const PushConsts& K) {
    double sx = P.sx[p], sy = P.sy[p], sz = P.sz[p];
    // particle.cpp:321-326: the fields divided back by q/m dt/2
    const double E[3] = {ebq[0] / K.qm, ebq[1] / K.qm, ebq[2] / K.qm};
    const double B[3] = {ebq[3] / K.qm, ebq[4] / K.qm, ebq[5] / K.qm};
    if (!spin_precess<double>(E, B, K.qm, ux, uy, uz, gam, K.spin, sx, sy, sz)) return 1;
    P.sx[p] = sx; P.sy[p] = sy; P.sz[p] = sz;
    return 0;
}

// Esirkepov deposit of one particle moving (xo,yo,zo) -> (xn,yn,zn)
// (particle.cpp:381-514).  The sink D receives D.base(bz, by, bx) once
// (the box origin, buffer indices) and then D.put(comp, k, j, i, a, b) with
// box-relative compile-time (k, j, i); comp 0..2 = J, 3 = rho.  The
// contribution is the product a * b, where a already carries the sink's
// scale D.scale(comp) (1 for a sink that adds values, 1/unit for one that
// adds fixed-point integers).
//
// The reference walks a box [min(i1,i2)-H, max(i1,i2)+H] per axis (M+1 or
// M+2 points).  Here the box is always L = M+2 points with the old/new
// shape weights placed at compile-time offsets (selected by d0/d1 in {0,1}),
// so the loops unroll into registers with immediate LDS offsets.  Callers
// guarantee |x_new - x_old| < 1 cell per axis, hence |i2 - i1| <= 1.  The
// shape weights and the order of each cumulative sum are the reference's;
// the transverse row weights are regrouped (below), so a term can differ
// from the reference's in its last bits (J and rho are compared at 1e-12:
// the order of the atomic adds differs anyway).  Points of the L-box skipped:
// (a) planes outside the reference's box along the non-cumulative axes,
// whose weights are exactly zero (adding +-0 changes no value), and (b) the
// last point of the reference's box along the cumulative axis (box point
// L-1 if the particle changes node along that axis, else L-2).  There the
// cumulative sum is sum(S1) - sum(S0), zero in exact arithmetic since both
// shapes sum to one: the reference adds a rounding residue of about 1e-16
// of the particle's contribution at that point, which is left out here (one
// LDS atomic per row saved).
template <int M>
__device__ __forceinline__ void place(const double* S, int d, double* A) {
#pragma unroll
    for (int t = 0; t < M + 2; t++) {
        const double a0 = (t <= M) ? S[t] : 0.0;
        const double a1 = (t >= 1) ? S[t - 1] : 0.0;
        A[t] = d ? a1 : a0;
    }
}

// deposit with the old-position weights (Sx0, Sy0, Sz0 at nodes i1, j1, k1)
// given.  STILL: the caller knows that the particle keeps its nearest node
// along every axis (i2 == i1, j2 == j1, k2 == k1); the box's last line along
// each axis then carries only zeros and is left out at compile time (the
// same contributions as the general form).
template <int M, typename D, bool STILL = false>
__device__ __forceinline__ void esirkepov_s0(double qw, const double* Sx0, const double* Sy0,
                                             const double* Sz0, int i1, int j1, int k1, double xn,
                                             double yn, double zn, const PushConsts& K, int m, D& dst) {
    constexpr int H = M / 2, L = M + 2;
    constexpr int LR = STILL ? L - 1 : L;  // box lines that can carry weight
    const double org = (double)(1 - m);
    double Sx1[M + 1], Sy1[M + 1], Sz1[M + 1];
    const int i2 = STILL ? (shape_at<M>(xn - org, Sx1), i1) : shape_at<M>(xn - org, Sx1);
    const int j2 = STILL ? (shape_at<M>(yn - org, Sy1), j1) : shape_at<M>(yn - org, Sy1);
    const int k2 = STILL ? (shape_at<M>(zn - org, Sz1), k1) : shape_at<M>(zn - org, Sz1);
    const int bx = min(i1, i2) - H, by = min(j1, j2) - H, bz = min(k1, k2) - H;
    const bool crx = !STILL && i1 != i2, cry = !STILL && j1 != j2, crz = !STILL && k1 != k2;
    double X0[L], X1[L], Y0[L], Y1[L], Z0[L], Z1[L];
    place<M>(Sx0, i1 - H - bx, X0); place<M>(Sx1, i2 - H - bx, X1);
    place<M>(Sy0, j1 - H - by, Y0); place<M>(Sy1, j2 - H - by, Y1);
    place<M>(Sz0, k1 - H - bz, Z0); place<M>(Sz1, k2 - H - bz, Z1);
    const double cx = qw * K.dx / K.dt, cy = qw * K.dy / K.dt, cz = qw * K.dz / K.dt;
    // the cumulative sums along each axis (the reference's order) times the
    // current scale: the reference's term -(c * acc * w) is (-(c * acc)) * w
    double CX[L - 1], CY[L - 1], CZ[L - 1];
    {
        const double sx = cx * dst.scale(0), sy = cy * dst.scale(1), sz = cz * dst.scale(2);
        double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
        for (int t = 0; t < L - 1; t++) {
            ax += X1[t] - X0[t];
            ay += Y1[t] - Y0[t];
            az += Z1[t] - Z0[t];
            CX[t] = -(sx * ax);
            CY[t] = -(sy * ay);
            CZ[t] = -(sz * az);
        }
    }
    // Transverse row weights: the reference's s0 t0 + ds t0 / 2 + s0 dt / 2 +
    // ds dt / 3 (particle.cpp:432-436, s along one transverse axis, t along
    // the other, ds and dt the new-minus-old differences) regrouped as
    // s0 * (t0 + dt / 2) + ds * (t0 / 2 + dt / 3), the two factors in
    // parentheses precomputed per axis point (P, Q): two operations per row
    // instead of ten.  A row outside both supports still gets exactly 0.
    double PX[L], QX[L], PY[L], QY[L], DY[L], DZ[L];
#pragma unroll
    for (int t = 0; t < L; t++) {
        const double dx_ = X1[t] - X0[t], dy_ = Y1[t] - Y0[t], dz_ = Z1[t] - Z0[t];
        PX[t] = X0[t] + 0.5 * dx_; QX[t] = 0.5 * X0[t] + (1.0 / 3.0) * dx_;
        PY[t] = Y0[t] + 0.5 * dy_; QY[t] = 0.5 * Y0[t] + (1.0 / 3.0) * dy_;
        DY[t] = dy_; DZ[t] = dz_;
    }
    dst.base(bz, by, bx);
    // Jx: cumulative sum along x, weights W_yz (particle.cpp:427-450)
    double za[L], zb[L];
    copy_arr<L>(Z0, za); copy_arr<L>(DZ, zb);
#pragma unroll 1
    for (int k = 0; k < LR; k++) {
        const double sz0 = za[0], dsz = zb[0];
        shift_down<L>(za); shift_down<L>(zb);
#pragma unroll
        for (int j = 0; j < LR; j++) {
            const double wyz = __builtin_fma(dsz, QY[j], sz0 * PY[j]);
            if (wyz == 0.0) continue;  // a plane outside both supports: only +-0 to add
#pragma unroll
            for (int i = 0; i < L - 1; i++)
                if (i < L - 2 || crx) dst.put(0, k, j, i, CX[i], wyz);
        }
    }
    // Jy (particle.cpp:452-475)
    copy_arr<L>(Z0, za); copy_arr<L>(DZ, zb);
#pragma unroll 1
    for (int k = 0; k < LR; k++) {
        const double sz0 = za[0], dsz = zb[0];
        shift_down<L>(za); shift_down<L>(zb);
#pragma unroll
        for (int i = 0; i < LR; i++) {
            const double wxz = __builtin_fma(dsz, QX[i], sz0 * PX[i]);
            if (wxz == 0.0) continue;
#pragma unroll
            for (int j = 0; j < L - 1; j++)
                if (j < L - 2 || cry) dst.put(1, k, j, i, CY[j], wxz);
        }
    }
    // Jz (particle.cpp:477-500)
    double ya[L], yb[L];
    copy_arr<L>(Y0, ya); copy_arr<L>(DY, yb);
#pragma unroll 1
    for (int j = 0; j < LR; j++) {
        const double sy0 = ya[0], dsy = yb[0];
        shift_down<L>(ya); shift_down<L>(yb);
#pragma unroll
        for (int i = 0; i < LR; i++) {
            const double wxy = __builtin_fma(dsy, QX[i], sy0 * PX[i]);
            if (wxy == 0.0) continue;
#pragma unroll
            for (int k = 0; k < L - 1; k++)
                if (k < L - 2 || crz) dst.put(2, k, j, i, CZ[k], wxy);
        }
    }
    // rho at the new position (particle.cpp:502-514): nodes i2-H .. i2+H.
    // The reference skips planes whose weight is exactly zero; adding +-0
    // there changes nothing, so every node is deposited.
    // The term is ((qw * Sx) * Sy) * Sz.
    dst.base(k2 - H, j2 - H, i2 - H);
    double rx[M + 1], zr[M + 1];
    const double qs = qw * dst.scale(3);
#pragma unroll
    for (int i = 0; i <= M; i++) rx[i] = qs * Sx1[i];
    copy_arr<M + 1>(Sz1, zr);
#pragma unroll 1
    for (int k = 0; k <= M; k++) {
        const double sz = zr[0];
        shift_down<M + 1>(zr);
#pragma unroll
        for (int j = 0; j <= M; j++)
#pragma unroll
            for (int i = 0; i <= M; i++) dst.put(3, k, j, i, rx[i] * Sy1[j], sz);
    }
}

template <int M, typename D, bool STILL = false>
__device__ __forceinline__ void esirkepov(double qw, double xo, double yo, double zo, double xn,
                                          double yn, double zn, const PushConsts& K, int m, D& dst) {
    const double org = (double)(1 - m);
    double Sx0[M + 1], Sy0[M + 1], Sz0[M + 1];
    const int i1 = shape_at<M>(xo - org, Sx0), j1 = shape_at<M>(yo - org, Sy0), k1 = shape_at<M>(zo - org, Sz0);
    esirkepov_s0<M, D, STILL>(qw, Sx0, Sy0, Sz0, i1, j1, k1, xn, yn, zn, K, m, dst);
}

// deposit sinks ---------------------------------------------------------
struct HbmSink {  // global FP64 atomics on the ghosted J/rho arrays
    double* jp[4];
    Grid g;
    int bz, by, bx;
    __device__ __forceinline__ double scale(int) const { return 1.0; }
    __device__ __forceinline__ void base(int z, int y, int x) { bz = z; by = y; bx = x; }
    __device__ __forceinline__ void put(int c, int k, int j, int i, double a, double b) {
        hbm_add(&jp[c][g.ji(bz + k, by + j, bx + i)], a * b);
    }
};

// LDS accumulator box with origin (oz, oy, ox), int64 fixed point: each
// contribution a * b (a scaled by 1/unit[c]) becomes the nearest integer:
// |a * b| <= 2^48 units, so one fused a * b + 1.5 * 2^52 rounds it into the
// low mantissa bits, and subtracting that constant's bit pattern leaves the
// two's-complement integer.  Added with ds_add_u64 (measured about 20 %
// cheaper per LDS instruction than ds_add_f64); a chunk of at most 2^12
// particles sums to below 2^60 units.
template <typename BOX>
struct LdsSink {
    unsigned long long* acc;
    const double* inv_unit;
    int oz, oy, ox;
    int off;
    __device__ __forceinline__ double scale(int c) const { return inv_unit[c]; }
    __device__ __forceinline__ void base(int z, int y, int x) {
        off = (z - oz) * BOX::PXY + (y - oy) * BOX::PX + (x - ox);
    }
    __device__ __forceinline__ void put(int c, int k, int j, int i, double a, double b) {
        const double t = __builtin_fma(a, b, 0x1.8p52);
        const unsigned long long q =
            (unsigned long long)__double_as_longlong(t) - 0x4338000000000000ull;  // bits of 1.5 * 2^52
        lds_add_u64(&acc[off + BOX::aoff(c) + k * BOX::PXY + j * BOX::PX + i], q);
    }
};

// Full push of one particle against HBM fields and HBM currents (the
// direct kernel, and the tiled kernel's path for a particle outside its
// tile).  Returns 1 if the particle was skipped as out of bounds.
template <int M, bool SPIN, bool LOWREG = false>
__device__ __forceinline__ int push_one_hbm(const P64& P, int64_t p, double wp, const Fields64& F,
                                         const Grid& g, const PushConsts& K, int& spin_bad) {
    double x = P.x[p], y = P.y[p], z = P.z[p];
    if (!wrap_check(x, y, z, K)) return 1;
    const double* fp[6] = {F.ex, F.ey, F.ez, F.bx, F.by, F.bz};
    auto fr = [&](int c, int k, int j, int i) -> double { return fp[c][g.fi(k, j, i)]; };
    double ux = P.ux[p], uy = P.uy[p], uz = P.uz[p], xn, yn, zn;
    double Sx0[M + 1], Sy0[M + 1], Sz0[M + 1];
    int i1, j1, k1;
    double ebq[6], gam;
    gather_boris<M, decltype(fr), LOWREG>(x, y, z, ux, uy, uz, xn, yn, zn, K, g.m, fr, Sx0, Sy0, Sz0, i1, j1, k1,
                                          ebq, gam);
    P.x[p] = xn; P.y[p] = yn; P.z[p] = zn;
    P.ux[p] = ux; P.uy[p] = uy; P.uz[p] = uz;
    if (P.xo) { P.xo[p] = x; P.yo[p] = y; P.zo[p] = z; }
    if (SPIN) spin_bad += spin_step(P, p, ebq, ux, uy, uz, gam, K);
    // a non-finite momentum would put the deposit box outside the arrays
    if (!(fabs(xn - x) < 1.0 && fabs(yn - y) < 1.0 && fabs(zn - z) < 1.0)) return 1;
    HbmSink d{{F.jx, F.jy, F.jz, F.rho}, g, 0, 0, 0};
    esirkepov_s0<M>(K.q * wp, Sx0, Sy0, Sz0, i1, j1, k1, xn, yn, zn, K, g.m, d);
    return 0;
}

// ---------------------------------------------------------------------
// currents_particle (particle.cpp:372-515) alone: the Esirkepov deposit of
// particles whose old and new positions are given, through HBM atomics.
// A particle is skipped (and counted) when its old position lies outside
// [0, N] on an axis or it moved a cell or more: its boxes would leave the
// ghosted arrays.
// ---------------------------------------------------------------------
template <int M>
__global__ __launch_bounds__(256) void k_deposit64(const double* __restrict__ xo,
                                                   const double* __restrict__ yo,
                                                   const double* __restrict__ zo,
                                                   const double* __restrict__ xn,
                                                   const double* __restrict__ yn,
                                                   const double* __restrict__ zn,
                                                   const double* __restrict__ w, int64_t n,
                                                   Fields64 F, Grid g, PushConsts K,
                                                   int64_t* __restrict__ skipped) {
    int64_t bad = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        const double x = xo[p], y = yo[p], z = zo[p];
        const bool inside = x >= 0.0 && x <= (double)K.N[0] && y >= 0.0 && y <= (double)K.N[1] &&
                            z >= 0.0 && z <= (double)K.N[2];
        if (!inside || !(fabs(xn[p] - x) < 1.0 && fabs(yn[p] - y) < 1.0 && fabs(zn[p] - z) < 1.0)) {
            bad++;
            continue;
        }
        HbmSink d{{F.jx, F.jy, F.jz, F.rho}, g, 0, 0, 0};
        esirkepov<M>(K.q * w[p], x, y, z, xn[p], yn[p], zn[p], K, g.m, d);
    }
    if (bad) atomicAdd((unsigned long long*)skipped, (unsigned long long)bad);
}

// ---------------------------------------------------------------------
// direct kernel: any particle order, global atomics
// ---------------------------------------------------------------------
template <int M, bool SPIN>
__global__ __launch_bounds__(256) void k_push64_direct(P64 P, Fields64 F, Grid g, PushConsts K) {
    const int64_t n = P.cnt[CNT_SLOTS];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t oob = 0;
    int spin_bad = 0;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        const double wp = P.w[p];
        if (!(wp == wp)) continue;
        oob += push_one_hbm<M, SPIN>(P, p, wp, F, g, K, spin_bad);
    }
    if (oob) atomicAdd((unsigned long long*)&P.cnt[CNT_OOB], (unsigned long long)oob);
    if (SPIN && spin_bad) atomicAdd((unsigned long long*)&P.cnt[CNT_SPIN], (unsigned long long)spin_bad);
}

// ---------------------------------------------------------------------
// tiled kernel: particles sorted by tile; J and rho accumulated in LDS
// ---------------------------------------------------------------------
// LDS accumulator box for a tile whose first cell is (T0x, T0y, T0z) with
// edges (tx, ty, tz): buffer indices [T0 + m - 2 - H, T0 + t - 1 + m + 1 + H],
// i.e. t + m + 3 points per axis (particles move < 1 cell per step).
//
// LDS bank layout.  LDS serves the 8-byte reads and atomics of 16 lanes at
// a time over 16 bank pairs (8-byte word index mod 16).  Each wave's
// particles are dealt over its lanes so that the 16 lanes of a group have
// deposit bases on distinct bank pairs (bank_spread_lane).  The field tile
// uses row and plane pitches congruent mod 16 to the accumulator box's, so
// the same dealing also puts the gather's primal bases (the deposit base
// for a particle that stays on its nodes) on distinct bank pairs; only the
// half-cell shifts of the staggered components remain.  Each component's
// field box is one point shorter along its staggered (dual) axes, each
// current box one point shorter along its cumulative axis, and boxes are
// packed back to back, so two workgroups of the quartic kernel fit a CU.
__host__ __device__ constexpr int pitch16(int at_least, int residue) {
    return at_least + ((residue - at_least) % 16 + 16) % 16;
}
template <int M, int TX, int TY, int TZ>
struct AccBox {
    static constexpr int H = M / 2;
    // accumulator box: tile + deposit halo, x pitch 1, row pitch BX, plane pitch BX * BY
    static constexpr int BX = TX + M + 3, BY = TY + M + 3, BZ = TZ + M + 3;
    static constexpr int PX = BX, PXY = BX * BY;
    // extents per component (Jx, Jy, Jz one shorter along their cumulative axis; rho)
    static constexpr int asize(int c) {
        return (BZ - (c == 2) - 1) * PXY + (BY - (c == 1) - 1) * PX + (BX - (c == 0));
    }
    static constexpr int aoff(int c) { return c == 0 ? 0 : aoff(c - 1) + asize(c - 1); }
    static constexpr int AVOL = aoff(4);
    // field tile: buffer indices [T0 + m - 1 - H, T0 + T + m - 1 + H] on a
    // primal axis, one fewer on a dual axis.  Staggering (x, y, z) per
    // component: Ex (d,p,p) Ey (p,d,p) Ez (p,p,d) Bx (p,d,d) By (d,p,d) Bz (d,d,p)
    static constexpr int FX = TX + M + 1, FY = TY + M + 1, FZ = TZ + M + 1;
    static constexpr int dual(int c, int axis) {
        return (c == 0 && axis == 0) || (c == 1 && axis == 1) || (c == 2 && axis == 2) ||
               (c == 3 && axis != 0) || (c == 4 && axis != 1) || (c == 5 && axis != 2);
    }
    static constexpr int fext(int c, int axis) { return (axis == 0 ? FX : axis == 1 ? FY : FZ) - dual(c, axis); }
    static constexpr int FPY = pitch16(FX, PX);                                   // row pitch
    static constexpr int fpz(int c) { return pitch16(FPY * fext(c, 1), PXY); }  // plane pitch
    static constexpr int fsize(int c) { return (fext(c, 2) - 1) * fpz(c) + (fext(c, 1) - 1) * FPY + fext(c, 0); }
    static constexpr int foff(int c) { return c == 0 ? 0 : foff(c - 1) + fsize(c - 1); }
    static constexpr int FVOL = foff(6);
    static constexpr size_t LDS = (size_t)(FVOL + AVOL) * sizeof(double);
};
static_assert(2 * (AccBox<4, 4, 4, 4>::LDS + 256) <= 160 * 1024, "two quartic workgroups per CU");

template <int M, int TX, int TY, int TZ, bool SPIN>
__global__ __launch_bounds__(256, 2) void k_push64_tiled(P64 P, Fields64 F, Grid g, PushConsts K,
                                                      Tiling tl, const int64_t* __restrict__ tile_off,
                                                      const int64_t* __restrict__ work,
                                                      const int64_t* __restrict__ nwork, int chunk) {
    using B = AccBox<M, TX, TY, TZ>;
    extern __shared__ double lds64[];
    double* fld = lds64;  // FVOL: the tile's fields
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(lds64 + B::FVOL);  // AVOL: J, rho
    const int64_t wi = blockIdx.x;
    if (wi >= *nwork) return;
    const int64_t item = work[wi];            // tile << 20 | chunk index
    const int tile = (int)(item >> 20);
    const int64_t cidx = item & ((1 << 20) - 1);
    const int64_t t_beg = tile_off[tile], t_end = tile_off[tile + 1];
    const int64_t beg = t_beg + cidx * chunk;
    const int64_t end = min(t_end, beg + chunk);
    const int tzi = tile / (tl.nt[0] * tl.nt[1]);
    const int tyi = (tile / tl.nt[0]) % tl.nt[1];
    const int txi = tile % tl.nt[0];
    const int T0x = txi * TX, T0y = tyi * TY, T0z = tzi * TZ;
    const double* fp[6] = {F.ex, F.ey, F.ez, F.bx, F.by, F.bz};
    const int fox = T0x + g.m - 1 - B::H, foy = T0y + g.m - 1 - B::H, foz = T0z + g.m - 1 - B::H;
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const int ex = B::fext(c, 0), ey = B::fext(c, 1), n = ex * ey * B::fext(c, 2);
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const int lx = e % ex, ly = (e / ex) % ey, lz = e / (ex * ey);
            // reads past a partial last tile are clamped (never used)
            const int gx = min(fox + lx, g.fe[0] - 1), gy = min(foy + ly, g.fe[1] - 1), gz = min(foz + lz, g.fe[2] - 1);
            fld[B::foff(c) + lz * B::fpz(c) + ly * B::FPY + lx] = fp[c][g.fi(gz, gy, gx)];
        }
    }
    for (int e = threadIdx.x; e < B::AVOL; e += blockDim.x) acc[e] = 0ull;
    __shared__ unsigned rg_hist[256 / 64][8];  // per wave: bank-spreading counters
    if (lane_id() < 8) rg_hist[threadIdx.x >> 6][lane_id()] = 0u;
    __syncthreads();
    // gather from the LDS field tile (an in-tile particle's stencil lies in it)
    auto fr = [&](int c, int k, int j, int i) -> double {
        return fld[B::foff(c) + (k - foz) * B::fpz(c) + (j - foy) * B::FPY + (i - fox)];
    };
    LdsSink<B> sink{acc, K.inv_unit, T0z + g.m - 2 - B::H, T0y + g.m - 2 - B::H, T0x + g.m - 2 - B::H, 0};
    int64_t oob = 0;
    int spin_bad = 0;
    uint32_t strays = 0;  // iterations whose particle was outside the tile
    int it = 0;
    const int len = (int)(end - beg);
    // Each wave takes 64 consecutive slots (coalesced loads), then deals its
    // particles out again over its lanes by the LDS bank of their deposit
    // base (bank_spread_lane); results go back to the particles' own slots.
    const int wv = threadIdx.x >> 6, lane = lane_id();
    for (int t0 = threadIdx.x & ~63; t0 < len; t0 += blockDim.x, it++) {
        const int t = t0 + lane;
        double wp = 0.0, x = 0.0, y = 0.0, z = 0.0, ux = 0.0, uy = 0.0, uz = 0.0;
        bool ok = false;
        int r = 32;
        if (t < len) {
            const int64_t p = beg + t;
            wp = P.w[p];
            x = P.x[p]; y = P.y[p]; z = P.z[p];
            ok = wp == wp;
            if (ok && !wrap_check(x, y, z, K)) { oob++; ok = false; }
            if (ok) {
                // cell after the wrap, from the shape functions' own argument
                // x + m - 1 (rounded: x just below a cell edge can land on
                // it), so that the dual stencil (which starts at its floor)
                // stays in the tile's field box.  A particle exactly at
                // x == N (kept by the strict wrap test) or rounded onto the
                // next tile's first cell is sorted into this tile but takes
                // the HBM pass like a particle outside its tile.
                const double org = (double)(1 - g.m);
                const int cx = (int)(x - org) - (g.m - 1), cy = (int)(y - org) - (g.m - 1),
                          cz = (int)(z - org) - (g.m - 1);
                if ((unsigned)(cx - T0x) >= (unsigned)TX || (unsigned)(cy - T0y) >= (unsigned)TY ||
                    (unsigned)(cz - T0z) >= (unsigned)TZ) {
                    strays |= 1u << it;  // left its tile since the last sort: HBM pass below
                    ok = false;
                }
            }
            if (ok) {
                ux = P.ux[p]; uy = P.uy[p]; uz = P.uz[p];
                // deposit box base: the nearest node of the old position, in 8-byte words
                const double org = (double)(1 - g.m);
                const int ib = (int)(x - org + 0.5) - B::H, jb = (int)(y - org + 0.5) - B::H,
                          kb = (int)(z - org + 0.5) - B::H;
                const int ofs = (kb - (T0z + g.m - 2 - B::H)) * B::PXY + (jb - (T0y + g.m - 2 - B::H)) * B::PX +
                                (ib - (T0x + g.m - 2 - B::H));
                r = (2 * ofs) & 31;
            }
        }
        const int to = bank_spread_lane(r, rg_hist[wv]);
        const int okv = lane_send(to, ok ? 1 : 0);
        const int tt = lane_send(to, t);
        x = lane_send_d(to, x); y = lane_send_d(to, y); z = lane_send_d(to, z);
        ux = lane_send_d(to, ux); uy = lane_send_d(to, uy); uz = lane_send_d(to, uz);
        wp = lane_send_d(to, wp);
        if (!okv) continue;
        const int64_t p = beg + tt;
        double xn, yn, zn;
        double Sx0[M + 1], Sy0[M + 1], Sz0[M + 1];
        int i1, j1, k1;
        double ebq[6], gam;
        gather_boris<M>(x, y, z, ux, uy, uz, xn, yn, zn, K, g.m, fr, Sx0, Sy0, Sz0, i1, j1, k1, ebq, gam);
        P.x[p] = xn; P.y[p] = yn; P.z[p] = zn;
        P.ux[p] = ux; P.uy[p] = uy; P.uz[p] = uz;
        if (P.xo) { P.xo[p] = x; P.yo[p] = y; P.zo[p] = z; }
        if (SPIN) spin_bad += spin_step(P, p, ebq, ux, uy, uz, gam, K);
        if (!(fabs(xn - x) < 1.0 && fabs(yn - y) < 1.0 && fabs(zn - z) < 1.0)) { oob++; continue; }
        // in-tile particle, move < 1 cell: its whole box lies in the LDS box
        esirkepov_s0<M>(K.q * wp, Sx0, Sy0, Sz0, i1, j1, k1, xn, yn, zn, K, g.m, sink);
    }
    if (strays) {
        it = 0;
        for (int t = threadIdx.x; t < len; t += blockDim.x, it++)
            if (strays & (1u << it)) {
                const int64_t p = beg + t;
                oob += push_one_hbm<M, SPIN>(P, p, P.w[p], F, g, K, spin_bad);
            }
    }
    __syncthreads();
    // flush: one global atomic per nonzero accumulator point
    double* jp[4] = {F.jx, F.jy, F.jz, F.rho};
    const int ox = T0x + g.m - 2 - B::H, oy = T0y + g.m - 2 - B::H, oz = T0z + g.m - 2 - B::H;
    for (int e = threadIdx.x; e < B::AVOL; e += blockDim.x) {
        const long long q = (long long)acc[e];
        if (q == 0) continue;
        const int c = e >= B::aoff(3) ? 3 : e >= B::aoff(2) ? 2 : e >= B::aoff(1) ? 1 : 0;
        const int r = e - (c == 3 ? B::aoff(3) : c == 2 ? B::aoff(2) : c == 1 ? B::aoff(1) : 0);
        const double v = (double)q * K.unit[c];
        const int lx = r % B::PX, ly = (r / B::PX) % B::BY, lz = r / B::PXY;
        const int gx = ox + lx, gy = oy + ly, gz = oz + lz;
        // a partial last tile's box reaches past the arrays; those points
        // are never deposited into, but keep the flush in bounds regardless
        if (gx > g.je[0] - 2 || gy > g.je[1] - 2 || gz > g.je[2] - 2) continue;
        hbm_add(&jp[c][g.ji(gz, gy, gx)], v);
    }
    if (oob) atomicAdd((unsigned long long*)&P.cnt[CNT_OOB], (unsigned long long)oob);
    if (SPIN && spin_bad) atomicAdd((unsigned long long*)&P.cnt[CNT_SPIN], (unsigned long long)spin_bad);
}

// ---------------------------------------------------------------------
// Two-kernel push over cell-ordered tiles (the sort keeps each tile's
// particles in cell order).
//  k_push64_gather: gather, Boris and advance of a tile chunk's particles
//   in storage order, the field tile alone in LDS.  The 64 lanes of a wave
//   hold the particles of a few neighbouring cells, so most of a lane's
//   field reads fall on the LDS words its cell-mates read (one broadcast
//   read, no bank conflict).  It writes the particle state back and, for
//   every particle the deposit kernel is to deposit, its wrapped old
//   position (dxo, dyo, dzo; NaN: nothing to deposit; dxo negative: the
//   particle changed its nearest node along some axis).  A particle outside
//   its tile takes the HBM path of k_push64_tiled, deposit included.
//  k_push64_deposit: the Esirkepov deposit of those particles into the
//   accumulator box alone in LDS.  The chunk's particles are listed in two
//   classes, those that kept their nearest node along every axis (a box
//   one line shorter per axis, fixed at compile time) and the rest.  Each
//   list is deposited in passes of 256 where thread t takes entry
//   spread_pick(t), so a wave's lanes sit in different cells, then the
//   particles are dealt by the bank of their deposit base.  A wave runs the
//   union of its lanes' box lines, so a class of its own for the still
//   particles saves the lines only moving neighbours need.
// Each kernel holds half of the one-kernel form's LDS and fewer registers,
// so more workgroups share a CU.  The particle arithmetic is the same code:
// the deposit's old-position weights come from the same shape function of
// the same argument.
// ---------------------------------------------------------------------
template <int M, int TX, int TY, int TZ, bool SPIN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_push64_gather(
    P64 P, Fields64 F, Grid g, PushConsts K, Tiling tl, const int64_t* __restrict__ tile_off,
    const int64_t* __restrict__ work, const int64_t* __restrict__ nwork, int chunk, double* __restrict__ dxo,
    double* __restrict__ dyo, double* __restrict__ dzo) {
    using B = AccBox<M, TX, TY, TZ>;
    extern __shared__ double lds64[];
    double* fld = lds64;  // FVOL: the tile's fields
    const int64_t wi = blockIdx.x;
    if (wi >= *nwork) return;
    const int64_t item = work[wi];
    const int tile = (int)(item >> 20);
    const int64_t cidx = item & ((1 << 20) - 1);
    const int64_t beg = tile_off[tile] + cidx * chunk;
    const int64_t end = min(tile_off[tile + 1], beg + chunk);
    const int tzi = tile / (tl.nt[0] * tl.nt[1]);
    const int tyi = (tile / tl.nt[0]) % tl.nt[1];
    const int txi = tile % tl.nt[0];
    const int T0x = txi * TX, T0y = tyi * TY, T0z = tzi * TZ;
    const double* fp[6] = {F.ex, F.ey, F.ez, F.bx, F.by, F.bz};
    const int fox = T0x + g.m - 1 - B::H, foy = T0y + g.m - 1 - B::H, foz = T0z + g.m - 1 - B::H;
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const int ex = B::fext(c, 0), ey = B::fext(c, 1), n = ex * ey * B::fext(c, 2);
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const int lx = e % ex, ly = (e / ex) % ey, lz = e / (ex * ey);
            const int gx = min(fox + lx, g.fe[0] - 1), gy = min(foy + ly, g.fe[1] - 1), gz = min(foz + lz, g.fe[2] - 1);
            fld[B::foff(c) + lz * B::fpz(c) + ly * B::FPY + lx] = fp[c][g.fi(gz, gy, gx)];
        }
    }
    __syncthreads();
    auto fr = [&](int c, int k, int j, int i) -> double {
        return fld[B::foff(c) + (k - foz) * B::fpz(c) + (j - foy) * B::FPY + (i - fox)];
    };
    int64_t oob = 0;
    int spin_bad = 0;
    uint32_t strays = 0;  // passes whose particle was outside the tile
    const int len = (int)(end - beg);
    const double org = (double)(1 - g.m);
    const double nan = __longlong_as_double(0x7ff8000000000000ll);
    int it = 0;
    for (int t0 = 0; t0 < len; t0 += blockDim.x, it++) {
        const int t = t0 + threadIdx.x;
        if (t >= len) continue;
        const int64_t p = beg + t;
        const double wp = P.w[p];
        double x = P.x[p], y = P.y[p], z = P.z[p];
        double ox = nan, oy = nan, oz = nan;
        bool ok = wp == wp;
        if (ok && !wrap_check(x, y, z, K)) { oob++; ok = false; }
        if (ok) {
            // the cell from the shape functions' own argument (k_push64_tiled)
            const int cx = (int)(x - org) - (g.m - 1), cy = (int)(y - org) - (g.m - 1),
                      cz = (int)(z - org) - (g.m - 1);
            if ((unsigned)(cx - T0x) >= (unsigned)TX || (unsigned)(cy - T0y) >= (unsigned)TY ||
                (unsigned)(cz - T0z) >= (unsigned)TZ) {
                strays |= 1u << it;  // left its tile since the last sort: HBM pass below
                ok = false;
            }
        }
        if (ok) {
            double ux = P.ux[p], uy = P.uy[p], uz = P.uz[p], xn, yn, zn;
            double Sx0[M + 1], Sy0[M + 1], Sz0[M + 1];
            int i1, j1, k1;
            double ebq[6], gam;
            gather_boris<M, decltype(fr), true>(x, y, z, ux, uy, uz, xn, yn, zn, K, g.m, fr, Sx0, Sy0, Sz0, i1, j1,
                                                k1, ebq, gam);
            P.x[p] = xn; P.y[p] = yn; P.z[p] = zn;
            P.ux[p] = ux; P.uy[p] = uy; P.uz[p] = uz;
            if (P.xo) { P.xo[p] = x; P.yo[p] = y; P.zo[p] = z; }
            if (SPIN) spin_bad += spin_step(P, p, ebq, ux, uy, uz, gam, K);
            // a non-finite momentum would put the deposit box outside the arrays
            if (fabs(xn - x) < 1.0 && fabs(yn - y) < 1.0 && fabs(zn - z) < 1.0) {
                // sign of dxo: the particle changed its nearest node (x >= 0 here)
                const bool moved = i1 != (int)(xn - org + 0.5) || j1 != (int)(yn - org + 0.5) ||
                                   k1 != (int)(zn - org + 0.5);
                ox = moved ? -fabs(x) : fabs(x); oy = y; oz = z;
            } else {
                oob++;
            }
        }
        dxo[p] = ox; dyo[p] = oy; dzo[p] = oz;
    }
    if (strays) {
        it = 0;
        for (int t = threadIdx.x; t < len; t += blockDim.x, it++)
            if (strays & (1u << it)) {
                const int64_t p = beg + t;
                oob += push_one_hbm<M, SPIN, true>(P, p, P.w[p], F, g, K, spin_bad);
            }
    }
    if (oob) atomicAdd((unsigned long long*)&P.cnt[CNT_OOB], (unsigned long long)oob);
    if (SPIN && spin_bad) atomicAdd((unsigned long long*)&P.cnt[CNT_SPIN], (unsigned long long)spin_bad);
}

template <int M, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_push64_deposit(
    P64 P, Fields64 F, Grid g, PushConsts K, Tiling tl, const int64_t* __restrict__ tile_off,
    const int64_t* __restrict__ work, const int64_t* __restrict__ nwork, int chunk,
    const double* __restrict__ dxo, const double* __restrict__ dyo, const double* __restrict__ dzo) {
    using B = AccBox<M, TX, TY, TZ>;
    static_assert(PIC_PUSH_CHUNK <= 65536, "chunk slots are listed as 16-bit offsets");
    extern __shared__ double lds64[];
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(lds64);  // AVOL: J, rho
    uint16_t* lst = reinterpret_cast<uint16_t*>(lds64 + B::AVOL);           // chunk slots by class
    __shared__ int ncls[2];
    const int64_t wi = blockIdx.x;
    if (wi >= *nwork) return;
    const int64_t item = work[wi];
    const int tile = (int)(item >> 20);
    const int64_t cidx = item & ((1 << 20) - 1);
    const int64_t beg = tile_off[tile] + cidx * chunk;
    const int64_t end = min(tile_off[tile + 1], beg + chunk);
    const int tzi = tile / (tl.nt[0] * tl.nt[1]);
    const int tyi = (tile / tl.nt[0]) % tl.nt[1];
    const int txi = tile % tl.nt[0];
    const int T0x = txi * TX, T0y = tyi * TY, T0z = tzi * TZ;
    for (int e = threadIdx.x; e < B::AVOL; e += blockDim.x) acc[e] = 0ull;
    __shared__ unsigned rg_hist[256 / 64][8];  // per wave: bank-spreading counters
    const int lane = lane_id();
    if (lane < 8) rg_hist[threadIdx.x >> 6][lane] = 0u;
    if (threadIdx.x < 2) ncls[threadIdx.x] = 0;
    __syncthreads();
    const int len = (int)(end - beg);
    // the two lists: still particles from the front, the others from the back
    const uint64_t below = (1ull << lane) - 1ull;
    for (int t = threadIdx.x; t - lane < len; t += blockDim.x) {
        const double d = t < len ? dxo[beg + t] : __longlong_as_double(0x7ff8000000000000ll);
        const bool some = d == d, mv = some && signbit(d);
        const uint64_t m0 = __ballot(some && !mv), m1 = __ballot(mv);
        int o0 = 0, o1 = 0;
        if (lane == 0) {
            o0 = m0 ? atomicAdd(&ncls[0], __popcll(m0)) : 0;
            o1 = m1 ? atomicAdd(&ncls[1], __popcll(m1)) : 0;
        }
        o0 = __shfl(o0, 0);
        o1 = __shfl(o1, 0);
        if (some && !mv) lst[o0 + __popcll(m0 & below)] = (uint16_t)t;
        if (mv) lst[len - 1 - (o1 + __popcll(m1 & below))] = (uint16_t)t;
    }
    __syncthreads();
    const int ox = T0x + g.m - 2 - B::H, oy = T0y + g.m - 2 - B::H, oz = T0z + g.m - 2 - B::H;
    LdsSink<B> sink{acc, K.inv_unit, oz, oy, ox, 0};
    const int wv = threadIdx.x >> 6;
    const double org = (double)(1 - g.m);
    auto run = [&](auto still) {
        constexpr bool STILL = decltype(still)::value;
        const int n = ncls[STILL ? 0 : 1];
        for (int pass = 0; 4 * pass < (n + 63) >> 6; pass++) {
            const int e = spread_row(threadIdx.x, pass, n), pk = e;
            if ((threadIdx.x >> 6) + 4 * pass >= (n + 63) >> 6) break;  // this wave's rows are done
            double x0 = 0.0, y0 = 0.0, z0 = 0.0, xn = 0.0, yn = 0.0, zn = 0.0, qw = 0.0;
            if (pk >= 0) {
                const int64_t p = beg + lst[STILL ? e : len - 1 - e];
                x0 = fabs(dxo[p]); y0 = dyo[p]; z0 = dzo[p];
                xn = P.x[p]; yn = P.y[p]; zn = P.z[p];
                qw = K.q * P.w[p];
            }
            // deal by the bank of the deposit box base, the nearest node of
            // the old position (q w == 0: nothing to add)
            int r = 32;
            if (qw != 0.0) {
                const int ib = (int)(x0 - org + 0.5) - B::H, jb = (int)(y0 - org + 0.5) - B::H,
                          kb = (int)(z0 - org + 0.5) - B::H;
                r = (2 * ((kb - oz) * B::PXY + (jb - oy) * B::PX + (ib - ox))) & 31;
            }
            const int to = bank_spread_lane(r, rg_hist[wv]);
            x0 = lane_send_d(to, x0); y0 = lane_send_d(to, y0); z0 = lane_send_d(to, z0);
            xn = lane_send_d(to, xn); yn = lane_send_d(to, yn); zn = lane_send_d(to, zn);
            qw = lane_send_d(to, qw);
            // an in-tile particle that moved < 1 cell: its whole box lies in the LDS box
            if (qw != 0.0) esirkepov<M, LdsSink<B>, STILL>(qw, x0, y0, z0, xn, yn, zn, K, g.m, sink);
        }
    };
    run(std::true_type{});
    run(std::false_type{});
    __syncthreads();
    // flush: one global atomic per nonzero accumulator point
    double* jp[4] = {F.jx, F.jy, F.jz, F.rho};
    for (int e = threadIdx.x; e < B::AVOL; e += blockDim.x) {
        const long long q = (long long)acc[e];
        if (q == 0) continue;
        const int c = e >= B::aoff(3) ? 3 : e >= B::aoff(2) ? 2 : e >= B::aoff(1) ? 1 : 0;
        const int r = e - (c == 3 ? B::aoff(3) : c == 2 ? B::aoff(2) : c == 1 ? B::aoff(1) : 0);
        const double v = (double)q * K.unit[c];
        const int lx = r % B::PX, ly = (r / B::PX) % B::BY, lz = r / B::PXY;
        const int gx = ox + lx, gy = oy + ly, gz = oz + lz;
        if (gx > g.je[0] - 2 || gy > g.je[1] - 2 || gz > g.je[2] - 2) continue;
        hbm_add(&jp[c][g.ji(gz, gy, gx)], v);
    }
}

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------
PushConsts push_consts64(const pic_ctx* c, const Species& S);
static PushConsts make_consts(const pic_ctx* c, const Species& S) { return push_consts64(c, S); }
PushConsts push_consts64(const pic_ctx* c, const Species& S) {
    PushConsts K;
    K.q = S.q;
    K.qm = S.q / (S.m) * (c->p.dt / 2.0);
    K.dt = c->p.dt; K.dx = c->p.dx; K.dy = c->p.dy; K.dz = c->p.dz;
    K.cx = K.cy = K.cz = 0.0;
    for (int a = 0; a < 3; a++) { K.N[a] = c->g.n[a]; K.per[a] = c->p.periodic[a]; }
    K.spin = c->spin;
    const double qw = std::fabs(S.q) * (S.wmax > 0 ? S.wmax : 1.0);
    const double bound[4] = {qw * c->p.dx / c->p.dt, qw * c->p.dy / c->p.dt, qw * c->p.dz / c->p.dt, qw};
    for (int k = 0; k < 4; k++) {
        K.unit[k] = std::ldexp(bound[k], -48);
        K.inv_unit[k] = 1.0 / K.unit[k];
    }
    return K;
}

// Particle pointers of the species' current buffer (read after any sort,
// which swaps the buffers).
static P64 p64(pic_ctx* c, Species& S) {
    const int cur = S.cur;
    return P64{S.x[cur], S.y[cur], S.z[cur], S.ux[cur], S.uy[cur], S.uz[cur], S.w[cur],
               c->keep_old ? S.xo[cur] : nullptr, c->keep_old ? S.yo[cur] : nullptr,
               c->keep_old ? S.zo[cur] : nullptr, S.d_cnt,
               (double*)S.spin_v[0][cur], (double*)S.spin_v[1][cur], (double*)S.spin_v[2][cur]};
}

template <int M, bool SPIN>
static int launch_tiled64(pic_ctx* c, Species& S, const P64& P, const Fields64& Fd,
                          const PushConsts& K, int64_t max_items) {
    using TS = TileShape<0, M>;
    using B = AccBox<M, TS::X, TS::Y, TS::Z>;
    PIC_HIP(c, hipFuncSetAttribute((const void*)k_push64_tiled<M, TS::X, TS::Y, TS::Z, SPIN>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)B::LDS));
    k_push64_tiled<M, TS::X, TS::Y, TS::Z, SPIN><<<(unsigned)max_items, 256, B::LDS, c->stream>>>(
        P, Fd, c->g, K, c->tl, S.tile_off, c->work, c->nwork, PIC_PUSH_CHUNK);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

template <int M, bool SPIN>
static int launch_split64(pic_ctx* c, Species& S, const P64& P, const Fields64& Fd, const PushConsts& K,
                          int64_t max_items) {
    using TS = TileShape<0, M>;
    using B = AccBox<M, TS::X, TS::Y, TS::Z>;
    constexpr size_t lds_g = (size_t)B::FVOL * sizeof(double),
                     lds_d = (size_t)B::AVOL * sizeof(double) + (size_t)PIC_PUSH_CHUNK * sizeof(uint16_t);
    PIC_HIP(c, hipFuncSetAttribute((const void*)k_push64_gather<M, TS::X, TS::Y, TS::Z, SPIN>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_g));
    PIC_HIP(c, hipFuncSetAttribute((const void*)k_push64_deposit<M, TS::X, TS::Y, TS::Z>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_d));
    k_push64_gather<M, TS::X, TS::Y, TS::Z, SPIN><<<(unsigned)max_items, 256, lds_g, c->stream>>>(
        P, Fd, c->g, K, c->tl, S.tile_off, c->work, c->nwork, PIC_PUSH_CHUNK, S.dxo, S.dyo, S.dzo);
    PIC_LAUNCH_CHECK(c);
    k_push64_deposit<M, TS::X, TS::Y, TS::Z><<<(unsigned)max_items, 256, lds_d, c->stream>>>(
        P, Fd, c->g, K, c->tl, S.tile_off, c->work, c->nwork, PIC_PUSH_CHUNK, S.dxo, S.dyo, S.dzo);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

int launch_deposit(pic_ctx* c, int s, int64_t n, const double* const* d6, const double* w,
                   int64_t* d_skipped) {
    Species& S = c->sp[s];
    Fields64 Fd{(const double*)c->F[0], (const double*)c->F[1], (const double*)c->F[2],
                (const double*)c->F[3], (const double*)c->F[4], (const double*)c->F[5],
                (double*)c->J[0], (double*)c->J[1], (double*)c->J[2], (double*)c->rho[s]};
    const PushConsts K = make_consts(c, S);
    const int blocks = (int)std::min<int64_t>(blocks_for(std::max<int64_t>(n, 1), 256), 4096);
    if (c->g.m == 2)
        k_deposit64<2><<<blocks, 256, 0, c->stream>>>(d6[0], d6[1], d6[2], d6[3], d6[4], d6[5], w, n,
                                                       Fd, c->g, K, d_skipped);
    else
        k_deposit64<4><<<blocks, 256, 0, c->stream>>>(d6[0], d6[1], d6[2], d6[3], d6[4], d6[5], w, n,
                                                       Fd, c->g, K, d_skipped);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

int launch_push64(pic_ctx* c, int s) {
    Species& S = c->sp[s];
    Fields64 Fd{(const double*)c->F[0], (const double*)c->F[1], (const double*)c->F[2],
                (const double*)c->F[3], (const double*)c->F[4], (const double*)c->F[5],
                (double*)c->J[0], (double*)c->J[1], (double*)c->J[2], (double*)c->rho[s]};
    PushConsts K = make_consts(c, S);
    const int kernel = c->push_kernel == 0 ? 2 : c->push_kernel;
    if (kernel == 1) {
        PhaseTimer t(c, PIC_T_PUSH);
        int blocks = (int)std::min<int64_t>(blocks_for(std::max<int64_t>(S.cap, 1), 256), 256 * 64);
        const P64 P = p64(c, S);
        if (S.spin) {
            if (c->g.m == 2) k_push64_direct<2, true><<<blocks, 256, 0, c->stream>>>(P, Fd, c->g, K);
            else k_push64_direct<4, true><<<blocks, 256, 0, c->stream>>>(P, Fd, c->g, K);
        } else {
            if (c->g.m == 2) k_push64_direct<2, false><<<blocks, 256, 0, c->stream>>>(P, Fd, c->g, K);
            else k_push64_direct<4, false><<<blocks, 256, 0, c->stream>>>(P, Fd, c->g, K);
        }
        PIC_LAUNCH_CHECK(c);
        return PIC_OK;
    }
    // tiled: particles grouped by tile (a stale grouping only costs speed:
    // a particle found outside its tile is pushed through HBM)
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
    const P64 P = p64(c, S);  // after the sort
    if (S.cellsorted) {  // cell-ordered tiles: the two-kernel push
        if (!S.dxo) {
            const size_t bytes = (size_t)std::max<int64_t>(S.cap, 1) * sizeof(double);
            if (hipMalloc((void**)&S.dxo, bytes) != hipSuccess || hipMalloc((void**)&S.dyo, bytes) != hipSuccess ||
                hipMalloc((void**)&S.dzo, bytes) != hipSuccess) {
                for (double** q : {&S.dxo, &S.dyo, &S.dzo}) {
                    if (*q) (void)hipFree(*q);
                    *q = nullptr;
                }
                return set_err(c, PIC_ERR_NOMEM, "deposit records for %lld particles", (long long)S.cap);
            }
        }
        if (S.spin)
            return c->g.m == 2 ? launch_split64<2, true>(c, S, P, Fd, K, max_items)
                               : launch_split64<4, true>(c, S, P, Fd, K, max_items);
        return c->g.m == 2 ? launch_split64<2, false>(c, S, P, Fd, K, max_items)
                           : launch_split64<4, false>(c, S, P, Fd, K, max_items);
    }
    if (S.spin)
        return c->g.m == 2 ? launch_tiled64<2, true>(c, S, P, Fd, K, max_items)
                           : launch_tiled64<4, true>(c, S, P, Fd, K, max_items);
    return c->g.m == 2 ? launch_tiled64<2, false>(c, S, P, Fd, K, max_items)
                       : launch_tiled64<4, false>(c, S, P, Fd, K, max_items);
}

}  // namespace pic
