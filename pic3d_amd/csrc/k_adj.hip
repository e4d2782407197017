// k_adj.hip -- differentiation form of the FP64 hot loop.
//
// Reverse-mode (adjoint) kernels for every operation of one time step --
// Yee half steps, halo wraps, current fold, laser/Silver-Mueller boundary,
// and the particle push (gather, Boris, position update, Esirkepov
// deposit) -- and the checkpoint/recompute driver behind pic_diff_*.  The
// objective is the final kinetic energy of one species (the FP64
// diagnostic sum m w (gamma - 1) dV); the parameters are the x-min laser
// envelope value of every step (what generate_laser_pulse(t, 0) returns,
// boundary.cpp:8-27) and the initial particle weights (the density
// profile the particles sample).
//
// The field operators are linear; their adjoints are the transposed
// stencils, written as gathers over the adjoint arrays (no atomics).  The
// particle adjoint recomputes the step of each particle from its saved
// state and differentiates it: dual numbers carry d/d(x0,y0,z0) through the
// gather, d/d(u0,E,B) through Boris and d/d(x0..z1) through the deposit;
// the field adjoint of the gather is scattered with HBM atomics.
//
// Particles keep their slots during a differentiated run: the species is
// sorted by tile once when the run begins and never again (the tiled push
// sends particles that left their tile through its HBM path; the direct
// push ignores order), and the boundary marks leavers dead without
// compacting.  So slot p at step n and slot p at step n+1 are the same
// particle.
#include <algorithm>
#include <cmath>
#include <map>

#include "atomics.h"
#include "laser.h"
#include "pic_internal.h"
#include "push64.h"
#include "shapes.h"
#include "wave.h"

namespace pic {

// ---------------------------------------------------------------------
// forward-mode dual numbers (value + N partial derivatives, scalar S)
// ---------------------------------------------------------------------
template <int N, typename S = double> struct Dual {
    S v;
    S d[N];
    __device__ Dual() {}
    __device__ Dual(double x) : v((S)x) {
#pragma unroll
        for (int i = 0; i < N; i++) d[i] = S(0);
    }
    __device__ Dual(float x) : v((S)x) {
#pragma unroll
        for (int i = 0; i < N; i++) d[i] = S(0);
    }
};
template <int N, typename S = double> __device__ __forceinline__ Dual<N, S> dvar(S x, int k) {
    Dual<N, S> r(x);
    r.d[k] = S(1);
    return r;
}
#define DUAL_T template <int N, typename S>
DUAL_T __device__ __forceinline__ Dual<N, S> operator+(const Dual<N, S>& a, const Dual<N, S>& b) {
    Dual<N, S> r(a.v + b.v);
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i];
    return r;
}
DUAL_T __device__ __forceinline__ Dual<N, S> operator-(const Dual<N, S>& a, const Dual<N, S>& b) {
    Dual<N, S> r(a.v - b.v);
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i];
    return r;
}
DUAL_T __device__ __forceinline__ Dual<N, S> operator-(const Dual<N, S>& a) {
    Dual<N, S> r(-a.v);
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
    return r;
}
DUAL_T __device__ __forceinline__ Dual<N, S> operator*(const Dual<N, S>& a, const Dual<N, S>& b) {
    Dual<N, S> r(a.v * b.v);
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
DUAL_T __device__ __forceinline__ Dual<N, S> operator/(const Dual<N, S>& a, const Dual<N, S>& b) {
    const S inv = S(1) / b.v;
    Dual<N, S> r(a.v * inv);
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
    return r;
}
DUAL_T __device__ __forceinline__ Dual<N, S> operator+(const Dual<N, S>& a, double b) { return a + Dual<N, S>(b); }
DUAL_T __device__ __forceinline__ Dual<N, S> operator+(double a, const Dual<N, S>& b) { return Dual<N, S>(a) + b; }
DUAL_T __device__ __forceinline__ Dual<N, S> operator-(const Dual<N, S>& a, double b) { return a - Dual<N, S>(b); }
DUAL_T __device__ __forceinline__ Dual<N, S> operator-(double a, const Dual<N, S>& b) { return Dual<N, S>(a) - b; }
DUAL_T __device__ __forceinline__ Dual<N, S> operator*(const Dual<N, S>& a, double b) {
    const S bs = (S)b;
    Dual<N, S> r(a.v * bs);
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * bs;
    return r;
}
DUAL_T __device__ __forceinline__ Dual<N, S> operator*(double a, const Dual<N, S>& b) { return b * a; }
DUAL_T __device__ __forceinline__ Dual<N, S> operator/(const Dual<N, S>& a, double b) { return a * (1.0 / b); }
DUAL_T __device__ __forceinline__ Dual<N, S> operator/(double a, const Dual<N, S>& b) { return Dual<N, S>(a) / b; }
DUAL_T __device__ __forceinline__ Dual<N, S>& operator+=(Dual<N, S>& a, const Dual<N, S>& b) { return a = a + b; }
DUAL_T __device__ __forceinline__ Dual<N, S> sqrt(const Dual<N, S>& a) {
    const S s = ::sqrt(a.v);
    Dual<N, S> r(s);
    const S k = S(0.5) / s;
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * k;
    return r;
}
#undef DUAL_T
__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ float val(float x) { return x; }
template <int N, typename S> __device__ __forceinline__ S val(const Dual<N, S>& x) { return x.v; }

template <int M, typename T> __device__ __forceinline__ int shape_at_t(T x, T* W) {
    const int ic = (int)(val(x) + 0.5);
    if (M == 2) shape2_delta<T>(x - T((double)ic), W);
    else shape4_d<T>(x - T((double)(ic - 2)), W);
    return ic;
}

// Boris rotation of particle.cpp:271-311 (E and B as gathered, before the
// q dt / 2m scaling)
template <typename T>
__device__ __forceinline__ void boris_t(const T* u, const T* E, const T* B, double qm, T* out) {
    using ::sqrt;
    const T e0 = E[0] * qm, e1 = E[1] * qm, e2 = E[2] * qm;
    const T b0 = B[0] * qm, b1 = B[1] * qm, b2 = B[2] * qm;
    const T px1 = u[0] + e0, py1 = u[1] + e1, pz1 = u[2] + e2;
    const T g1 = sqrt(1.0 + (px1 * px1 + py1 * py1 + pz1 * pz1));
    const T tx = b0 / g1, ty = b1 / g1, tz = b2 / g1;
    const T tt = 2.0 / (1.0 + tx * tx + ty * ty + tz * tz);
    const T sx = tx * tt, sy = ty * tt, sz = tz * tt;
    const T px2 = px1 + (py1 * tz - pz1 * ty);
    const T py2 = py1 + (pz1 * tx - px1 * tz);
    const T pz2 = pz1 + (px1 * ty - py1 * tx);
    out[0] = px1 + (py2 * sz - pz2 * sy) + e0;
    out[1] = py1 + (pz2 * sx - px2 * sz) + e1;
    out[2] = pz1 + (px2 * sy - py2 * sx) + e2;
}

// Transpose of boris_t's Jacobian applied to lo (reverse mode): given the
// adjoint lo of the output momentum, the adjoints of u (lu), E and B (lF,
// E then B).  The forward intermediates are recomputed here, so the
// caller keeps nothing of the forward live across its other work.
template <typename T>
__device__ __forceinline__ void boris_vjp(const T* u, const T* E, const T* B, T qm, const T* lo, T* lu, T* lF) {
    using ::sqrt;
    const T e[3] = {E[0] * qm, E[1] * qm, E[2] * qm};
    const T b[3] = {B[0] * qm, B[1] * qm, B[2] * qm};
    const T p1[3] = {u[0] + e[0], u[1] + e[1], u[2] + e[2]};
    const T g1 = sqrt(T(1) + (p1[0] * p1[0] + p1[1] * p1[1] + p1[2] * p1[2]));
    const T t[3] = {b[0] / g1, b[1] / g1, b[2] / g1};
    const T tt = T(2) / (T(1) + t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    const T sv[3] = {t[0] * tt, t[1] * tt, t[2] * tt};
    const T p2[3] = {p1[0] + (p1[1] * t[2] - p1[2] * t[1]), p1[1] + (p1[2] * t[0] - p1[0] * t[2]),
                     p1[2] + (p1[0] * t[1] - p1[1] * t[0])};
    auto cross = [](const T* a, const T* c, T* r) {
        r[0] = a[1] * c[2] - a[2] * c[1];
        r[1] = a[2] * c[0] - a[0] * c[2];
        r[2] = a[0] * c[1] - a[1] * c[0];
    };
    // out = p1 + p2 x s + e
    T le[3] = {lo[0], lo[1], lo[2]}, lp1[3] = {lo[0], lo[1], lo[2]}, lp2[3], ls[3];
    cross(sv, lo, lp2);  // d/dp2 of (p2 x s).lo = s x lo
    cross(lo, p2, ls);   // d/ds = lo x p2
    // p2 = p1 + p1 x t
    T tl[3], lt[3];
    cross(t, lp2, tl);
    cross(lp2, p1, lt);
#pragma unroll
    for (int a = 0; a < 3; a++) lp1[a] += lp2[a] + tl[a];
    // s = t tt, tt = 2 / (1 + t.t): d tt / dt = -tt^2 t
    const T ltt = t[0] * ls[0] + t[1] * ls[1] + t[2] * ls[2];
#pragma unroll
    for (int a = 0; a < 3; a++) lt[a] += tt * ls[a] - tt * tt * ltt * t[a];
    // t = b / g1, g1 = sqrt(1 + p1.p1)
    const T lg1 = -(t[0] * lt[0] + t[1] * lt[1] + t[2] * lt[2]) / g1;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lp1[a] += lg1 * p1[a] / g1;
        le[a] += lp1[a];
        lu[a] = lp1[a];
        lF[a] = qm * le[a];
        lF[3 + a] = qm * (lt[a] / g1);
    }
}

// Esirkepov deposit of particle.cpp:372-500 (J only: rho feeds nothing)
// as a function of the six positions; 'visit(c, k, j, i, value)' gets the
// contribution to component c at buffer node (k, j, i).  The points the
// push kernels leave out (the last of the particle's box along the
// cumulative axis, whose value is a rounding residue) are left out here.
template <int M, typename T, typename V>
__device__ __forceinline__ void esirkepov_t(double q, T xo, T yo, T zo, T xn, T yn, T zn,
                                            const PushConsts& K, int m, V&& visit) {
    constexpr int H = M / 2, L = M + 2;
    const double org = (double)(1 - m);
    T Sx0[M + 1], Sx1[M + 1], Sy0[M + 1], Sy1[M + 1], Sz0[M + 1], Sz1[M + 1];
    const int i1 = shape_at_t<M>(xo - org, Sx0), i2 = shape_at_t<M>(xn - org, Sx1);
    const int j1 = shape_at_t<M>(yo - org, Sy0), j2 = shape_at_t<M>(yn - org, Sy1);
    const int k1 = shape_at_t<M>(zo - org, Sz0), k2 = shape_at_t<M>(zn - org, Sz1);
    const int bx = min(i1, i2) - H, by = min(j1, j2) - H, bz = min(k1, k2) - H;
    T X0[L], X1[L], Y0[L], Y1[L], Z0[L], Z1[L];
#pragma unroll
    for (int t = 0; t < L; t++) {
        const int a0 = t - (i1 - H - bx), a1 = t - (i2 - H - bx);
        X0[t] = (a0 >= 0 && a0 <= M) ? Sx0[a0] : T(0.0);
        X1[t] = (a1 >= 0 && a1 <= M) ? Sx1[a1] : T(0.0);
        const int b0 = t - (j1 - H - by), b1 = t - (j2 - H - by);
        Y0[t] = (b0 >= 0 && b0 <= M) ? Sy0[b0] : T(0.0);
        Y1[t] = (b1 >= 0 && b1 <= M) ? Sy1[b1] : T(0.0);
        const int c0 = t - (k1 - H - bz), c1 = t - (k2 - H - bz);
        Z0[t] = (c0 >= 0 && c0 <= M) ? Sz0[c0] : T(0.0);
        Z1[t] = (c1 >= 0 && c1 <= M) ? Sz1[c1] : T(0.0);
    }
    const double cx = q * K.dx / K.dt, cy = q * K.dy / K.dt, cz = q * K.dz / K.dt;
    const double third = 1.0 / 3.0;
    for (int k = 0; k < L; k++)
        for (int j = 0; j < L; j++) {
            const T sy0 = Y0[j], dsy = Y1[j] - sy0, sz0 = Z0[k], dsz = Z1[k] - sz0;
            const T wyz = sy0 * sz0 + 0.5 * dsy * sz0 + 0.5 * sy0 * dsz + third * dsy * dsz;
            T acc(0.0);
            for (int i = 0; i < L - 1; i++) {
                acc += X1[i] - X0[i];
                if (i == L - 2 && i1 == i2) break;
                visit(0, bz + k, by + j, bx + i, -(cx * acc * wyz));
            }
        }
    for (int k = 0; k < L; k++)
        for (int i = 0; i < L; i++) {
            const T sx0 = X0[i], dsx = X1[i] - sx0, sz0 = Z0[k], dsz = Z1[k] - sz0;
            const T wxz = sx0 * sz0 + 0.5 * dsx * sz0 + 0.5 * sx0 * dsz + third * dsx * dsz;
            T acc(0.0);
            for (int j = 0; j < L - 1; j++) {
                acc += Y1[j] - Y0[j];
                if (j == L - 2 && j1 == j2) break;
                visit(1, bz + k, by + j, bx + i, -(cy * acc * wxz));
            }
        }
    for (int j = 0; j < L; j++)
        for (int i = 0; i < L; i++) {
            const T sx0 = X0[i], dsx = X1[i] - sx0, sy0 = Y0[j], dsy = Y1[j] - sy0;
            const T wxy = sx0 * sy0 + 0.5 * dsx * sy0 + 0.5 * sx0 * dsy + third * dsx * dsy;
            T acc(0.0);
            for (int k = 0; k < L - 1; k++) {
                acc += Z1[k] - Z0[k];
                if (k == L - 2 && k1 == k2) break;
                visit(2, bz + k, by + j, bx + i, -(cz * acc * wxy));
            }
        }
}

// ---------------------------------------------------------------------
// particle adjoint
// ---------------------------------------------------------------------
struct AdjPush {
    // state at the start of the push (slot order)
    const double *x, *y, *z, *ux, *uy, *uz, *w;
    // fields at the push
    const double* F[6];
    // adjoint of J after the fold transpose (read) and of E/B (scatter)
    const double* lJ[3];
    double* lF[6];
    // per-particle adjoint: in = after the push, out = before it
    double *lx, *ly, *lz, *lux, *luy, *luz;
    double* lw;  // accumulated d objective / d w
    int64_t n;
};

template <int M>
__global__ __launch_bounds__(256) void k_adj_push64(AdjPush A, Grid g, PushConsts K) {
    constexpr int H = M / 2;
    using D3 = Dual<3>;
    using D9 = Dual<9>;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < A.n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const double w = A.w[p];
        if (!(w == w)) continue;  // dead slot
        double x = A.x[p], y = A.y[p], z = A.z[p];
        if (!wrap_check(x, y, z, K)) continue;  // not pushed: adjoint passes through
        // ---- gather at (x, y, z), with d/d(x, y, z)
        const double org = (double)(1 - g.m);
        const D3 xc = dvar<3>(x - org, 0), yc = dvar<3>(y - org, 1), zc = dvar<3>(z - org, 2);
        D3 Wxp[M + 1], Wyp[M + 1], Wzp[M + 1], Wxd[M + 1], Wyd[M + 1], Wzd[M + 1];
        const int ip = shape_at_t<M>(xc, Wxp), jp = shape_at_t<M>(yc, Wyp), kp = shape_at_t<M>(zc, Wzp);
        shape_at_t<M>(xc - 0.5, Wxd);
        shape_at_t<M>(yc - 0.5, Wyd);
        shape_at_t<M>(zc - 0.5, Wzd);
        const int id = (int)xc.v, jd = (int)yc.v, kd = (int)zc.v;
        const D3* WX[6] = {Wxd, Wxp, Wxp, Wxp, Wxd, Wxd};
        const D3* WY[6] = {Wyp, Wyd, Wyp, Wyd, Wyp, Wyd};
        const D3* WZ[6] = {Wzp, Wzp, Wzd, Wzd, Wzd, Wzp};
        const int I0[6] = {id, ip, ip, ip, id, id}, J0[6] = {jp, jd, jp, jd, jp, jd},
                  K0[6] = {kp, kp, kd, kd, kd, kp};
        D3 Fp[6];
        for (int c = 0; c < 6; c++) {
            D3 e(0.0);
            for (int k = 0; k <= M; k++)
                for (int j = 0; j <= M; j++)
                    for (int i = 0; i <= M; i++)
                        e += WX[c][i] * WY[c][j] * WZ[c][k] *
                             A.F[c][g.fi(K0[c] + k - H, J0[c] + j - H, I0[c] + i - H)];
            Fp[c] = e;
        }
        // ---- Boris with d/d(u0, E, B)
        D9 u0[3] = {dvar<9>(A.ux[p], 0), dvar<9>(A.uy[p], 1), dvar<9>(A.uz[p], 2)};
        D9 E9[3] = {dvar<9>(Fp[0].v, 3), dvar<9>(Fp[1].v, 4), dvar<9>(Fp[2].v, 5)};
        D9 B9[3] = {dvar<9>(Fp[3].v, 6), dvar<9>(Fp[4].v, 7), dvar<9>(Fp[5].v, 8)};
        D9 u1[3];
        boris_t<D9>(u0, E9, B9, K.qm, u1);
        const double v1[3] = {u1[0].v, u1[1].v, u1[2].v};
        const double gam = ::sqrt(1.0 + (v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]));
        const double dtd[3] = {K.dt / K.dx, K.dt / K.dy, K.dt / K.dz};
        const double xn = x + v1[0] / gam * K.dt / K.dx;
        const double yn = y + v1[1] / gam * K.dt / K.dy;
        const double zn = z + v1[2] / gam * K.dt / K.dz;
        // ---- adjoint of the state after the push
        double lx1[3] = {A.lx[p], A.ly[p], A.lz[p]};
        double lu1[3] = {A.lux[p], A.luy[p], A.luz[p]};
        double lx0[3] = {0.0, 0.0, 0.0};
        double gw = 0.0;
        // ---- deposit (skipped by the forward for moves of a cell or more)
        if (fabs(xn - x) < 1.0 && fabs(yn - y) < 1.0 && fabs(zn - z) < 1.0) {
            using D6 = Dual<6>;
            D6 G(0.0);
            const double* lJ0 = A.lJ[0];
            const double* lJ1 = A.lJ[1];
            const double* lJ2 = A.lJ[2];
            esirkepov_t<M, D6>(K.q, dvar<6>(x, 0), dvar<6>(y, 1), dvar<6>(z, 2), dvar<6>(xn, 3),
                               dvar<6>(yn, 4), dvar<6>(zn, 5), K, g.m,
                               [&](int c, int k, int j, int i, const D6& v) {
                                   const double* l = c == 0 ? lJ0 : (c == 1 ? lJ1 : lJ2);
                                   G += v * l[g.ji(k, j, i)];
                               });
            for (int a = 0; a < 3; a++) {
                lx0[a] += w * G.d[a];
                lx1[a] += w * G.d[3 + a];
            }
            gw = G.v;
        }
        // ---- x1 = x0 + u1 / gamma(u1) * dt / d
        const double ig = 1.0 / gam, ig3 = ig * ig * ig;
        for (int b = 0; b < 3; b++) {
            double s = 0.0;
            for (int a = 0; a < 3; a++)
                s += lx1[a] * dtd[a] * ((a == b ? ig : 0.0) - v1[a] * v1[b] * ig3);
            lu1[b] += s;
        }
        for (int a = 0; a < 3; a++) lx0[a] += lx1[a];
        // ---- Boris transpose
        double lu0[3], lF[6];
        for (int b = 0; b < 9; b++) {
            const double s = lu1[0] * u1[0].d[b] + lu1[1] * u1[1].d[b] + lu1[2] * u1[2].d[b];
            if (b < 3) lu0[b] = s;
            else lF[b - 3] = s;
        }
        // ---- gather transpose: d/dx0 and the scatter onto the field grid
        for (int c = 0; c < 6; c++) {
            for (int a = 0; a < 3; a++) lx0[a] += lF[c] * Fp[c].d[a];
            if (lF[c] == 0.0) continue;
            for (int k = 0; k <= M; k++)
                for (int j = 0; j <= M; j++)
                    for (int i = 0; i <= M; i++)
                        hbm_add(&A.lF[c][g.fi(K0[c] + k - H, J0[c] + j - H, I0[c] + i - H)],
                                WX[c][i].v * WY[c][j].v * WZ[c][k].v * lF[c]);
        }
        A.lx[p] = lx0[0]; A.ly[p] = lx0[1]; A.lz[p] = lx0[2];
        A.lux[p] = lu0[0]; A.luy[p] = lu0[1]; A.luz[p] = lu0[2];
        A.lw[p] += gw;
    }
}

// ---------------------------------------------------------------------
// particle adjoint, separable form
//
// The same derivatives as k_adj_push64, evaluated by separable
// contractions instead of dual numbers: B-spline weights come with their
// derivatives (one direction per axis), the gather and its position
// gradient are contracted axis by axis, and the deposit's adjoint is
//   G = sum lambda_J * J(x0, x1) over the Esirkepov box,
// with, for Jx = -cx A_i W_jk (A the cumulative x sum, W the y/z weight):
//   dG/dx0,1 = -cx sum_i dA_i/dx0,1 P_i,        P_i = sum_jk lJ_ijk W_jk
//   dG/dy0,1 = -cx sum_j dY0,1_j sum_k q_jk dW/dY0,1,   q_jk = sum_i lJ_ijk A_i
//   dG/dz0,1 likewise; Jy and Jz by permuting the axes.
// Templated on the scalar T (double: the FP64 form; float: the FP32 form).
// A position along an axis is an integer buffer node 'base' plus a
// fraction 'frac' (FP64: base 0 and the whole buffer coordinate; FP32: the
// particle's cell node and its in-cell offset, so no precision is lost far
// from the origin).
// ---------------------------------------------------------------------

// weights and derivatives of the shape at buffer coordinate base + frac;
// returns the centre node (int)(coordinate + 0.5)
template <int M, typename T>
__device__ __forceinline__ int shape_dv(int base, T frac, T* W, T* dW) {
    const int t = (int)floor(val(frac) + T(0.5));
    Dual<1, T> Wd[M + 1];
    const Dual<1, T> x = dvar<1, T>(frac, 0);
    if (M == 2) shape2_delta<Dual<1, T>>(x - Dual<1, T>((double)t), Wd);
    else shape4_d<Dual<1, T>>(x - Dual<1, T>((double)(t - 2)), Wd);
#pragma unroll
    for (int q = 0; q <= M; q++) {
        W[q] = Wd[q].v;
        dW[q] = Wd[q].d[0];
    }
    return base + t;
}

// register rotations for rolled loops (compile-time indices only):
// rot_down reads element 0 per iteration; rot_push appends
template <int L, typename T> __device__ __forceinline__ void rot_down(T* A) {
#pragma unroll
    for (int t = 0; t < L - 1; t++) A[t] = A[t + 1];
}
template <int L, typename T> __device__ __forceinline__ void rot_push(T* A, T v) {
#pragma unroll
    for (int t = 0; t < L - 1; t++) A[t] = A[t + 1];
    A[L - 1] = v;
}

// weights S placed at offset d in a box of M+2
template <int M, typename T> __device__ __forceinline__ void place_t(const T* S, int d, T* A) {
#pragma unroll
    for (int t = 0; t < M + 2; t++) {
        const T a0 = (t <= M) ? S[t] : T(0);
        const T a1 = (t >= 1) ? S[t - 1] : T(0);
        A[t] = d ? a1 : a0;
    }
}

// Adjoint of one Esirkepov current component along axis a (cumulative
// sum along a; weights over the other two axes b (middle loop) and c
// (outer loop)).  lj(oc, ob, oa) reads lambda of this component at the box
// point.  Adds to g[0..5] (d/dx0,y0,z0,x1,y1,z1) and returns G.  As in the
// forward deposit, the last point of the particle's box along a (box point
// L-1 if it changes node along a, 'moved', else L-2) is not deposited, so
// it takes no lambda.
template <int L, typename T, typename LJ>
__device__ __forceinline__ T esirkepov_adj_axis(int a, const T* A0, const T* A1, const T* dA0,
                                                const T* dA1, const T* B0, const T* B1, const T* dB0,
                                                const T* dB1, const T* C0, const T* C1, const T* dC0,
                                                const T* dC1, T coef, LJ lj, T* g, int ab, int ac,
                                                bool moved) {
    T S[L], dS0[L], dS1[L];
    T s = T(0), s0 = T(0), s1 = T(0);
#pragma unroll
    for (int t = 0; t < L; t++) {
        s += A1[t] - A0[t];
        s0 -= dA0[t];
        s1 += dA1[t];
        S[t] = s; dS0[t] = s0; dS1[t] = s1;
    }
    T P[L], R0[L], R1[L], T0[L], T1[L], Cr0[L], Cr1[L];
#pragma unroll
    for (int t = 0; t < L; t++) {
        P[t] = T(0); R0[t] = T(0); R1[t] = T(0); T0[t] = T(0); T1[t] = T(0);
        Cr0[t] = C0[t]; Cr1[t] = C1[t];
    }
    const T sixth = T(1.0 / 6.0), third = T(1.0 / 3.0), half = T(0.5);
#pragma unroll 1
    for (int oc = 0; oc < L; oc++) {
        const T c0 = Cr0[0], dc = Cr1[0] - c0;
        rot_down<L>(Cr0);
        rot_down<L>(Cr1);
        T t0 = T(0), t1 = T(0);
#pragma unroll
        for (int ob = 0; ob < L; ob++) {
            const T b0 = B0[ob], db = B1[ob] - b0;
            const T W = b0 * c0 + half * db * c0 + half * b0 * dc + third * db * dc;
            T q = T(0);
#pragma unroll
            for (int oa = 0; oa < L - 1; oa++) {
                const T lam = (oa < L - 2 || moved) ? lj(oc, ob, oa) : T(0);
                P[oa] += lam * W;
                q += lam * S[oa];
            }
            R0[ob] += q * (half * c0 + sixth * dc);
            R1[ob] += q * (half * c0 + third * dc);
            t0 += q * (half * b0 + sixth * db);
            t1 += q * (half * b0 + third * db);
        }
        rot_push<L>(T0, t0);
        rot_push<L>(T1, t1);
    }
    T G = T(0), ga0 = T(0), ga1 = T(0), gb0 = T(0), gb1 = T(0), gc0 = T(0), gc1 = T(0);
#pragma unroll
    for (int t = 0; t < L; t++) {
        G += S[t] * P[t];
        ga0 += dS0[t] * P[t];
        ga1 += dS1[t] * P[t];
        gb0 += R0[t] * dB0[t];
        gb1 += R1[t] * dB1[t];
        gc0 += T0[t] * dC0[t];
        gc1 += T1[t] * dC1[t];
    }
    g[a] += -coef * ga0;
    g[3 + a] += -coef * ga1;
    g[ab] += -coef * gb0;
    g[3 + ab] += -coef * gb1;
    g[ac] += -coef * gc0;
    g[3 + ac] += -coef * gc1;
    return -coef * G;
}

// Constants of the differentiated push in the scalar of the form.
template <typename T> struct AdjK {
    T q, qm;        // charge; q dt / (2 m)
    T dtd[3];       // dt / dx, dt / dy, dt / dz
    T jc[3];        // q dx / dt, q dy / dt, q dz / dt (Esirkepov's factor per unit weight)
};

// g[0..5] += d/d(x0..z1) of sum lambda_J * J (per unit weight); returns
// the sum itself (d/dw).  Old position base + f0, new base + f1 (buffer
// coordinates).  lj(c, k, j, i) reads lambda_J of component c at buffer
// node (k, j, i).
template <int M, typename T, typename LJ>
__device__ __forceinline__ T deposit_adj(const int* base, const T* f0, const T* f1,
                                         const AdjK<T>& K, LJ lj, T* g) {
    constexpr int H = M / 2, L = M + 2;
    T S[6][M + 1], dS[6][M + 1];
    const int i1 = shape_dv<M, T>(base[0], f0[0], S[0], dS[0]), i2 = shape_dv<M, T>(base[0], f1[0], S[3], dS[3]);
    const int j1 = shape_dv<M, T>(base[1], f0[1], S[1], dS[1]), j2 = shape_dv<M, T>(base[1], f1[1], S[4], dS[4]);
    const int k1 = shape_dv<M, T>(base[2], f0[2], S[2], dS[2]), k2 = shape_dv<M, T>(base[2], f1[2], S[5], dS[5]);
    const int bx = min(i1, i2) - H, by = min(j1, j2) - H, bz = min(k1, k2) - H;
    T X0[L], X1[L], Y0[L], Y1[L], Z0[L], Z1[L];
    T dX0[L], dX1[L], dY0[L], dY1[L], dZ0[L], dZ1[L];
    place_t<M>(S[0], i1 - H - bx, X0); place_t<M>(dS[0], i1 - H - bx, dX0);
    place_t<M>(S[3], i2 - H - bx, X1); place_t<M>(dS[3], i2 - H - bx, dX1);
    place_t<M>(S[1], j1 - H - by, Y0); place_t<M>(dS[1], j1 - H - by, dY0);
    place_t<M>(S[4], j2 - H - by, Y1); place_t<M>(dS[4], j2 - H - by, dY1);
    place_t<M>(S[2], k1 - H - bz, Z0); place_t<M>(dS[2], k1 - H - bz, dZ0);
    place_t<M>(S[5], k2 - H - bz, Z1); place_t<M>(dS[5], k2 - H - bz, dZ1);
    T G = T(0);
    // Jx: cumulative in x, weights over y (middle) and z (outer)
    G += esirkepov_adj_axis<L, T>(0, X0, X1, dX0, dX1, Y0, Y1, dY0, dY1, Z0, Z1, dZ0, dZ1, K.jc[0],
                                  [&](int oz, int oy, int ox) { return lj(0, bz + oz, by + oy, bx + ox); },
                                  g, 1, 2, i1 != i2);
    // Jy: cumulative in y, weights over x (middle) and z (outer)
    G += esirkepov_adj_axis<L, T>(1, Y0, Y1, dY0, dY1, X0, X1, dX0, dX1, Z0, Z1, dZ0, dZ1, K.jc[1],
                                  [&](int oz, int ox, int oy) { return lj(1, bz + oz, by + oy, bx + ox); },
                                  g, 0, 2, j1 != j2);
    // Jz: cumulative in z, weights over x (middle) and y (outer)
    G += esirkepov_adj_axis<L, T>(2, Z0, Z1, dZ0, dZ1, X0, X1, dX0, dX1, Y0, Y1, dY0, dY1, K.jc[2],
                                  [&](int oy, int ox, int oz) { return lj(2, bz + oz, by + oy, bx + ox); },
                                  g, 0, 1, k1 != k2);
    return G;
}

// Adjoint of one particle's push from its state before the push:
// position base + frac (buffer coordinates), momentum u0, weight w.
// fr(c, k, j, i): field c at a buffer node; lj(c, k, j, i): lambda_J;
// lf(c, k, j, i, v): adds v to lambda_F.  lx, lu: in = adjoint after the
// push, out = before it.  Returns this step's d objective / d w.
// SKIP_FAST: a move of a cell or more leaves the particle untouched (the
// FP32 push) instead of moving it without a deposit (the FP64 push).
template <int M, typename T, bool SKIP_FAST, typename FR, typename LJ, typename LF>
__device__ __forceinline__ T adj_particle(const int* base, const T* frac, const T* u0v, T w,
                                          const AdjK<T>& K, FR fr, LJ lj, LF lf, T* lx, T* lu) {
    constexpr int H = M / 2;
    // gather with position gradient, separably
    T Wp[3][M + 1], dWp[3][M + 1], Wd[3][M + 1], dWd[3][M + 1];
    int ipn[3], idn[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        ipn[a] = shape_dv<M, T>(base[a], frac[a], Wp[a], dWp[a]);
        idn[a] = shape_dv<M, T>(base[a], frac[a] - T(0.5), Wd[a], dWd[a]);
    }
    // staggering per component: 1 = dual along that axis
    constexpr int SX[6] = {1, 0, 0, 0, 1, 1}, SY[6] = {0, 1, 0, 1, 0, 1}, SZ[6] = {0, 0, 1, 1, 1, 0};
    T Fv[6], Fg[6][3];
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const T* WX = SX[c] ? Wd[0] : Wp[0];
        const T* dWX = SX[c] ? dWd[0] : dWp[0];
        const T* WY = SY[c] ? Wd[1] : Wp[1];
        const T* dWY = SY[c] ? dWd[1] : dWp[1];
        const T* WZ = SZ[c] ? Wd[2] : Wp[2];
        const T* dWZ = SZ[c] ? dWd[2] : dWp[2];
        const int i0 = (SX[c] ? idn[0] : ipn[0]) - H, j0 = (SY[c] ? idn[1] : ipn[1]) - H,
                  k0 = (SZ[c] ? idn[2] : ipn[2]) - H;
        T v = T(0), gx = T(0), gy = T(0), gz = T(0);
        T wzr[M + 1], dwzr[M + 1];
#pragma unroll
        for (int t = 0; t <= M; t++) { wzr[t] = WZ[t]; dwzr[t] = dWZ[t]; }
#pragma unroll 1
        for (int k = 0; k <= M; k++) {
            const T wz = wzr[0], dwz = dwzr[0];
            rot_down<M + 1>(wzr);
            rot_down<M + 1>(dwzr);
#pragma unroll
            for (int j = 0; j <= M; j++) {
                T s = T(0), sd = T(0);
#pragma unroll
                for (int i = 0; i <= M; i++) {
                    const T f = fr(c, k0 + k, j0 + j, i0 + i);
                    s += WX[i] * f;
                    sd += dWX[i] * f;
                }
                const T yz = WY[j] * wz;
                v += yz * s;
                gx += yz * sd;
                gy += dWY[j] * wz * s;
                gz += WY[j] * dwz * s;
            }
        }
        Fv[c] = v;
        Fg[c][0] = gx; Fg[c][1] = gy; Fg[c][2] = gz;
    }
    // Boris (its transpose below, by boris_vjp)
    T v1[3];
    boris_t<T>(u0v, Fv, Fv + 3, (double)K.qm, v1);
    const T gam = ::sqrt(T(1) + (v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]));
    T f1[3];
#pragma unroll
    for (int a = 0; a < 3; a++) f1[a] = frac[a] + v1[a] / gam * K.dtd[a];
    T lx0[3] = {T(0), T(0), T(0)};
    T gw = T(0);
    const bool moved = fabs(f1[0] - frac[0]) < T(1) && fabs(f1[1] - frac[1]) < T(1) && fabs(f1[2] - frac[2]) < T(1);
    if (SKIP_FAST && !moved) return T(0);
    if (moved) {
        T g[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
        gw = deposit_adj<M, T>(base, frac, f1, K, lj, g);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lx0[a] += w * g[a];
            lx[a] += w * g[3 + a];
        }
    }
    // x1 = x0 + u1 / gamma(u1) * dt / d
    const T ig = T(1) / gam, ig3 = ig * ig * ig;
    T lu1[3] = {lu[0], lu[1], lu[2]};
#pragma unroll
    for (int b = 0; b < 3; b++) {
        T s = T(0);
#pragma unroll
        for (int a = 0; a < 3; a++) s += lx[a] * K.dtd[a] * ((a == b ? ig : T(0)) - v1[a] * v1[b] * ig3);
        lu1[b] += s;
    }
#pragma unroll
    for (int a = 0; a < 3; a++) lx0[a] += lx[a];
    // Boris transpose
    T lF[6];
    boris_vjp<T>(u0v, Fv, Fv + 3, (T)K.qm, lu1, lu, lF);
    // gather transpose
#pragma unroll
    for (int c = 0; c < 6; c++) {
#pragma unroll
        for (int a = 0; a < 3; a++) lx0[a] += lF[c] * Fg[c][a];
        const T* WX = SX[c] ? Wd[0] : Wp[0];
        const T* WY = SY[c] ? Wd[1] : Wp[1];
        const T* WZ = SZ[c] ? Wd[2] : Wp[2];
        const int i0 = (SX[c] ? idn[0] : ipn[0]) - H, j0 = (SY[c] ? idn[1] : ipn[1]) - H,
                  k0 = (SZ[c] ? idn[2] : ipn[2]) - H;
        T wzr[M + 1];
#pragma unroll
        for (int t = 0; t <= M; t++) wzr[t] = WZ[t];
#pragma unroll 1
        for (int k = 0; k <= M; k++) {
            const T wz = wzr[0] * lF[c];
            rot_down<M + 1>(wzr);
#pragma unroll
            for (int j = 0; j <= M; j++) {
                const T yz = WY[j] * wz;
#pragma unroll
                for (int i = 0; i <= M; i++) lf(c, k0 + k, j0 + j, i0 + i, WX[i] * yz);
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) lx[a] = lx0[a];
    return gw;
}

__host__ __device__ inline AdjK<double> adj_consts64(const PushConsts& P) {
    AdjK<double> K;
    K.q = P.q; K.qm = P.qm;
    K.dtd[0] = P.dt / P.dx; K.dtd[1] = P.dt / P.dy; K.dtd[2] = P.dt / P.dz;
    K.jc[0] = P.q * P.dx / P.dt; K.jc[1] = P.q * P.dy / P.dt; K.jc[2] = P.q * P.dz / P.dt;
    return K;
}

// FP32 state at the start of the push (slot order): linear cell (-1 =
// removed) + in-cell offsets, fp32 momenta and weights; fp32 fields and
// adjoints, the weight adjoint accumulated in FP64.
struct AdjPush32 {
    const int32_t* cell;
    const float *fx, *fy, *fz, *ux, *uy, *uz, *w;
    const float* F[6];
    const float* lJ[3];
    float* lF[6];
    float *lx, *ly, *lz, *lux, *luy, *luz;  // lx..lz: adjoint of the offsets (cell units)
    double* lw;
    int64_t n;
};

// Particle views of the adjoint push, one per precision.  locate(p):
// whether slot p is pushed this step, its cell per axis and its position
// as a buffer node plus a fraction; adjoint<M>(p, ...): that particle's
// adjoint through adj_particle, written back to the slot.
struct AdjView64 {
    using T = double;
    // waves per SIMD the tiled adjoint asks of the register allocator (two
    // would spill heavily)
    static constexpr int MIN_WAVES = 1;
    AdjPush A;
    PushConsts P;
    AdjK<double> K;
    int m;
    __device__ __forceinline__ bool locate(int64_t p, int* cell, int* base, double* frac) const {
        const double w = A.w[p];
        if (!(w == w)) return false;  // dead slot
        double x = A.x[p], y = A.y[p], z = A.z[p];
        if (!wrap_check(x, y, z, P)) return false;  // not pushed: the adjoint passes through
        const double org = (double)(1 - m);
        const double v[3] = {x, y, z};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            cell[a] = min((int)v[a], (int)P.N[a] - 1);
            base[a] = 0;
            frac[a] = v[a] - org;
        }
        return true;
    }
    template <int M, typename FR, typename LJ, typename LF>
    __device__ __forceinline__ void adjoint(int64_t p, const int* base, const double* frac, FR fr, LJ lj,
                                            LF lf) const {
        const double u0[3] = {A.ux[p], A.uy[p], A.uz[p]};
        double lx[3] = {A.lx[p], A.ly[p], A.lz[p]};
        double lu[3] = {A.lux[p], A.luy[p], A.luz[p]};
        const double gw = adj_particle<M, double, false>(base, frac, u0, A.w[p], K, fr, lj, lf, lx, lu);
        A.lx[p] = lx[0]; A.ly[p] = lx[1]; A.lz[p] = lx[2];
        A.lux[p] = lu[0]; A.luy[p] = lu[1]; A.luz[p] = lu[2];
        A.lw[p] += gw;
    }
};

struct AdjView32 {
    using T = float;
    // two waves per SIMD: the quartic 4x4x4 tile's boxes (68.5 KB) let two
    // workgroups share a CU, and the register cap at 256 (444 bytes of
    // spills per lane) still measured faster than one wave per SIMD
    // (backward 5.90 -> 5.14 ms per step at the 4.17M-electron ramp case);
    // the quadratic kernels fit 256 registers without spills
    static constexpr int MIN_WAVES = 2;
    AdjPush32 A;
    AdjK<float> K;
    int n[3];
    int m;
    CellSplit cs;
    __device__ __forceinline__ bool locate(int64_t p, int* cell, int* base, float* frac) const {
        const int c = A.cell[p];
        if (c < 0) return false;  // removed at an open face
        cs.split(c, cell[0], cell[1], cell[2]);
#pragma unroll
        for (int a = 0; a < 3; a++) base[a] = cell[a] + m - 1;
        frac[0] = A.fx[p]; frac[1] = A.fy[p]; frac[2] = A.fz[p];
        return true;
    }
    template <int M, typename FR, typename LJ, typename LF>
    __device__ __forceinline__ void adjoint(int64_t p, const int* base, const float* frac, FR fr, LJ lj,
                                            LF lf) const {
        const float u0[3] = {A.ux[p], A.uy[p], A.uz[p]};
        float lx[3] = {A.lx[p], A.ly[p], A.lz[p]};
        float lu[3] = {A.lux[p], A.luy[p], A.luz[p]};
        const float gw = adj_particle<M, float, true>(base, frac, u0, A.w[p], K, fr, lj, lf, lx, lu);
        A.lx[p] = lx[0]; A.ly[p] = lx[1]; A.lz[p] = lx[2];
        A.lux[p] = lu[0]; A.luy[p] = lu[1]; A.luz[p] = lu[2];
        A.lw[p] += (double)gw;
    }
};

// separable adjoint push through HBM (gathers from the field arrays,
// global atomics into lambda_F): every slot (list = null, any order), or
// the slots of a list (the tiled kernel's strays)
template <int M, typename V>
__global__ __launch_bounds__(256) void k_adj_push_hbm(V v, Grid g, const int64_t* __restrict__ list,
                                                      const int64_t* __restrict__ nlist) {
    using T = typename V::T;
    const int64_t cnt = list ? *nlist : v.A.n;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < cnt;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = list ? list[q] : q;
        int cell[3], base[3];
        T frac[3];
        if (!v.locate(p, cell, base, frac)) continue;
        v.template adjoint<M>(
            p, base, frac, [&](int c, int k, int j, int i) { return v.A.F[c][g.fi(k, j, i)]; },
            [&](int c, int k, int j, int i) { return v.A.lJ[c][g.ji(k, j, i)]; },
            [&](int c, int k, int j, int i, T x) { hbm_add(&v.A.lF[c][g.fi(k, j, i)], x); });
    }
}

// LDS-tiled adjoint push: one workgroup per (tile, chunk) work item of the
// frozen tile grouping.  lambda_J of the tile's deposit box and the
// fields of its gather box are staged in LDS; lambda_F of the gather box
// accumulates in LDS (LACC: double, ds_add_f64; float, ds_add_f32) and is
// flushed with one global atomic per non-zero point.  Particles that left
// their tile take the HBM path (k_adj_push_hbm over the stray list).
template <int M, int TX, int TY, int TZ, typename T, typename LACC>
struct AdjBox {
    static constexpr int H = M / 2;
    static constexpr int AX = TX + M + 3, AY = TY + M + 3, AZ = TZ + M + 3, AVOL = AX * AY * AZ;
    static constexpr int FX = TX + M + 1, FY = TY + M + 1, FZ = TZ + M + 1, FVOL = FX * FY * FZ;
    // accumulators first (8-byte aligned), then lambda_J and the fields
    static constexpr size_t LDS = (size_t)6 * FVOL * sizeof(LACC) + (size_t)(3 * AVOL + 6 * FVOL) * sizeof(T);
};

//
// SPLIT > 1: the tile is cut along x into SPLIT parts, each work item runs
// as SPLIT workgroups, and a workgroup takes the particles of its part only
// (particles outside the whole tile go to the stray list from part 0).
// The LDS boxes shrink with the part, so more workgroups fit a CU.
template <int M, int TX, int TY, int TZ, typename V, typename LACC, int SPLIT = 1>
__global__ __launch_bounds__(256, V::MIN_WAVES) void k_adj_push_tiled(V v, Grid g, Tiling tl,
                                                        const int64_t* __restrict__ tile_off,
                                                        const int64_t* __restrict__ work,
                                                        const int64_t* __restrict__ nwork, int chunk,
                                                        int64_t* __restrict__ stray_list,
                                                        int64_t* __restrict__ stray_n) {
    using T = typename V::T;
    constexpr int PX = TX / SPLIT;  // x cells of this workgroup's part
    static_assert(PX * SPLIT == TX, "the tile splits evenly");
    using B = AdjBox<M, PX, TY, TZ, T, LACC>;
    extern __shared__ double lds_adj[];
    LACC* lfb = reinterpret_cast<LACC*>(lds_adj);    // 6 * FVOL: lambda_F accumulator
    T* lj = reinterpret_cast<T*>(lfb + 6 * B::FVOL);  // 3 * AVOL: lambda_J of the deposit box
    T* fld = lj + 3 * B::AVOL;                        // 6 * FVOL: the fields of the gather box
    const int64_t wi = blockIdx.x / SPLIT;
    const int part = (int)(blockIdx.x % SPLIT);
    if (wi >= *nwork) return;
    const int64_t item = work[wi];
    const int tile = (int)(item >> 20);
    const int64_t cidx = item & ((1 << 20) - 1);
    const int64_t beg = tile_off[tile] + cidx * chunk;
    const int64_t end = min(tile_off[tile + 1], beg + chunk);
    const int txi = tile % tl.nt[0], tyi = (tile / tl.nt[0]) % tl.nt[1], tzi = tile / (tl.nt[0] * tl.nt[1]);
    const int Tx = txi * TX, T0x = Tx + part * PX, T0y = tyi * TY, T0z = tzi * TZ;
    const int m = g.m;
    // deposit box origin (J buffer indices) and gather box origin (E/B)
    const int aox = T0x + m - 2 - B::H, aoy = T0y + m - 2 - B::H, aoz = T0z + m - 2 - B::H;
    const int fox = T0x + m - 1 - B::H, foy = T0y + m - 1 - B::H, foz = T0z + m - 1 - B::H;
    for (int e = threadIdx.x; e < 3 * B::AVOL; e += blockDim.x) {
        const int c = e / B::AVOL, r = e % B::AVOL;
        const int lx = r % B::AX, ly = (r / B::AX) % B::AY, lz = r / (B::AX * B::AY);
        // the box of a partial last tile reaches past the array: clamp (unused)
        const int gx = min(aox + lx, g.je[0] - 2), gy = min(aoy + ly, g.je[1] - 2), gz = min(aoz + lz, g.je[2] - 2);
        lj[e] = v.A.lJ[c][g.ji(gz, gy, gx)];
    }
    __shared__ unsigned rg_hist[256 / 64][8];  // per wave: bank-spreading counters
    if (lane_id() < 8) rg_hist[threadIdx.x >> 6][lane_id()] = 0u;
    for (int e = threadIdx.x; e < 6 * B::FVOL; e += blockDim.x) {
        lfb[e] = LACC(0);
        const int c = e / B::FVOL, r = e % B::FVOL;
        const int lx = r % B::FX, ly = (r / B::FX) % B::FY, lz = r / (B::FX * B::FY);
        const int gx = min(fox + lx, g.fe[0] - 1), gy = min(foy + ly, g.fe[1] - 1), gz = min(foz + lz, g.fe[2] - 1);
        fld[e] = v.A.F[c][g.fi(gz, gy, gx)];
    }
    __syncthreads();
    uint32_t strays = 0;
    int it = 0;
    const int len = (int)(end - beg);
    // Each wave takes 64 consecutive slots, then deals its particles out
    // over its lanes by the LDS bank of their lambda_F base
    // (bank_spread_lane), as the forward push does.
    const int wv = threadIdx.x >> 6, lane = lane_id();
    for (int t0 = threadIdx.x & ~63; t0 < len; t0 += blockDim.x, it++) {
        const int t = t0 + lane;
        int cell[3] = {0, 0, 0}, base[3] = {0, 0, 0};
        T frac[3] = {T(0), T(0), T(0)};
        bool ok = t < len && v.locate(beg + t, cell, base, frac);
        if (ok && ((unsigned)(cell[0] - Tx) >= (unsigned)TX || (unsigned)(cell[1] - T0y) >= (unsigned)TY ||
                   (unsigned)(cell[2] - T0z) >= (unsigned)TZ)) {
            if (part == 0) strays |= 1u << it;
            ok = false;
        }
        if (ok && (unsigned)(cell[0] - T0x) >= (unsigned)PX) ok = false;  // another part's particle
        int r = 32;
        if (ok) {
            int nn[3];
#pragma unroll
            for (int a = 0; a < 3; a++) nn[a] = base[a] + (int)floor(frac[a] + T(0.5));
            const int idx = ((nn[2] - foz) * B::FY + (nn[1] - foy)) * B::FX + (nn[0] - fox);
            r = (idx * (int)(sizeof(LACC) / 4)) & 31;
        }
        const int to = bank_spread_lane(r, rg_hist[wv]);
        const int okv = lane_send(to, ok ? 1 : 0);
        const int tt = lane_send(to, t);
#pragma unroll
        for (int a = 0; a < 3; a++) { base[a] = lane_send(to, base[a]); frac[a] = lane_send(to, frac[a]); }
        if (!okv) continue;
        const int64_t p = beg + tt;
        v.template adjoint<M>(
            p, base, frac,
            [&](int c, int k, int j, int i) {
                return fld[c * B::FVOL + ((k - foz) * B::FY + (j - foy)) * B::FX + (i - fox)];
            },
            [&](int c, int k, int j, int i) {
                return lj[c * B::AVOL + ((k - aoz) * B::AY + (j - aoy)) * B::AX + (i - aox)];
            },
            [&](int c, int k, int j, int i, T x) {
                lds_add(&lfb[c * B::FVOL + ((k - foz) * B::FY + (j - foy)) * B::FX + (i - fox)], (LACC)x);
            });
    }
    // particles that left their tile: appended to the list the HBM kernel
    // processes after this one
    if (strays) {
        it = 0;
        for (int t = threadIdx.x; t < len; t += blockDim.x, it++)
            if (strays & (1u << it)) {
                const unsigned long long at = atomicAdd((unsigned long long*)stray_n, 1ull);
                stray_list[at] = beg + t;
            }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 6 * B::FVOL; e += blockDim.x) {
        const LACC x = lfb[e];
        if (x == LACC(0)) continue;
        const int c = e / B::FVOL, r = e % B::FVOL;
        const int gx = fox + r % B::FX, gy = foy + (r / B::FX) % B::FY, gz = foz + r / (B::FX * B::FY);
        if (gx >= g.fe[0] || gy >= g.fe[1] || gz >= g.fe[2]) continue;
        hbm_add(&v.A.lF[c][g.fi(gz, gy, gx)], (T)x);
    }
}

// adjoint of the final kinetic energy sum_p m w (gamma - 1) dV
__global__ __launch_bounds__(256) void k_adj_objective(AdjPush A, double mdv, double gout) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < A.n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const double w = A.w[p];
        double lu[3] = {0.0, 0.0, 0.0}, lw = 0.0;
        if (w == w) {
            const double ux = A.ux[p], uy = A.uy[p], uz = A.uz[p];
            const double gam = ::sqrt(1.0 + (ux * ux + uy * uy + uz * uz));
            lu[0] = gout * mdv * w * ux / gam;
            lu[1] = gout * mdv * w * uy / gam;
            lu[2] = gout * mdv * w * uz / gam;
            lw = gout * mdv * (gam - 1.0);
        }
        A.lx[p] = 0.0; A.ly[p] = 0.0; A.lz[p] = 0.0;
        A.lux[p] = lu[0]; A.luy[p] = lu[1]; A.luz[p] = lu[2];
        A.lw[p] = lw;
    }
}

// FP32: the same from the fp32 state (removed slots carry no adjoint)
__global__ __launch_bounds__(256) void k_adj_objective32(AdjPush32 A, double mdv, double gout) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < A.n;
         p += (int64_t)gridDim.x * blockDim.x) {
        double lu[3] = {0.0, 0.0, 0.0}, lw = 0.0;
        if (A.cell[p] >= 0) {
            const double w = A.w[p], ux = A.ux[p], uy = A.uy[p], uz = A.uz[p];
            const double gam = ::sqrt(1.0 + (ux * ux + uy * uy + uz * uz));
            lu[0] = gout * mdv * w * ux / gam;
            lu[1] = gout * mdv * w * uy / gam;
            lu[2] = gout * mdv * w * uz / gam;
            lw = gout * mdv * (gam - 1.0);
        }
        A.lx[p] = 0.0f; A.ly[p] = 0.0f; A.lz[p] = 0.0f;
        A.lux[p] = (float)lu[0]; A.luy[p] = (float)lu[1]; A.luz[p] = (float)lu[2];
        A.lw[p] = lw;
    }
}

// ---------------------------------------------------------------------
// field adjoints (gathers over the adjoint arrays)
// ---------------------------------------------------------------------
__device__ __forceinline__ bool inner(const Grid& g, int k, int j, int i) {
    return i >= g.m && i < g.m + g.n[0] && j >= g.m && j < g.m + g.n[1] && k >= g.m && k < g.m + g.n[2];
}

// transpose of E += c (curl B - J) (k_e_half): lB += .., lJ -= c lE
template <typename T>
__global__ __launch_bounds__(256) void k_adj_e(const T* __restrict__ lEx, const T* __restrict__ lEy,
                                               const T* __restrict__ lEz, T* __restrict__ lBx,
                                               T* __restrict__ lBy, T* __restrict__ lBz,
                                               T* __restrict__ lJx, T* __restrict__ lJy,
                                               T* __restrict__ lJz, Grid g, T c, T dx, T dy, T dz) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.fsize) return;
    const int i = (int)(t % g.fe[0]), j = (int)((t / g.fe[0]) % g.fe[1]), k = (int)(t / ((int64_t)g.fe[0] * g.fe[1]));
    auto L = [&](const T* a, int kk, int jj, int ii) -> T {
        return inner(g, kk, jj, ii) ? a[g.fi(kk, jj, ii)] : T(0);
    };
    lBz[t] += c * (L(lEx, k, j, i) / dy - L(lEx, k, j + 1, i) / dy - L(lEy, k, j, i) / dx + L(lEy, k, j, i + 1) / dx);
    lBy[t] += c * (-L(lEx, k, j, i) / dz + L(lEx, k + 1, j, i) / dz + L(lEz, k, j, i) / dx - L(lEz, k, j, i + 1) / dx);
    lBx[t] += c * (L(lEy, k, j, i) / dz - L(lEy, k + 1, j, i) / dz - L(lEz, k, j, i) / dy + L(lEz, k, j + 1, i) / dy);
    if (inner(g, k, j, i)) {
        const int64_t q = g.ji(k, j, i);
        lJx[q] -= c * lEx[t];
        lJy[q] -= c * lEy[t];
        lJz[q] -= c * lEz[t];
    }
}

// transpose of B -= c curl E (k_b_half): lE += ..
template <typename T>
__global__ __launch_bounds__(256) void k_adj_b(T* __restrict__ lEx, T* __restrict__ lEy,
                                               T* __restrict__ lEz, const T* __restrict__ lBx,
                                               const T* __restrict__ lBy, const T* __restrict__ lBz,
                                               Grid g, T c, T dx, T dy, T dz) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.fsize) return;
    const int i = (int)(t % g.fe[0]), j = (int)((t / g.fe[0]) % g.fe[1]), k = (int)(t / ((int64_t)g.fe[0] * g.fe[1]));
    auto L = [&](const T* a, int kk, int jj, int ii) -> T {
        return inner(g, kk, jj, ii) ? a[g.fi(kk, jj, ii)] : T(0);
    };
    lEx[t] += c * (-L(lBy, k - 1, j, i) / dz + L(lBy, k, j, i) / dz + L(lBz, k, j - 1, i) / dy - L(lBz, k, j, i) / dy);
    lEy[t] += c * (L(lBx, k - 1, j, i) / dz - L(lBx, k, j, i) / dz - L(lBz, k, j, i - 1) / dx + L(lBz, k, j, i) / dx);
    lEz[t] += c * (-L(lBx, k, j - 1, i) / dy + L(lBx, k, j, i) / dy + L(lBy, k, j, i - 1) / dx - L(lBy, k, j, i) / dx);
}

// transpose of the halo wrap of three components along one axis
// (k_wrap): the ghost's adjoint moves onto its source
template <typename T>
__global__ __launch_bounds__(256) void k_adj_wrap(T* __restrict__ a0, T* __restrict__ a1,
                                                  T* __restrict__ a2, Grid g, int axis) {
    const int m = g.m, N = g.n[axis];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int e0, e1;
    if (axis == 0) { e0 = g.fe[1]; e1 = g.fe[2]; }
    else if (axis == 1) { e0 = g.fe[0]; e1 = g.fe[2]; }
    else { e0 = g.fe[0]; e1 = g.fe[1]; }
    if (t >= (int64_t)2 * m * e0 * e1) return;
    const int gi = (int)(t % (2 * m));
    const int64_t r = t / (2 * m);
    const int u = (int)(r % e0), v = (int)(r / e0);
    const int dst = gi < m ? gi : m + N + (gi - m);
    const int src = gi < m ? N + gi : m + (gi - m);
    int64_t od, os;
    if (axis == 0) { od = g.fi(v, u, dst); os = g.fi(v, u, src); }
    else if (axis == 1) { od = g.fi(v, dst, u); os = g.fi(v, src, u); }
    else { od = g.fi(dst, v, u); os = g.fi(src, v, u); }
    T* A[3] = {a0, a1, a2};
    for (int q = 0; q < 3; q++) {
        A[q][os] += A[q][od];
        A[q][od] = T(0);
    }
}

// transpose of the current fold (k_fold): phase 1 = sync, phase 0 = accumulate
template <typename T>
__global__ __launch_bounds__(256) void k_adj_fold(T* __restrict__ a0, T* __restrict__ a1,
                                                  T* __restrict__ a2, Grid g, int axis, int phase) {
    const int m = g.m, N = g.n[axis];
    const int st = m, end = m + N - 1;
    int e0, e1;
    if (axis == 0) { e0 = g.je[1]; e1 = g.je[2]; }
    else if (axis == 1) { e0 = g.je[0]; e1 = g.je[2]; }
    else { e0 = g.je[0]; e1 = g.je[1]; }
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)(m + 1) * e0 * e1) return;
    const int l = (int)(t % (m + 1));
    const int64_t r = t / (m + 1);
    const int u = (int)(r % e0) - 1, v = (int)(r / e0) - 1;
    auto idx = [&](int s) -> int64_t {
        if (axis == 0) return g.ji(v, u, s);
        if (axis == 1) return g.ji(v, s, u);
        return g.ji(s, v, u);
    };
    const int64_t i_lo_int = idx(st + l), i_hi_gh = idx(end + 1 + l);
    const int64_t i_hi_int = idx(end - m + l), i_lo_gh = idx(-1 + l);
    T* A[3] = {a0, a1, a2};
    for (int q = 0; q < 3; q++) {
        T* D = A[q];
        if (phase == 1) {
            D[i_hi_int] += D[i_lo_gh];
            D[i_lo_gh] = T(0);
            D[i_lo_int] += D[i_hi_gh];
            D[i_hi_gh] = T(0);
        } else {
            D[i_hi_gh] += D[i_lo_int];
            D[i_lo_gh] += D[i_hi_int];
        }
    }
}

// transpose of the laser / Silver-Mueller writes (k_laser) and the
// derivative of the objective with respect to this step's x-min envelope
// value; g_row(y, z) = Gpulse(t, 0, y, z) / envelope.
template <typename T>
__global__ __launch_bounds__(256) void k_adj_laser(T* __restrict__ lEy, T* __restrict__ lEz,
                                                   T* __restrict__ lBx, T* __restrict__ lBy,
                                                   T* __restrict__ lBz, T* __restrict__ lJy,
                                                   T* __restrict__ lJz, Grid g, LaserArgs L,
                                                   double* __restrict__ genv) {
    const int j = g.m + blockIdx.x * blockDim.x + threadIdx.x;
    const int k = g.m + blockIdx.y;
    double ge = 0.0;
    if (j < g.m + g.n[1]) {
        const int m = g.m, N = g.n[0];
        const double dt = L.dt;
        const double alpha = 1 / (1 + dt / L.dx);
        const double beta = 1 - dt / L.dx;
        const double dtdz = dt / L.dz, dtdy = dt / L.dy;
        const double yy = ((double)j + 0.0) * L.dy - L.yfocus;
        const double zz = ((double)k + 0.0) * L.dz - L.zfocus;
        // left: column m-1 from column m
        {
            const int64_t o = g.fi(k, j, m), gl = g.fi(k, j, m - 1);
            const double a = lBz[gl], b = lBy[gl];
            lBz[gl] = T(0);
            lBy[gl] = T(0);
            hbm_add(&lEy[o], (T)(-2.0 * alpha * a));
            hbm_add(&lBx[o], (T)(-alpha * dtdz * a + alpha * dtdy * b));
            hbm_add(&lBx[g.fi(k - 1, j, m)], (T)(alpha * dtdz * a));
            hbm_add(&lBx[g.fi(k, j - 1, m)], (T)(-alpha * dtdy * b));
            hbm_add(&lBz[o], (T)(-alpha * beta * a));
            hbm_add(&lBy[o], (T)(-alpha * beta * b));
            hbm_add(&lEz[o], (T)(2.0 * alpha * b));
            hbm_add(&lJy[g.ji(k, j, m)], (T)(-alpha * dt * a));
            hbm_add(&lJz[g.ji(k, j, m)], (T)(alpha * dt * b));
            if (L.left) {
                // S_l = P_l = envelope * g_row for the enabled polarisations
                LaserArgs U = L;
                U.env_on = 1;
                U.env_val = 1.0;
                const double gr = d_gpulse(U, L.time, 0.0, yy, zz);
                if (L.pol == 1 || L.pol == 3) ge += 4.0 * alpha * a * gr;
                if (L.pol == 2 || L.pol == 3) ge += -4.0 * alpha * b * gr;
            }
        }
        // right: column L.xw from columns L.xre (E, Bx, J) and L.xrb (By,
        // Bz) (the sources there do not depend on the x-min envelope)
        {
            const int64_t o = g.fi(k, j, L.xre), ob = g.fi(k, j, L.xrb), gr = g.fi(k, j, L.xw);
            const double a = lBz[gr], b = lBy[gr];
            lBz[gr] = T(0);
            lBy[gr] = T(0);
            hbm_add(&lEy[o], (T)(2.0 * alpha * a));
            hbm_add(&lBx[o], (T)(-alpha * dtdz * a + alpha * dtdy * b));
            hbm_add(&lBx[g.fi(k - 1, j, L.xre)], (T)(alpha * dtdz * a));
            hbm_add(&lBx[g.fi(k, j - 1, L.xre)], (T)(-alpha * dtdy * b));
            hbm_add(&lBz[ob], (T)(-alpha * beta * a));
            hbm_add(&lBy[ob], (T)(-alpha * beta * b));
            hbm_add(&lEz[o], (T)(-2.0 * alpha * b));
            hbm_add(&lJy[g.ji(k, j, L.xre)], (T)(alpha * dt * a));
            hbm_add(&lJz[g.ji(k, j, L.xre)], (T)(-alpha * dt * b));
        }
    }
    // block sum of ge
    __shared__ double red[256];
    red[threadIdx.x] = ge;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] != 0.0) hbm_add(genv, red[0]);
}

}  // namespace pic

// ---------------------------------------------------------------------
// objective: final kinetic energy of the species (FP64 sum)
// ---------------------------------------------------------------------
namespace pic {

__global__ __launch_bounds__(256) void k_adj_kinetic(const double* __restrict__ w, const double* __restrict__ ux,
                                                     const double* __restrict__ uy, const double* __restrict__ uz,
                                                     int64_t n, double* __restrict__ out) {
    double s = 0.0;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const double wp = w[p];
        if (!(wp == wp)) continue;
        s += wp * (::sqrt(1.0 + (ux[p] * ux[p] + uy[p] * uy[p] + uz[p] * uz[p])) - 1.0);
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) hbm_add(out, red[0]);
}

// set the weights of live slots (dead slots stay dead)
__global__ __launch_bounds__(256) void k_adj_set_w(double* __restrict__ w, const double* __restrict__ win, int64_t n) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        if (w[p] == w[p]) w[p] = win[p];
}

// FP32 state: sum w (gamma - 1) over live slots, in FP64
__global__ __launch_bounds__(256) void k_adj_kinetic32(const int32_t* __restrict__ cell, const float* __restrict__ w,
                                                       const float* __restrict__ ux, const float* __restrict__ uy,
                                                       const float* __restrict__ uz, int64_t n,
                                                       double* __restrict__ out) {
    double s = 0.0;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        if (cell[p] < 0) continue;
        const double x = ux[p], y = uy[p], z = uz[p];
        s += (double)w[p] * (::sqrt(1.0 + (x * x + y * y + z * z)) - 1.0);
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) hbm_add(out, red[0]);
}

__global__ __launch_bounds__(256) void k_adj_set_w32(const int32_t* __restrict__ cell, float* __restrict__ w,
                                                     const double* __restrict__ win, int64_t n) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        if (cell[p] >= 0) w[p] = (float)win[p];
}

// ---------------------------------------------------------------------
// checkpoint / recompute driver
// ---------------------------------------------------------------------
// the particle adjoints of a differentiated run (one precision)
template <typename T> struct DiffLam {
    T* l[6];
    double* w;
};

struct DiffState {
    int s = -1;
    int f32 = 0;                     // FP32 context: fp32 state, fp32 adjoints
    int kind = 0;                    // tile shape kind of the run's grouping
    int64_t slots = 0;               // particle slots of species s (live and dead)
    int first = 1, nsteps = 0, K = 1;
    int forward_done = 0;
    int saved_push_kernel = 0;
    int saved_sort_interval = 1;
    int kernel = 2;                  // push kernel of the run (1 direct, 2 tiled)
    // Tiled runs sort the species every sort_every steps (before the push)
    // with a keep-dead sort, which preserves the slot count: the forward
    // records each sort's permutation and tile offsets, the recompute
    // replays them (so forward and recompute agree on every slot's
    // particle), and the adjoint carries the particle adjoints back
    // through each permutation.
    int sort_every = 1;
    int32_t* perms = nullptr;        // one permutation (slots entries) per sort step
    int64_t* toffs = nullptr;        // tile offsets (ntiles + 2) after each sort step
    int nsorts = 0;
    bool sorts_at(int i) const { return kernel == 2 && i % sort_every == 0; }
    std::vector<double> env;         // envelope per step (empty: the analytic one)
    double objective = 0.0;
    // snapshots: [0] the run's initial state, [i] the state before step first + i*K
    std::vector<char*> ck;
    size_t ck_bytes = 0;
    // the species' scrambled flag (in-tile storage order) of each snapshot
    // taken, restored with it
    std::map<const char*, int> scr;
    char* fin = nullptr;             // the state at the end of the forward
    // per-step buffers of the segments being recomputed and reversed
    // (double-buffered: segment k - 1 is recomputed on the context's stream
    // while the adjoint of segment k runs on the adjoint stream)
    char* segb[2] = {};
    int seg_steps = 0;
    size_t step_bytes = 0;
    hipStream_t as = nullptr;        // the adjoint stream (default priority)
    hipEvent_t ev_rec[2] = {};       // segment buffer b recomputed
    hipEvent_t ev_adj[2] = {};       // segment buffer b reversed (free again)
    hipEvent_t ev_obj = nullptr;     // the final-state adjoint taken
    int64_t* atoff = nullptr;        // the adjoint's tile offsets and work list
    int64_t* awork = nullptr;
    int64_t* anwork = nullptr;
    // adjoint state, in the precision of the context
    void* lF[6] = {};
    void* lJ[3] = {};
    void* lp[6] = {};                // x y z (FP32: the offsets) ux uy uz
    double* lw = nullptr;            // accumulated d objective / d w (FP64)
    void* lp_tmp[6] = {};            // targets of the un-permutation
    double* lw_tmp = nullptr;
    double* genv = nullptr;          // per step (device)
    int64_t* stray = nullptr;        // tiled adjoint: slots outside their tile (+ count)
    double* red = nullptr;           // reduction scratch
    size_t es() const { return f32 ? 4 : 8; }  // bytes of a field / particle value
    // particle columns of a snapshot: FP64 x y z ux uy uz w;
    // FP32 cell fx fy fz ux uy uz w
    int ncol() const { return f32 ? 8 : 7; }
    int wcol() const { return f32 ? 7 : 6; }
};

namespace {

size_t up256(size_t b) { return (b + 255) / 256 * 256; }

int dalloc(pic_ctx* c, void** p, size_t bytes) {
    if (hipMalloc(p, std::max<size_t>(up256(bytes), 256)) != hipSuccess) {
        *p = nullptr;
        return set_err(c, PIC_ERR_NOMEM, "differentiated run: device allocation of %zu bytes failed", bytes);
    }
    return PIC_OK;
}

int fill0(pic_ctx* c, void* p, size_t bytes) { return launch_fill(c, p, (int64_t)bytes); }

__global__ __launch_bounds__(256) void k_adj_zero(uint4* __restrict__ p, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// zero 'bytes' of a dalloc'd buffer (whole 16-byte words) on the adjoint stream
int zero_adj(pic_ctx* c, DiffState& D, void* p, size_t bytes) {
    const int64_t n16 = (int64_t)((bytes + 15) / 16);
    if (n16 == 0) return PIC_OK;
    k_adj_zero<<<(int)std::min<int64_t>(blocks_for(n16, 256), 8192), 256, 0, D.as>>>((uint4*)p, n16);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

// particle column f of species s (current buffer), in snapshot order
void* pcol(pic_ctx* c, int s, int f) {
    Species& S = c->sp[s];
    const int b = S.cur;
    if (c->fp32) {
        void* a[8] = {S.cell[b], S.fx[b], S.fy[b], S.fz[b], S.vx[b], S.vy[b], S.vz[b], S.fw[b]};
        return a[f];
    }
    void* a[7] = {S.x[b], S.y[b], S.z[b], S.ux[b], S.uy[b], S.uz[b], S.w[b]};
    return a[f];
}

// state <-> snapshot: 6 fields, 3 currents, the particle columns, counters
int copy_state(pic_ctx* c, DiffState& D, char* snap, bool save) {
    const Grid& g = c->g;
    const size_t es = D.es(), fb = g.fsize * es, jb = g.jsize * es, pb = D.slots * es;
    size_t off = 0;
    auto mv = [&](void* live, size_t bytes) -> int {
        void* dst = save ? (void*)(snap + off) : live;
        const void* src = save ? (const void*)live : (const void*)(snap + off);
        if (bytes && hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
            return set_err(c, PIC_ERR_HIP, "differentiated run: snapshot copy failed");
        off += up256(bytes);
        return PIC_OK;
    };
    int r = PIC_OK;
    for (int f = 0; f < 6 && !r; f++) r = mv(c->F[f], fb);
    for (int f = 0; f < 3 && !r; f++) r = mv(c->J[f], jb);
    for (int f = 0; f < D.ncol() && !r; f++) r = mv(pcol(c, D.s, f), pb);
    if (!r) r = mv(c->sp[D.s].d_cnt, CNT_N * 8);
    if (!r) r = mv(c->sp[D.s].tile_off, (size_t)(c->tl.ntiles + 2) * 8);
    if (!r) {
        int& f = c->sp[D.s].scrambled;
        if (save) D.scr[snap] = f;
        else if (D.scr.count(snap)) f = D.scr[snap];
    }
    return r;
}

size_t snapshot_bytes(const pic_ctx* c, const DiffState& D) {
    const size_t es = D.es();
    return 6 * up256(c->g.fsize * es) + 3 * up256(c->g.jsize * es) + D.ncol() * up256(D.slots * es) +
           up256(CNT_N * 8) + up256((size_t)(c->tl.ntiles + 2) * 8);
}

// FP64/FP32 particle columns moved by a sort: the snapshot columns, the
// ids and (FP64, when kept) x_old
struct GatherCols {
    const void* src[12];
    void* dst[12];
    int es[12];
    int n;
};

__global__ __launch_bounds__(256) void k_gather_cols(GatherCols G, const int32_t* __restrict__ perm, int64_t slots) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < slots;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = perm[q];
        for (int k = 0; k < G.n; k++) {
            if (G.es[k] == 8) ((int64_t*)G.dst[k])[q] = ((const int64_t*)G.src[k])[p];
            else ((int32_t*)G.dst[k])[q] = ((const int32_t*)G.src[k])[p];
        }
    }
}

// the adjoint of a permutation: lambda_before[perm[q]] = lambda_after[q]
template <typename T>
__global__ __launch_bounds__(256) void k_adj_unpermute(DiffLam<T> a, DiffLam<T> b, const int32_t* __restrict__ perm,
                                                       int64_t slots) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < slots;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = perm[q];
        for (int k = 0; k < 6; k++) b.l[k][p] = a.l[k][q];
        b.w[p] = a.w[q];
    }
}

// Replay of a recorded sort: the columns permuted into the other buffer,
// which becomes current, and the tile offsets restored.
int replay_sort(pic_ctx* c, DiffState& D, const int32_t* perm, const int64_t* toff) {
    Species& S = c->sp[D.s];
    GatherCols G{};
    const int a = S.cur, b = 1 - S.cur;
    auto add = [&](const void* x, void* y, int es) {
        if (x && y) { G.src[G.n] = x; G.dst[G.n] = y; G.es[G.n] = es; G.n++; }
    };
    if (c->fp32) {
        add(S.cell[a], S.cell[b], 4); add(S.fx[a], S.fx[b], 4); add(S.fy[a], S.fy[b], 4); add(S.fz[a], S.fz[b], 4);
        add(S.vx[a], S.vx[b], 4); add(S.vy[a], S.vy[b], 4); add(S.vz[a], S.vz[b], 4); add(S.fw[a], S.fw[b], 4);
    } else {
        add(S.x[a], S.x[b], 8); add(S.y[a], S.y[b], 8); add(S.z[a], S.z[b], 8); add(S.ux[a], S.ux[b], 8);
        add(S.uy[a], S.uy[b], 8); add(S.uz[a], S.uz[b], 8); add(S.w[a], S.w[b], 8);
        if (c->keep_old) { add(S.xo[a], S.xo[b], 8); add(S.yo[a], S.yo[b], 8); add(S.zo[a], S.zo[b], 8); }
    }
    add(S.id[a], S.id[b], 8);
    if (D.slots) {
        k_gather_cols<<<(int)std::min<int64_t>(blocks_for(D.slots, 256), 8192), 256, 0, c->stream>>>(G, perm, D.slots);
        PIC_LAUNCH_CHECK(c);
    }
    PIC_HIP(c, hipMemcpyAsync(S.tile_off, toff, (size_t)(c->tl.ntiles + 2) * 8, hipMemcpyDeviceToDevice, c->stream));
    S.cur = b;
    S.sorted = 1;
    S.cellsorted = 0;  // a recorded tile sort
    S.pushes_since_sort = 0;
    return PIC_OK;
}

// offset of particle column f in a snapshot
size_t snap_col(const pic_ctx* c, const DiffState& D, int f) {
    const size_t es = D.es();
    return 6 * up256(c->g.fsize * es) + 3 * up256(c->g.jsize * es) + f * up256(D.slots * es);
}

// one step of run_main_time_loop (simulation.cpp:159-183) with the
// envelope value e; 'save' (may be null) receives the particle state and
// the fields at the push
// The sort before step i's push: recorded (forward) or replayed (recompute).
int diff_sort(pic_ctx* c, DiffState& D, int i, bool replay) {
    if (!D.sorts_at(i)) return PIC_OK;
    const int k = i / D.sort_every;
    int32_t* perm = D.perms + (size_t)k * D.slots;
    int64_t* toff = D.toffs + (size_t)k * (c->tl.ntiles + 2);
    if (replay) return replay_sort(c, D, perm, toff);
    int r = launch_sort(c, D.s, perm);
    if (r) return r;
    PIC_HIP(c, hipMemcpyAsync(toff, c->sp[D.s].tile_off, (size_t)(c->tl.ntiles + 2) * 8, hipMemcpyDeviceToDevice,
                              c->stream));
    return PIC_OK;
}

// bytes of one step's saved push inputs: particle columns, fields, tile offsets
size_t step_state_bytes(const pic_ctx* c, const DiffState& D) {
    const size_t es = D.es();
    return D.ncol() * up256(D.slots * es) + 6 * up256(c->g.fsize * es) + up256((size_t)(c->tl.ntiles + 2) * 8);
}

int diff_step(pic_ctx* c, DiffState& D, int st, char* save, bool replay) {
    int r;
    c->env_on = D.env.empty() ? 0 : 1;
    c->env_val = D.env.empty() ? 0.0 : D.env[st - D.first];
    r = pic_laser_boundary(c, st);
    c->env_on = 0;
    if (r) return r;
    if ((r = pic_fields_first(c))) return r;
    if ((r = pic_clear_currents(c))) return r;
    if ((r = diff_sort(c, D, st - D.first, replay))) return r;
    if (save) {
        const size_t es = D.es(), pb = up256(D.slots * es), fb = up256(c->g.fsize * es);
        for (int f = 0; f < D.ncol(); f++)
            if (D.slots)
                PIC_HIP(c, hipMemcpyAsync(save + f * pb, pcol(c, D.s, f), D.slots * es, hipMemcpyDeviceToDevice, c->stream));
        for (int f = 0; f < 6; f++)
            PIC_HIP(c, hipMemcpyAsync(save + D.ncol() * pb + f * fb, c->F[f], c->g.fsize * es,
                                      hipMemcpyDeviceToDevice, c->stream));
        PIC_HIP(c, hipMemcpyAsync(save + D.ncol() * pb + 6 * fb, c->sp[D.s].tile_off,
                                  (size_t)(c->tl.ntiles + 2) * 8, hipMemcpyDeviceToDevice, c->stream));
    }
    if ((r = launch_push(c, D.s))) return r;
    if ((r = pic_particle_boundary(c, D.s, nullptr))) return r;
    if ((r = pic_fold_currents(c))) return r;
    return pic_fields_second(c);
}

template <typename T> int adj_wrap3(pic_ctx* c, DiffState& D, void* const* a) {
    const Grid& g = c->g;
    for (int ax = 2; ax >= 0; ax--) {
        if (!c->p.periodic[ax]) continue;
        const int e0 = ax == 0 ? g.fe[1] : g.fe[0], e1 = ax == 2 ? g.fe[1] : g.fe[2];
        const int64_t total = (int64_t)2 * g.m * e0 * e1;
        k_adj_wrap<T><<<blocks_for(total, 256), 256, 0, D.as>>>((T*)a[0], (T*)a[1], (T*)a[2], g, ax);
        PIC_LAUNCH_CHECK(c);
    }
    return PIC_OK;
}

template <typename T> int adj_e(pic_ctx* c, DiffState& D, double coef) {
    const Grid& g = c->g;
    T* const* l = (T* const*)D.lF;
    T* const* j = (T* const*)D.lJ;
    k_adj_e<T><<<blocks_for(g.fsize, 256), 256, 0, D.as>>>(l[0], l[1], l[2], l[3], l[4], l[5], j[0], j[1],
                                                                 j[2], g, (T)coef, (T)c->p.dx, (T)c->p.dy,
                                                                 (T)c->p.dz);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

template <typename T> int adj_b(pic_ctx* c, DiffState& D, double coef) {
    const Grid& g = c->g;
    T* const* l = (T* const*)D.lF;
    k_adj_b<T><<<blocks_for(g.fsize, 256), 256, 0, D.as>>>(l[0], l[1], l[2], l[3], l[4], l[5], g, (T)coef,
                                                                 (T)c->p.dx, (T)c->p.dy, (T)c->p.dz);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

template <typename T> int adj_fold(pic_ctx* c, DiffState& D) {
    const Grid& g = c->g;
    for (int phase = 1; phase >= 0; phase--)
        for (int ax = 2; ax >= 0; ax--) {
            if (!c->p.periodic[ax]) continue;
            const int e0 = ax == 0 ? g.je[1] : g.je[0], e1 = ax == 2 ? g.je[1] : g.je[2];
            const int64_t total = (int64_t)(g.m + 1) * e0 * e1;
            k_adj_fold<T><<<blocks_for(total, 256), 256, 0, D.as>>>((T*)D.lJ[0], (T*)D.lJ[1], (T*)D.lJ[2], g,
                                                                         ax, phase);
            PIC_LAUNCH_CHECK(c);
        }
    return PIC_OK;
}

// FP64 particle view of the push at one step: the saved state (save) or
// the live one (null)
AdjPush adj_view(pic_ctx* c, DiffState& D, const char* save) {
    AdjPush A;
    const size_t pb = up256(D.slots * 8), fb = up256(c->g.fsize * 8);
    const double* P[7];
    for (int f = 0; f < 7; f++) P[f] = save ? (const double*)(save + f * pb) : (const double*)pcol(c, D.s, f);
    A.x = P[0]; A.y = P[1]; A.z = P[2]; A.ux = P[3]; A.uy = P[4]; A.uz = P[5]; A.w = P[6];
    for (int f = 0; f < 6; f++) A.F[f] = save ? (const double*)(save + 7 * pb + f * fb) : (const double*)c->F[f];
    for (int f = 0; f < 3; f++) A.lJ[f] = (const double*)D.lJ[f];
    for (int f = 0; f < 6; f++) A.lF[f] = (double*)D.lF[f];
    double* const* lp = (double* const*)D.lp;
    A.lx = lp[0]; A.ly = lp[1]; A.lz = lp[2]; A.lux = lp[3]; A.luy = lp[4]; A.luz = lp[5];
    A.lw = D.lw;
    A.n = D.slots;
    return A;
}

// FP32 particle view, likewise
AdjPush32 adj_view32(pic_ctx* c, DiffState& D, const char* save) {
    AdjPush32 A;
    const size_t pb = up256(D.slots * 4), fb = up256(c->g.fsize * 4);
    const void* P[8];
    for (int f = 0; f < 8; f++) P[f] = save ? (const void*)(save + f * pb) : (const void*)pcol(c, D.s, f);
    A.cell = (const int32_t*)P[0];
    A.fx = (const float*)P[1]; A.fy = (const float*)P[2]; A.fz = (const float*)P[3];
    A.ux = (const float*)P[4]; A.uy = (const float*)P[5]; A.uz = (const float*)P[6]; A.w = (const float*)P[7];
    for (int f = 0; f < 6; f++) A.F[f] = save ? (const float*)(save + 8 * pb + f * fb) : (const float*)c->F[f];
    for (int f = 0; f < 3; f++) A.lJ[f] = (const float*)D.lJ[f];
    for (int f = 0; f < 6; f++) A.lF[f] = (float*)D.lF[f];
    float* const* lp = (float* const*)D.lp;
    A.lx = lp[0]; A.ly = lp[1]; A.lz = lp[2]; A.lux = lp[3]; A.luy = lp[4]; A.luz = lp[5];
    A.lw = D.lw;
    A.n = D.slots;
    return A;
}

AdjK<float> adj_consts32(const pic_ctx* c, const Species& S) {
    AdjK<float> K;
    K.q = (float)S.q;
    K.qm = (float)(S.q / S.m * (c->p.dt / 2.0));
    const double d[3] = {c->p.dx, c->p.dy, c->p.dz};
    for (int a = 0; a < 3; a++) {
        K.dtd[a] = (float)(c->p.dt / d[a]);
        K.jc[a] = (float)(S.q * d[a] / c->p.dt);
    }
    return K;
}

int tile_kind(const pic_ctx* c) { return c->fp32 ? (c->fix32 ? 2 : 1) : 0; }

// Parts per tile (split along x).  Whole tiles: a split of the 8x4x4
// quartic tile that let two workgroups' LDS fit a CU (the FP32 adjoint then
// held 357 registers, one wave per SIMD) needed the registers capped at two
// waves per SIMD, which spilled and measured slower; the 4x4x4 quartic tile
// of the fixed-point FP32 push fits two workgroups whole (AdjView32).
template <int M, int KIND> struct AdjSplit { static constexpr int S = 1; };

template <int M, int KIND, typename V, typename LACC = double>
int launch_adj_tiled(pic_ctx* c, DiffState& D, const V& v) {
    using TS = TileShape<KIND, M>;
    constexpr int S = AdjSplit<M, KIND>::S;
    using B = AdjBox<M, TS::X / S, TS::Y, TS::Z, typename V::T, LACC>;
    Species& Sp = c->sp[D.s];
    int r = build_work_on(c, D.atoff, PIC_PUSH_CHUNK, D.awork, D.anwork, D.as);
    if (r) return r;
    PIC_HIP(c, hipFuncSetAttribute((const void*)k_adj_push_tiled<M, TS::X, TS::Y, TS::Z, V, LACC, S>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)B::LDS));
    const int64_t max_items = (int64_t)c->tl.ntiles + Sp.cap / PIC_PUSH_CHUNK + 1;
    int64_t* cnt = D.stray + std::max<int64_t>(D.slots, 1);
    if ((r = zero_adj(c, D, cnt, 16))) return r;
    k_adj_push_tiled<M, TS::X, TS::Y, TS::Z, V, LACC, S><<<(unsigned)(max_items * S), 256, B::LDS, D.as>>>(
        v, c->g, c->tl, D.atoff, D.awork, D.anwork, PIC_PUSH_CHUNK, D.stray, cnt);
    PIC_LAUNCH_CHECK(c);
    k_adj_push_hbm<M, V><<<1024, 256, 0, D.as>>>(v, c->g, D.stray, cnt);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

// adjoint push: the LDS-tiled kernel when the run's forward is tiled, the
// separable HBM kernel for the direct forward (or PIC_OPT_ADJ_PUSH = 2),
// and (FP64, PIC_OPT_ADJ_PUSH = 1) the dual-number kernel, an independent
// evaluation of the same derivatives
int launch_adj_push(pic_ctx* c, DiffState& D, const char* save) {
    const int blocks = (int)std::min<int64_t>(blocks_for(std::max<int64_t>(D.slots, 1), 256), 8192);
    const bool q = c->g.m == 2;
    const bool tiled = D.kernel == 2 && c->adj_push == 0;
    if (tiled && tile_kind(c) != D.kind)
        return set_err(c, PIC_ERR_STATE, "differentiated run: the deposit accumulator option changed during the run");
    if (D.f32) {
        AdjView32 v{adj_view32(c, D, save), adj_consts32(c, c->sp[D.s]), {c->g.n[0], c->g.n[1], c->g.n[2]}, c->g.m,
                    CellSplit(c->g.n[0], c->g.n[1])};
        if (tiled) {
            if (D.kind == 2) return q ? launch_adj_tiled<2, 2>(c, D, v) : launch_adj_tiled<4, 2>(c, D, v);
            return q ? launch_adj_tiled<2, 1>(c, D, v) : launch_adj_tiled<4, 1>(c, D, v);
        }
        if (q) k_adj_push_hbm<2><<<blocks, 256, 0, D.as>>>(v, c->g, nullptr, nullptr);
        else k_adj_push_hbm<4><<<blocks, 256, 0, D.as>>>(v, c->g, nullptr, nullptr);
        PIC_LAUNCH_CHECK(c);
        return PIC_OK;
    }
    const PushConsts K = push_consts64(c, c->sp[D.s]);
    const AdjPush A = adj_view(c, D, save);
    if (c->adj_push == 1) {
        if (q) k_adj_push64<2><<<blocks, 256, 0, D.as>>>(A, c->g, K);
        else k_adj_push64<4><<<blocks, 256, 0, D.as>>>(A, c->g, K);
        PIC_LAUNCH_CHECK(c);
        return PIC_OK;
    }
    AdjView64 v{A, K, adj_consts64(K), c->g.m};
    if (tiled) return q ? launch_adj_tiled<2, 0>(c, D, v) : launch_adj_tiled<4, 0>(c, D, v);
    if (q) k_adj_push_hbm<2><<<blocks, 256, 0, D.as>>>(v, c->g, nullptr, nullptr);
    else k_adj_push_hbm<4><<<blocks, 256, 0, D.as>>>(v, c->g, nullptr, nullptr);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

// reverse of diff_step for step st, from the saved push-time state
template <typename T> int adj_step_t(pic_ctx* c, DiffState& D, int st, const char* save) {
    const double h = c->p.dt / 2.0;
    void* lE[3] = {D.lF[0], D.lF[1], D.lF[2]};
    void* lB[3] = {D.lF[3], D.lF[4], D.lF[5]};
    int r;
    // fields_second: B then E
    if ((r = adj_wrap3<T>(c, D, lE)) || (r = adj_e<T>(c, D, h))) return r;
    if ((r = adj_wrap3<T>(c, D, lB)) || (r = adj_b<T>(c, D, h))) return r;
    // fold, boundary (dead slots carry no adjoint), push
    if ((r = adj_fold<T>(c, D))) return r;
    if (D.kernel == 2) {  // the tile grouping of this step's push
        const size_t pb = up256(D.slots * sizeof(T)), fb = up256(c->g.fsize * sizeof(T));
        PIC_HIP(c, hipMemcpyAsync(D.atoff, save + D.ncol() * pb + 6 * fb, (size_t)(c->tl.ntiles + 2) * 8,
                                  hipMemcpyDeviceToDevice, D.as));
    }
    if ((r = launch_adj_push(c, D, save))) return r;
    // the sort before the push
    const int i = st - D.first;
    if (D.sorts_at(i) && D.slots) {
        const int32_t* perm = D.perms + (size_t)(i / D.sort_every) * D.slots;
        DiffLam<T> a{{(T*)D.lp[0], (T*)D.lp[1], (T*)D.lp[2], (T*)D.lp[3], (T*)D.lp[4], (T*)D.lp[5]}, D.lw};
        DiffLam<T> b{{(T*)D.lp_tmp[0], (T*)D.lp_tmp[1], (T*)D.lp_tmp[2], (T*)D.lp_tmp[3], (T*)D.lp_tmp[4],
                      (T*)D.lp_tmp[5]}, D.lw_tmp};
        k_adj_unpermute<T><<<(int)std::min<int64_t>(blocks_for(D.slots, 256), 8192), 256, 0, D.as>>>(
            a, b, perm, D.slots);
        PIC_LAUNCH_CHECK(c);
        for (int f = 0; f < 6; f++) std::swap(D.lp[f], D.lp_tmp[f]);
        std::swap(D.lw, D.lw_tmp);
    }
    // clear: the currents before it get their adjoint from this step's
    // E half step and laser boundary only
    for (int f = 0; f < 3; f++)
        if ((r = zero_adj(c, D, D.lJ[f], c->g.jsize * sizeof(T)))) return r;
    // fields_first: E then B
    if ((r = adj_wrap3<T>(c, D, lB)) || (r = adj_b<T>(c, D, h))) return r;
    if ((r = adj_wrap3<T>(c, D, lE)) || (r = adj_e<T>(c, D, h))) return r;
    // laser boundary (x open only)
    if (!c->p.periodic[0]) {
        LaserArgs L = laser_args(c, st, 1, c->p.laser_right);
        L.env_on = D.env.empty() ? 0 : 1;
        L.env_val = D.env.empty() ? 0.0 : D.env[st - D.first];
        dim3 grid(blocks_for(c->g.n[1], 256), c->g.n[2]);
        T* const* l = (T* const*)D.lF;
        k_adj_laser<T><<<grid, 256, 0, D.as>>>(l[1], l[2], l[3], l[4], l[5], (T*)D.lJ[1], (T*)D.lJ[2], c->g, L,
                                                    D.genv + (st - D.first));
        PIC_LAUNCH_CHECK(c);
    }
    return PIC_OK;
}

int adj_step(pic_ctx* c, DiffState& D, int st, const char* save) {
    return D.f32 ? adj_step_t<float>(c, D, st, save) : adj_step_t<double>(c, D, st, save);
}

void free_adjoint(DiffState& D) {
    for (auto& p : D.lF) { if (p) hipFree(p); p = nullptr; }
    for (auto& p : D.lJ) { if (p) hipFree(p); p = nullptr; }
    for (auto& p : D.lp) { if (p) hipFree(p); p = nullptr; }
    for (auto& p : D.lp_tmp) { if (p) hipFree(p); p = nullptr; }
    if (D.lw) hipFree(D.lw);
    D.lw = nullptr;
    if (D.lw_tmp) hipFree(D.lw_tmp);
    D.lw_tmp = nullptr;
    if (D.genv) hipFree(D.genv);
    D.genv = nullptr;
    if (D.stray) hipFree(D.stray);
    D.stray = nullptr;
    for (auto& b : D.segb) { if (b) hipFree(b); b = nullptr; }
    D.seg_steps = 0;
    for (int64_t** p : {&D.atoff, &D.awork, &D.anwork}) { if (*p) hipFree(*p); *p = nullptr; }
}

void free_checkpoints(DiffState& D, bool keep_initial) {
    for (size_t i = keep_initial ? 1 : 0; i < D.ck.size(); i++)
        if (D.ck[i]) hipFree(D.ck[i]);
    D.ck.resize(keep_initial && !D.ck.empty() ? 1 : 0);
    if (D.fin) hipFree(D.fin);
    D.fin = nullptr;
    if (D.perms) hipFree(D.perms);
    D.perms = nullptr;
    if (D.toffs) hipFree(D.toffs);
    D.toffs = nullptr;
    D.nsorts = 0;
}

}  // namespace

void diff_free(pic_ctx* c) {
    if (!c || !c->diff) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    DiffState& D = *c->diff;
    if (D.as) hipStreamSynchronize(D.as);
    free_adjoint(D);
    free_checkpoints(D, false);
    if (D.red) hipFree(D.red);
    for (int b = 0; b < 2; b++) {
        if (D.ev_rec[b]) hipEventDestroy(D.ev_rec[b]);
        if (D.ev_adj[b]) hipEventDestroy(D.ev_adj[b]);
    }
    if (D.ev_obj) hipEventDestroy(D.ev_obj);
    if (D.as) hipStreamDestroy(D.as);
    c->push_kernel = D.saved_push_kernel;
    c->sort_interval = D.saved_sort_interval;
    delete c->diff;
    c->diff = nullptr;
}

}  // namespace pic

using namespace pic;

extern "C" {

int pic_diff_begin(pic_ctx* c, int s) {
    if (c) gauss_invalidate(c);
    if (!c || s < 0 || s >= PIC_MAX_SPECIES) return PIC_ERR_ARG;
    Species& S = c->sp[s];
    if (!S.declared || S.immobile) return set_err(c, PIC_ERR_STATE, "species %d is not a mobile species", s);
    if (S.spin)
        return set_err(c, PIC_ERR_STATE, "differentiated run: species %d precesses its spin (not differentiated)", s);
    for (int o = 0; o < PIC_MAX_SPECIES; o++)
        if (o != s && c->sp[o].declared && !c->sp[o].immobile && c->sp[o].cap > 0)
            return set_err(c, PIC_ERR_STATE, "differentiated run: species %d is mobile too (one mobile species)", o);
    diff_free(c);
    PIC_HIP(c, hipSetDevice(c->device));
    // The tiled push keeps slots fixed as long as the species is not
    // re-sorted: sort once here (grouping by tile, dropping dead slots);
    // particles that later leave their tile take the push's HBM path.
    const int kernel = c->push_kernel == 1 ? 1 : 2;
    if (kernel == 2 && S.cap > 0) {
        int r0 = launch_sort(c, s);
        if (r0) return r0;
    }
    int64_t cnt[CNT_N];
    PIC_HIP(c, hipMemcpyAsync(cnt, S.d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    c->diff = new DiffState();
    DiffState& D = *c->diff;
    D.s = s;
    D.f32 = c->fp32;
    D.kind = tile_kind(c);
    D.slots = cnt[CNT_SLOTS];
    D.saved_push_kernel = c->push_kernel;
    D.saved_sort_interval = c->sort_interval;
    D.kernel = kernel;
    D.sort_every = std::max(1, c->sort_interval);
    D.ck_bytes = snapshot_bytes(c, D);
    void* p = nullptr;
    int r = dalloc(c, &p, D.ck_bytes);
    if (!r) {
        D.ck.push_back((char*)p);
        r = copy_state(c, D, D.ck[0], true);
    }
    if (!r) r = dalloc(c, (void**)&D.red, 256);
    if (!r && hipStreamCreateWithFlags(&D.as, hipStreamNonBlocking) != hipSuccess) {
        D.as = nullptr;
        r = set_err(c, PIC_ERR_HIP, "pic_diff_begin: adjoint stream");
    }
    for (int b = 0; b < 2 && !r; b++)
        if (hipEventCreateWithFlags(&D.ev_rec[b], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&D.ev_adj[b], hipEventDisableTiming) != hipSuccess)
            r = set_err(c, PIC_ERR_HIP, "pic_diff_begin: events");
    if (!r && hipEventCreateWithFlags(&D.ev_obj, hipEventDisableTiming) != hipSuccess)
        r = set_err(c, PIC_ERR_HIP, "pic_diff_begin: events");
    if (!r && hipStreamSynchronize(c->stream) != hipSuccess) r = set_err(c, PIC_ERR_HIP, "pic_diff_begin failed");
    if (r) diff_free(c);
    return r;
}

int pic_diff_slots(pic_ctx* c, int64_t* slots) {
    if (!c || !slots) return PIC_ERR_ARG;
    if (!c->diff) return set_err(c, PIC_ERR_STATE, "no differentiated run (pic_diff_begin)");
    *slots = c->diff->slots;
    return PIC_OK;
}

int pic_diff_initial(pic_ctx* c, double* x, double* y, double* z, double* w) {
    if (!c) return PIC_ERR_ARG;
    if (!c->diff) return set_err(c, PIC_ERR_STATE, "no differentiated run (pic_diff_begin)");
    DiffState& D = *c->diff;
    PIC_HIP(c, hipSetDevice(c->device));
    double* dst[4] = {x, y, z, w};
    if (!D.f32) {
        const int col[4] = {0, 1, 2, 6};
        for (int i = 0; i < 4; i++)
            if (dst[i] && D.slots)
                PIC_HIP(c, hipMemcpyAsync(dst[i], D.ck[0] + snap_col(c, D, col[i]), D.slots * 8, hipMemcpyDeviceToHost,
                                          c->stream));
        PIC_HIP(c, hipStreamSynchronize(c->stream));
        return PIC_OK;
    }
    // FP32: global position = cell + offset; a removed slot has a NaN weight
    std::vector<int32_t> cell(D.slots);
    std::vector<float> v[4];
    const int col[4] = {1, 2, 3, 7};
    if (D.slots) {
        PIC_HIP(c, hipMemcpyAsync(cell.data(), D.ck[0] + snap_col(c, D, 0), D.slots * 4, hipMemcpyDeviceToHost,
                                  c->stream));
        for (int i = 0; i < 4; i++) {
            v[i].resize(D.slots);
            PIC_HIP(c, hipMemcpyAsync(v[i].data(), D.ck[0] + snap_col(c, D, col[i]), D.slots * 4,
                                      hipMemcpyDeviceToHost, c->stream));
        }
    }
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    const int nx = c->g.n[0], ny = c->g.n[1];
    for (int64_t p = 0; p < D.slots; p++) {
        const int cl = cell[p];
        const int cc[3] = {cl % nx, (cl / nx) % ny, cl / (nx * ny)};
        for (int a = 0; a < 3; a++)
            if (dst[a]) dst[a][p] = cl < 0 ? 0.0 : cc[a] + (double)v[a][p];
        if (w) w[p] = cl < 0 ? __builtin_nan("") : (double)v[3][p];
    }
    return PIC_OK;
}

int pic_diff_forward(pic_ctx* c, int first_step, int nsteps, int K, const double* envelope,
                     const double* weights, double* objective) {
    if (c) gauss_invalidate(c);
    if (!c || nsteps < 0 || K < 1 || !objective) return PIC_ERR_ARG;
    if (!c->diff) return set_err(c, PIC_ERR_STATE, "no differentiated run (pic_diff_begin)");
    DiffState& D = *c->diff;
    PIC_HIP(c, hipSetDevice(c->device));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    free_adjoint(D);
    free_checkpoints(D, true);
    D.first = first_step;
    D.nsteps = nsteps;
    D.K = K;
    D.forward_done = 0;
    if (envelope) D.env.assign(envelope, envelope + nsteps);
    else D.env.clear();
    int r = copy_state(c, D, D.ck[0], false);
    if (!r && D.kernel == 2 && nsteps > 0) {  // the run's sort records
        if (D.slots > INT32_MAX)
            return set_err(c, PIC_ERR_ARG, "differentiated run: %lld slots exceed the int32 sort records",
                           (long long)D.slots);
        D.nsorts = (nsteps + D.sort_every - 1) / D.sort_every;
        r = dalloc(c, (void**)&D.perms, (size_t)D.nsorts * std::max<int64_t>(D.slots, 1) * 4);
        if (!r) r = dalloc(c, (void**)&D.toffs, (size_t)D.nsorts * (c->tl.ntiles + 2) * 8);
    }
    if (!r && weights && D.slots) {  // new initial weights: into the state and the snapshot
        // The FP32 deposit's fixed-point unit is sized from the largest
        // weight (k_push32.hip consts32): size it for these weights, so that
        // the forward, the recompute and the adjoint of this run all use it.
        double wmax = 0.0;
        for (int64_t q = 0; q < D.slots; q++)
            if (std::isfinite(weights[q])) wmax = std::max(wmax, std::fabs(weights[q]));
        c->sp[D.s].wmax = wmax;
        double* tmp = nullptr;
        r = dalloc(c, (void**)&tmp, D.slots * 8);
        if (!r && hipMemcpyAsync(tmp, weights, D.slots * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            r = set_err(c, PIC_ERR_HIP, "pic_diff_forward: weight upload failed");
        if (!r) {
            if (D.f32)
                k_adj_set_w32<<<blocks_for(D.slots, 256), 256, 0, c->stream>>>(
                    (const int32_t*)pcol(c, D.s, 0), (float*)pcol(c, D.s, D.wcol()), tmp, D.slots);
            else
                k_adj_set_w<<<blocks_for(D.slots, 256), 256, 0, c->stream>>>((double*)pcol(c, D.s, D.wcol()), tmp,
                                                                           D.slots);
            if (hipGetLastError() != hipSuccess) r = set_err(c, PIC_ERR_HIP, "pic_diff_forward: weight kernel failed");
        }
        if (!r) r = copy_state(c, D, D.ck[0], true);
        if (!r && hipStreamSynchronize(c->stream) != hipSuccess) r = set_err(c, PIC_ERR_HIP, "pic_diff_forward failed");
        if (tmp) hipFree(tmp);
    }
    // the push's own re-sort is off: diff_step re-sorts every D.sort_every
    // steps itself and records each permutation for the recompute
    c->push_kernel = D.kernel;
    c->sort_interval = 1 << 30;
    for (int st = first_step; st < first_step + nsteps && !r; st++) {
        const int i = st - first_step;
        if (i > 0 && i % K == 0) {
            void* p = nullptr;
            r = dalloc(c, &p, D.ck_bytes);
            if (!r) {
                D.ck.push_back((char*)p);
                r = copy_state(c, D, D.ck.back(), true);
            }
        }
        if (!r) r = diff_step(c, D, st, nullptr, false);
    }
    c->push_kernel = D.saved_push_kernel;
    c->sort_interval = D.saved_sort_interval;
    if (!r) r = dalloc(c, (void**)&D.fin, D.ck_bytes);
    if (!r) r = copy_state(c, D, D.fin, true);
    if (r) return r;
    // objective: sum m w (gamma - 1) dV
    Species& S = c->sp[D.s];
    if ((r = fill0(c, D.red, 8))) return r;
    if (D.slots) {
        const int blocks = (int)std::min<int64_t>(blocks_for(D.slots, 256), 2048);
        if (D.f32)
            k_adj_kinetic32<<<blocks, 256, 0, c->stream>>>(
                (const int32_t*)pcol(c, D.s, 0), (const float*)pcol(c, D.s, 7), (const float*)pcol(c, D.s, 4),
                (const float*)pcol(c, D.s, 5), (const float*)pcol(c, D.s, 6), D.slots, D.red);
        else
            k_adj_kinetic<<<blocks, 256, 0, c->stream>>>(
                (const double*)pcol(c, D.s, 6), (const double*)pcol(c, D.s, 3), (const double*)pcol(c, D.s, 4),
                (const double*)pcol(c, D.s, 5), D.slots, D.red);
        PIC_LAUNCH_CHECK(c);
    }
    double sum = 0.0;
    PIC_HIP(c, hipMemcpyAsync(&sum, D.red, 8, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    D.objective = sum * S.m * c->p.dx * c->p.dy * c->p.dz;
    *objective = D.objective;
    D.forward_done = 1;
    return PIC_OK;
}

int pic_diff_backward(pic_ctx* c, double grad_objective, double* grad_envelope, double* grad_weights) {
    if (c) gauss_invalidate(c);
    if (!c) return PIC_ERR_ARG;
    if (!c->diff || !c->diff->forward_done)
        return set_err(c, PIC_ERR_STATE, "pic_diff_backward needs a pic_diff_forward first");
    DiffState& D = *c->diff;
    PIC_HIP(c, hipSetDevice(c->device));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    const Grid& g = c->g;
    const size_t es = D.es();
    const int64_t ns = std::max<int64_t>(D.slots, 1);
    int r = copy_state(c, D, D.fin, false);  // the final state of the forward
    free_adjoint(D);
    for (int f = 0; f < 6 && !r; f++) r = dalloc(c, &D.lF[f], g.fsize * es);
    for (int f = 0; f < 3 && !r; f++) r = dalloc(c, &D.lJ[f], g.jsize * es);
    for (int f = 0; f < 6 && !r; f++) r = dalloc(c, &D.lp[f], ns * es);
    if (!r) r = dalloc(c, (void**)&D.lw, ns * 8);
    const bool sorts = D.kernel == 2;
    for (int f = 0; f < 6 && !r && sorts; f++) r = dalloc(c, &D.lp_tmp[f], ns * es);
    if (!r && sorts) r = dalloc(c, (void**)&D.lw_tmp, ns * 8);
    if (!r) r = dalloc(c, (void**)&D.genv, std::max(D.nsteps, 1) * 8);
    if (!r) r = dalloc(c, (void**)&D.stray, (ns + 2) * 8);  // slots + a 16-byte count
    D.step_bytes = step_state_bytes(c, D);
    D.seg_steps = std::min(D.K, std::max(D.nsteps, 1));
    const int nseg = (D.nsteps + D.K - 1) / D.K;
    for (int b = 0; b < std::min(nseg, 2) && !r; b++) r = dalloc(c, (void**)&D.segb[b], D.step_bytes * D.seg_steps);
    if (!r) r = dalloc(c, (void**)&D.atoff, (size_t)(c->tl.ntiles + 2) * 8);
    if (!r) r = dalloc(c, (void**)&D.awork, (size_t)std::max<int64_t>(c->work_cap, 1) * 8);
    if (!r) r = dalloc(c, (void**)&D.anwork, 8);
    for (int f = 0; f < 6 && !r; f++) r = fill0(c, D.lF[f], g.fsize * es);
    for (int f = 0; f < 3 && !r; f++) r = fill0(c, D.lJ[f], g.jsize * es);
    if (!r) r = fill0(c, D.genv, std::max(D.nsteps, 1) * 8);
    if (r) return r;
    // adjoint of the objective at the final state (the live state now)
    Species& S = c->sp[D.s];
    const double mdv = S.m * c->p.dx * c->p.dy * c->p.dz;
    if (D.f32)
        k_adj_objective32<<<blocks_for(ns, 256), 256, 0, c->stream>>>(adj_view32(c, D, nullptr), mdv, grad_objective);
    else
        k_adj_objective<<<blocks_for(ns, 256), 256, 0, c->stream>>>(adj_view(c, D, nullptr), mdv, grad_objective);
    PIC_LAUNCH_CHECK(c);
    PIC_HIP(c, hipEventRecord(D.ev_obj, c->stream));
    PIC_HIP(c, hipStreamWaitEvent(D.as, D.ev_obj, 0));
    c->push_kernel = D.kernel;
    c->sort_interval = 1 << 30;
    // Segment k is recomputed on the context's stream into buffer k % 2 and
    // reversed on the adjoint stream; segment k - 1 is recomputed while
    // segment k is reversed (its buffer is free once segment k + 1 has been
    // reversed).
    auto seg_range = [&](int sg, int& a, int& b) {
        a = D.first + sg * D.K;
        b = std::min(a + D.K, D.first + D.nsteps);
    };
    auto recompute = [&](int sg) -> int {
        int a, b;
        seg_range(sg, a, b);
        char* buf = D.segb[sg % 2];
        if (sg + 2 <= nseg - 1) PIC_HIP(c, hipStreamWaitEvent(c->stream, D.ev_adj[sg % 2], 0));
        int rr = copy_state(c, D, D.ck[sg], false);
        for (int st = a; st < b && !rr; st++) rr = diff_step(c, D, st, buf + (size_t)(st - a) * D.step_bytes, true);
        if (!rr) PIC_HIP(c, hipEventRecord(D.ev_rec[sg % 2], c->stream));
        return rr;
    };
    auto reverse = [&](int sg) -> int {
        int a, b;
        seg_range(sg, a, b);
        char* buf = D.segb[sg % 2];
        PIC_HIP(c, hipStreamWaitEvent(D.as, D.ev_rec[sg % 2], 0));
        int rr = PIC_OK;
        for (int st = b - 1; st >= a && !rr; st--) rr = adj_step(c, D, st, buf + (size_t)(st - a) * D.step_bytes);
        if (!rr) PIC_HIP(c, hipEventRecord(D.ev_adj[sg % 2], D.as));
        return rr;
    };
    if (nseg > 0) r = recompute(nseg - 1);
    for (int sg = nseg - 1; sg >= 0 && !r; sg--) {
        r = reverse(sg);
        if (!r && sg > 0) r = recompute(sg - 1);
    }
    const hipError_t e_as = hipStreamSynchronize(D.as);
    c->push_kernel = D.saved_push_kernel;
    c->sort_interval = D.saved_sort_interval;
    if (!r && e_as != hipSuccess) r = set_err(c, PIC_ERR_HIP, "adjoint stream: %s", hipGetErrorString(e_as));
    // leave the context at the run's initial state
    if (!r) r = copy_state(c, D, D.ck[0], false);
    if (r) {
        hipStreamSynchronize(c->stream);
        return r;
    }
    if (grad_envelope && D.nsteps)
        PIC_HIP(c, hipMemcpyAsync(grad_envelope, D.genv, D.nsteps * 8, hipMemcpyDeviceToHost, c->stream));
    if (grad_weights && D.slots)
        PIC_HIP(c, hipMemcpyAsync(grad_weights, D.lw, D.slots * 8, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_diff_end(pic_ctx* c) {
    if (!c) return PIC_ERR_ARG;
    diff_free(c);
    return PIC_OK;
}

}  // extern "C"
