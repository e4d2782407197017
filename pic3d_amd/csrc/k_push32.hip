// k_push32.hip -- the FP32 form of the particle hot path.
//
// Particles: linear cell index (int32, -1 = removed) + in-cell offsets
// fx, fy, fz in [0,1) (fp32) + momenta u = p/mc (fp32) + weight (fp32) + id.
// Fields, currents and charge are fp32.  Positions relative to the cell keep
// the shape-function arguments small, so fp32 keeps ~1e-7 of a cell of
// resolution anywhere in the domain.
//
// Same algorithm as the FP64 path (and the reference, particle.cpp:113-515):
// periodic wrap, staggered gather with quadratic/quartic B-splines,
// relativistic Boris, Esirkepov deposit of J and rho.  Differences, all
// exact in exact arithmetic: the gather is evaluated separably (x, then y,
// then z sums) with fused multiply-adds; the last point of each cumulative
// current sum (which is sum(S1) - sum(S0) = 0) is not deposited; periodic
// positions are wrapped right after the move instead of at the next push,
// and boundary crossings are classified by the push itself (the deposit
// still uses the unwrapped position, so the ghost fold carries the current).
#include <cmath>
#include <type_traits>
#include <vector>

#include "pic_internal.h"
#include "atomics.h"
#include "shapes.h"
#include "wave.h"

#define PIC_LD_FN __host__ __device__ static inline
#include "pic_loader_impl.h"

namespace pic {

struct Soa32 {
    int32_t* cell;
    float *fx, *fy, *fz, *ux, *uy, *uz, *w;
    int64_t* id;
    float *sx, *sy, *sz;  // spin (species with precession only)
};

static Soa32 soa32(Species& S, int b) {
    return Soa32{S.cell[b], S.fx[b], S.fy[b], S.fz[b], S.vx[b], S.vy[b], S.vz[b], S.fw[b], S.id[b],
                 (float*)S.spin_v[0][b], (float*)S.spin_v[1][b], (float*)S.spin_v[2][b]};
}

struct Consts32 {
    // LDS accumulation is int32 fixed point: component c (Jx, Jy, Jz, rho)
    // in units ufix[c] = (bound on one particle's contribution) * 2^-18, so
    // one chunk (<= 4096 = 2^12 particles, one contribution each per node)
    // stays below 2^30 and cannot overflow; rounding error per deposit is
    // <= 2^-19 of that bound.  The quadratic push's chunks of 2^13 use
    // 2^-17 (rounding <= 2^-18 of the bound).  (Float atomic adds to LDS are
    // an order of magnitude slower than integer ones on this part.)
    float ufix[4], inv_ufix[4];
    float qm;          // q/m * dt/2
    float q;
    float cfx, cfy, cfz;  // dt/dx, dt/dy, dt/dz
    float cjx, cjy, cjz;  // dx/dt, dy/dt, dz/dt (current scale per q w)
    int n[3];
    int per[3];
    int m;
    CellSplit cs;  // cell index -> (cx, cy, cz)
    SpinConsts spin;
};

// quadratic/quartic weights for an offset delta = x - round(x) in [-0.5, 0.5)
template <int M> __device__ __forceinline__ void shape32(float d, float* W) {
    if (M == 2) shape2_delta<float>(d, W);
    else shape4_d<float>(d + 2.0f, W);
}

template <int L>
__device__ __forceinline__ void shift_down32(float* A) {
#pragma unroll
    for (int t = 0; t < L - 1; t++) A[t] = A[t + 1];
}

// ---------------------------------------------------------------------
// gather (fp32): E and B at the particle.  FR(comp, k, j, i) reads a field
// value at buffer indices.  Separable sums (x, then y, then z) with FMAs;
// for quartic shapes the z loop is rolled (register rotation of the z
// weights) to keep the kernel inside the instruction cache.
// ---------------------------------------------------------------------
// The primal weights at the particle (Wxp, Wyp, Wzp) are returned: the
// deposit's old-position weights are the same values.
template <int M, typename FR>
__device__ __forceinline__ void gather32(int Cx, int Cy, int Cz, float fx, float fy, float fz,
                                         float* E, float* B, FR fr, float* Wxp, float* Wyp, float* Wzp) {
    constexpr int H = M / 2;
    float Wxd[M + 1], Wyd[M + 1], Wzd[M + 1];
    const int ox = fx >= 0.5f, oy = fy >= 0.5f, oz = fz >= 0.5f;
    shape32<M>(fx - (float)ox, Wxp);
    shape32<M>(fy - (float)oy, Wyp);
    shape32<M>(fz - (float)oz, Wzp);
    shape32<M>(fx - 0.5f, Wxd);
    shape32<M>(fy - 0.5f, Wyd);
    shape32<M>(fz - 0.5f, Wzd);
    const int ipx = Cx + ox - H, ipy = Cy + oy - H, ipz = Cz + oz - H;  // primal box start
    const int idx = Cx - H, idy = Cy - H, idz = Cz - H;                 // dual box start
    // stagger per component (x, y, z): Ex (d,p,p) Ey (p,d,p) Ez (p,p,d)
    // Bx (p,d,d) By (d,p,d) Bz (d,d,p)
    auto one = [&](int c, const float* WX, const float* WY, const float* WZ, int bx, int by,
                   int bz) -> float {
        float acc = 0.0f;
        float wz[M + 1];
#pragma unroll
        for (int t = 0; t <= M; t++) wz[t] = WZ[t];
#pragma unroll(M == 2 ? M + 1 : 1)
        for (int k = 0; k <= M; k++) {
            float row = 0.0f;
#pragma unroll
            for (int j = 0; j <= M; j++) {
                float s = 0.0f;
#pragma unroll
                for (int i = 0; i <= M; i++) s = fmaf(WX[i], fr(c, bz + k, by + j, bx + i), s);
                row = fmaf(WY[j], s, row);
            }
            acc = fmaf(wz[0], row, acc);
            shift_down32<M + 1>(wz);
        }
        return acc;
    };
    E[0] = one(0, Wxd, Wyp, Wzp, idx, ipy, ipz);
    E[1] = one(1, Wxp, Wyd, Wzp, ipx, idy, ipz);
    E[2] = one(2, Wxp, Wyp, Wzd, ipx, ipy, idz);
    B[0] = one(3, Wxp, Wyd, Wzd, ipx, idy, idz);
    B[1] = one(4, Wxd, Wyp, Wzd, idx, ipy, idz);
    B[2] = one(5, Wxd, Wyd, Wzp, idx, idy, ipz);
}

__device__ __forceinline__ void boris32(float* u, const float* E, const float* B, float qm) {
    const float e0 = E[0] * qm, e1 = E[1] * qm, e2 = E[2] * qm;
    const float b0 = B[0] * qm, b1 = B[1] * qm, b2 = B[2] * qm;
    const float px1 = u[0] + e0, py1 = u[1] + e1, pz1 = u[2] + e2;
    const float ig1 = rsqrtf(1.0f + (px1 * px1 + py1 * py1 + pz1 * pz1));
    const float tx = b0 * ig1, ty = b1 * ig1, tz = b2 * ig1;
    const float tt = 2.0f / (1.0f + tx * tx + ty * ty + tz * tz);
    const float sx = tx * tt, sy = ty * tt, sz = tz * tt;
    const float px2 = px1 + (py1 * tz - pz1 * ty);
    const float py2 = py1 + (pz1 * tx - px1 * tz);
    const float pz2 = pz1 + (px1 * ty - py1 * tx);
    u[0] = px1 + (py2 * sz - pz2 * sy) + e0;
    u[1] = py1 + (pz2 * sx - px2 * sz) + e1;
    u[2] = pz1 + (px2 * sy - py2 * sx) + e2;
}

// Esirkepov deposit (fp32) in the frame of the particle's old cell: old
// offsets f0 in [0,1), new offsets f1 = f0 + d with |d| < 1.  The sink D
// receives D.base(bz, by, bx) (box origin relative to the cell) and then
// D.put(comp, k, j, i, a, b) with box-relative (k, j, i): the contribution
// a * b, with a carrying D.scale(comp).  The box is always
// L = M+2 points per axis.  The last point of the particle's box along the
// cumulative axis (box point L-1 if the particle changes node along that
// axis, else L-2) carries sum(S1) - sum(S0), zero in exact arithmetic, and
// is not deposited; the extra planes along the other axes carry exact zeros.
// S0x, S0y, S0z: the weights at f0 (as the gather evaluates them)
// STILL: the caller knows that the particle keeps its nearest node along
// every axis (floor(f1 + 0.5) == (f0 >= 0.5)); the box's last line along
// each axis then carries only zeros and is left out at compile time (the
// same contributions as the general form).
template <int M, typename D, bool STILL = false>
__device__ __forceinline__ void esirkepov32(float qw, const float* f0, const float* f1, const float* S0x,
                                            const float* S0y, const float* S0z, const Consts32& K, D& dst) {
    constexpr int H = M / 2, L = M + 2;
    constexpr int LR = STILL ? L - 1 : L;  // box lines that can carry weight
    constexpr int R = (M == 2) ? L : 1;  // unroll factor of the outer loops
    const float* S0[3] = {S0x, S0y, S0z};
    float S1[3][M + 1];
    int b[3], d0[3], d1[3];
    bool moved[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int i1 = f0[a] >= 0.5f ? 1 : 0;
        const int i2 = STILL ? i1 : (int)floorf(f1[a] + 0.5f);  // -1 .. 2
        shape32<M>(f1[a] - (float)i2, S1[a]);
        b[a] = min(i1, i2) - H;
        moved[a] = STILL ? false : i1 != i2;
        d0[a] = i1 - H - b[a];
        d1[a] = i2 - H - b[a];
    }
    float X0[L], X1[L], Y0[L], Y1[L], Z0[L], Z1[L];
    auto place = [](const float* S, int d, float* A) {
#pragma unroll
        for (int t = 0; t < L; t++) {
            const float a0 = (t <= M) ? S[t] : 0.0f;
            const float a1 = (t >= 1) ? S[t - 1] : 0.0f;
            A[t] = d ? a1 : a0;
        }
    };
    place(S0[0], d0[0], X0); place(S1[0], d1[0], X1);
    place(S0[1], d0[1], Y0); place(S1[1], d1[1], Y1);
    place(S0[2], d0[2], Z0); place(S1[2], d1[2], Z1);
    // cumulative sums along each axis times the current scale (with the
    // sink's scale folded in): a contribution is CX[i] * W_yz
    float CX[L - 1], CY[L - 1], CZ[L - 1];
    {
        const float sx = -qw * K.cjx * dst.scale(0), sy = -qw * K.cjy * dst.scale(1),
                    sz = -qw * K.cjz * dst.scale(2);
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma unroll
        for (int t = 0; t < L - 1; t++) {
            ax += X1[t] - X0[t];
            ay += Y1[t] - Y0[t];
            az += Z1[t] - Z0[t];
            CX[t] = sx * ax;
            CY[t] = sy * ay;
            CZ[t] = sz * az;
        }
    }
    // transverse row weights s0 t0 + ds t0 / 2 + s0 dt / 2 + ds dt / 3
    // regrouped as s0 * (t0 + dt / 2) + ds * (t0 / 2 + dt / 3), the two
    // factors precomputed per axis point
    const float third = 1.0f / 3.0f;
    float PX[L], QX[L], PY[L], QY[L], DY[L], DZ[L];
#pragma unroll
    for (int t = 0; t < L; t++) {
        const float dx_ = X1[t] - X0[t], dy_ = Y1[t] - Y0[t], dz_ = Z1[t] - Z0[t];
        PX[t] = X0[t] + 0.5f * dx_; QX[t] = 0.5f * X0[t] + third * dx_;
        PY[t] = Y0[t] + 0.5f * dy_; QY[t] = 0.5f * Y0[t] + third * dy_;
        DY[t] = dy_; DZ[t] = dz_;
    }
    dst.base(b[2], b[1], b[0]);
    float za[L], zb[L];
#pragma unroll
    for (int t = 0; t < L; t++) { za[t] = Z0[t]; zb[t] = DZ[t]; }
    // Jx: cumulative along x, weights W_yz
#pragma unroll(R)
    for (int k = 0; k < LR; k++) {
        const float sz0 = za[0], dsz = zb[0];
#pragma unroll
        for (int j = 0; j < LR; j++) {
            const float wyz = __builtin_fmaf(dsz, QY[j], sz0 * PY[j]);
            if (wyz == 0.0f) continue;  // a plane outside both supports: only zeros to add
#pragma unroll
            for (int i = 0; i < L - 1; i++)
                if (i < L - 2 || moved[0]) dst.put(0, k, j, i, CX[i], wyz);
        }
        shift_down32<L>(za); shift_down32<L>(zb);
    }
    // Jy: cumulative along y, weights W_xz
#pragma unroll
    for (int t = 0; t < L; t++) { za[t] = Z0[t]; zb[t] = DZ[t]; }
#pragma unroll(R)
    for (int k = 0; k < LR; k++) {
        const float sz0 = za[0], dsz = zb[0];
#pragma unroll
        for (int i = 0; i < LR; i++) {
            const float wxz = __builtin_fmaf(dsz, QX[i], sz0 * PX[i]);
            if (wxz == 0.0f) continue;
#pragma unroll
            for (int j = 0; j < L - 1; j++)
                if (j < L - 2 || moved[1]) dst.put(1, k, j, i, CY[j], wxz);
        }
        shift_down32<L>(za); shift_down32<L>(zb);
    }
    // Jz: cumulative along z, weights W_xy
    float ya[L], yb[L];
#pragma unroll
    for (int t = 0; t < L; t++) { ya[t] = Y0[t]; yb[t] = DY[t]; }
#pragma unroll(R)
    for (int j = 0; j < LR; j++) {
        const float sy0 = ya[0], dsy = yb[0];
#pragma unroll
        for (int i = 0; i < LR; i++) {
            const float wxy = __builtin_fmaf(dsy, QX[i], sy0 * PX[i]);
            if (wxy == 0.0f) continue;
#pragma unroll
            for (int k = 0; k < L - 1; k++)
                if (k < L - 2 || moved[2]) dst.put(2, k, j, i, CZ[k], wxy);
        }
        shift_down32<L>(ya); shift_down32<L>(yb);
    }
    // rho at the new position: nodes i2-H .. i2+H
    dst.base(b[2] + d1[2], b[1] + d1[1], b[0] + d1[0]);
    float zr[M + 1];
    const float qs = qw * dst.scale(3);
#pragma unroll
    for (int t = 0; t <= M; t++) zr[t] = qs * S1[2][t];
#pragma unroll(M == 2 ? M + 1 : 1)
    for (int k = 0; k <= M; k++) {
#pragma unroll
        for (int j = 0; j <= M; j++) {
            const float syz = zr[0] * S1[1][j];
#pragma unroll
            for (int i = 0; i <= M; i++) dst.put(3, k, j, i, syz, S1[0][i]);
        }
        shift_down32<M + 1>(zr);
    }
}

// deposit sinks (fp32) ---------------------------------------------------
// Sinks take each contribution as a product a * b, a carrying the sink's
// scale(c).
struct HbmSink32 {  // global atomics; indices relative to the cell (Cz, Cy, Cx)
    float* jp[4];
    Grid g;
    int Cz, Cy, Cx;
    int bz, by, bx;
    __device__ __forceinline__ float scale(int) const { return 1.0f; }
    __device__ __forceinline__ void base(int z, int y, int x) { bz = Cz + z; by = Cy + y; bx = Cx + x; }
    __device__ __forceinline__ void put(int c, int k, int j, int i, float a, float b) {
        hbm_add(&jp[c][g.ji(bz + k, by + j, bx + i)], a * b);
    }
};

// LDS accumulator box (ACC = double: FP64 sums; ACC = int: fixed point in
// units ufix[c]); 'cell' = the particle's cell offset in the box.  Fixed
// point: a carries 1/ufix[c], |a * b| <= 2^18 units, and one fused
// a * b + 1.5 * 2^23 rounds the product to the nearest integer in the low
// mantissa bits; subtracting that constant's bit pattern leaves the
// two's-complement integer (valid below 2^22 units).
template <int AX, int AY, int AVOL, typename ACC>
struct LdsSink32 {
    ACC* acc;
    const float* inv_u;
    int cell;
    int off;
    __device__ __forceinline__ float scale(int c) const {
        if constexpr (sizeof(ACC) == 8) return 1.0f;
        else return inv_u[c];
    }
    __device__ __forceinline__ void base(int z, int y, int x) { off = cell + (z * AY + y) * AX + x; }
    __device__ __forceinline__ void put(int c, int k, int j, int i, float a, float b) {
        if constexpr (sizeof(ACC) == 8) {
            lds_add((double*)&acc[off + c * AVOL + (k * AY + j) * AX + i], (double)(a * b));
        } else {
            const float t = __builtin_fmaf(a, b, 0x1.8p23f);
            lds_add_i32((int*)&acc[off + c * AVOL + (k * AY + j) * AX + i], __float_as_int(t) - 0x4B400000);
        }
    }
};

// Move: new offsets, cell shifts, periodic wrap / open-boundary removal.
// Returns the new linear cell or -1 (removed); counts crossings in cnt12
// (lost x-,x+,y-,y+,z-,z+ then wrapped x-,...).
__device__ __forceinline__ int move32(int cx, int cy, int cz, float* f, const float* d,
                                      const Consts32& K, unsigned* cnt12) {
    int c[3] = {cx, cy, cz};
    bool lost = false;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float nf = f[a] + d[a];
        const float s = floorf(nf);
        nf -= s;
        int sh = (int)s;
        if (nf >= 1.0f) { nf = 0.0f; sh += 1; }  // rounding of a tiny negative offset
        f[a] = nf;
        int cc = c[a] + sh;
        // phases X, Y, Z in order: once removed at an open face, later axes
        // are not classified (mpi_utils.cpp:62-230)
        if (!lost && (cc < 0 || cc >= K.n[a])) {
            const int face = 2 * a + (cc < 0 ? 0 : 1);
            if (K.per[a]) {
                cc += cc < 0 ? K.n[a] : -K.n[a];
                atomicAdd(&cnt12[6 + face], 1u);
            } else {
                lost = true;
                atomicAdd(&cnt12[face], 1u);
            }
        }
        c[a] = cc;
    }
    if (lost) return -1;
    return (c[2] * K.n[1] + c[1]) * K.n[0] + c[0];
}

// ---------------------------------------------------------------------
// one particle: gather (through FR), Boris, deposit (into SINK), move.
// Returns false when the particle was skipped as out of bounds.
// ---------------------------------------------------------------------
//
// SPIN: the spin sp[3] precesses after the push (particle.cpp:318-368, in
// FP32); spin_bad counts steps without a g-factor value.
template <int M, bool SPIN, typename FR, typename SINK>
__device__ __forceinline__ bool push32_core(int cx, int cy, int cz, int m, float* f, float* u,
                                            float w, const Consts32& K, FR fr, SINK& sink,
                                            int& newcell, unsigned* cnt12, float* sp, int& spin_bad) {
    const int Cx = cx + m - 1, Cy = cy + m - 1, Cz = cz + m - 1;  // buffer index of the cell
    float E[3], Bf[3], W0[3][M + 1];
    gather32<M>(Cx, Cy, Cz, f[0], f[1], f[2], E, Bf, fr, W0[0], W0[1], W0[2]);
    boris32(u, E, Bf, K.qm);
    if constexpr (SPIN) {
        const float gam = sqrtf(1.0f + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
        if (!spin_precess<float>(E, Bf, K.qm, u[0], u[1], u[2], gam, K.spin, sp[0], sp[1], sp[2])) spin_bad++;
    }
    const float ig = rsqrtf(1.0f + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    float d[3] = {u[0] * ig * K.cfx, u[1] * ig * K.cfy, u[2] * ig * K.cfz};
    // a non-finite momentum would put the deposit box outside the arrays
    if (!(fabsf(d[0]) < 1.0f && fabsf(d[1]) < 1.0f && fabsf(d[2]) < 1.0f)) return false;
    float f1[3] = {f[0] + d[0], f[1] + d[1], f[2] + d[2]};
    esirkepov32<M>(K.q * w, f, f1, W0[0], W0[1], W0[2], K, sink);
    newcell = move32(cx, cy, cz, f, d, K, cnt12);
    return true;
}

// ---------------------------------------------------------------------
// tiled push: one workgroup per (tile, chunk) work item; field tile and
// current/charge accumulators in LDS
// ---------------------------------------------------------------------
template <int M, int TX, int TY, int TZ, typename ACC>
struct Box32 {
    static constexpr int H = M / 2;
    // field tile: buffer indices [T0 + m - 1 - H, T0 + T + m - 1 + H]
    static constexpr int FX = TX + M + 1, FY = TY + M + 1, FZ = TZ + M + 1;
    // At quartic order each component keeps its own box: along an axis
    // where it sits on the dual grid its stencil starts at the cell (not the
    // nearest node), so it needs one point fewer there.  Bit a of dual(c):
    // component c (Ex, Ey, Ez, Bx, By, Bz) is dual along axis a.  The saving
    // (3.6 KB at the 8x4x4 tile) let three workgroups share a CU: push
    // 10.0 -> 9.4 ms at 41.9M electrons.
    //
    // The fixed-point kernels (CONG) lay the field boxes out with row and
    // plane pitches congruent mod 32 (4-byte banks) to the accumulator
    // box's: lanes are dealt so that deposit bases fall on distinct banks
    // (bank_spread_lane), and with congruent pitches the gather's primal
    // bases then do too.  The quadratic one keeps one box shape for all six
    // components (one base register: trimmed boxes took it from 127 to 166
    // VGPRs, half the waves); at 8x8x8 this still fits two workgroups per
    // CU, and the quartic one at 4x4x4 three.
    static constexpr int AX = TX + M + 3, AY = TY + M + 3, AZ = TZ + M + 3;
    static constexpr bool CONG = sizeof(ACC) == 4;
    static constexpr int dual(int c) {
        return M != 4 ? 0 : c == 0 ? 1 : c == 1 ? 2 : c == 2 ? 4 : c == 3 ? 6 : c == 4 ? 5 : 3;
    }
    static constexpr int cdim(int c, int a) { return (a == 0 ? FX : a == 1 ? FY : FZ) - ((dual(c) >> a) & 1); }
    static constexpr int pitch32(int at_least, int residue) {
        return at_least + ((residue - at_least) % 32 + 32) % 32;
    }
    static constexpr int rowp(int c) { return CONG ? pitch32(FX, AX) : cdim(c, 0); }
    static constexpr int planep(int c) {
        return CONG ? pitch32(rowp(c) * cdim(c, 1), AX * AY) : cdim(c, 0) * cdim(c, 1);
    }
    static constexpr int csize(int c) {
        return (cdim(c, 2) - 1) * planep(c) + (cdim(c, 1) - 1) * rowp(c) + cdim(c, 0);
    }
    static constexpr int coff(int c) { return c == 0 ? 0 : coff(c - 1) + csize(c - 1); }
    static constexpr int FTOT = coff(6);
    // accumulator box: [T0 + m - 2 - H, T0 + T + m + H]
    static constexpr int AVOL = AX * AY * AZ;
    // accumulators first (8-byte aligned for ACC = double), then fields
    static constexpr size_t LDS = (size_t)4 * AVOL * sizeof(ACC) + (size_t)FTOT * sizeof(float);
};

struct Fields32 {
    const float *ex, *ey, *ez, *bx, *by, *bz;
    float *jx, *jy, *jz, *rho;
};

// Threads per workgroup of the tiled push.  The LDS boxes allow two
// workgroups per CU, so 256 threads leave each SIMD two waves.  The
// quadratic int32-accumulator kernel without spin takes 512 threads at the
// 128 registers it uses anyway (four waves per SIMD): 29 % faster at the
// 1.07e9-electron ramp target.  The quartic one measured 21 % slower at 384
// threads (three waves per SIMD at its 158 registers) and keeps 256.
template <int M, typename ACC, bool SPIN> struct Push32Threads {
    static constexpr int N = (M == 2 && !SPIN && std::is_same<ACC, int>::value) ? 512 : 256;
};

template <int M, int TX, int TY, int TZ, typename ACC, bool SPIN>
__global__ __launch_bounds__((Push32Threads<M, ACC, SPIN>::N), 2) void k_push32_tiled(Soa32 P, Fields32 F, Grid g, Consts32 K,
                                                      Tiling tl, const int64_t* __restrict__ tile_off,
                                                      const int64_t* __restrict__ work,
                                                      const int64_t* __restrict__ nwork, int chunk,
                                                      int64_t* __restrict__ cnt) {
    using B = Box32<M, TX, TY, TZ, ACC>;
    extern __shared__ double lds_d[];
    ACC* acc = reinterpret_cast<ACC*>(lds_d);             // 4 * AVOL accumulators
    float* fld = reinterpret_cast<float*>(acc + 4 * B::AVOL);  // the six field boxes
    __shared__ unsigned cnt12[12];
    __shared__ unsigned rg_hist[Push32Threads<M, ACC, SPIN>::N / 64][8];  // per wave: bank-spreading counters
    const int64_t wi = blockIdx.x;
    if (wi >= *nwork) return;
    const int64_t item = work[wi];
    const int tile = (int)(item >> 20);
    const int64_t cidx = item & ((1 << 20) - 1);
    const int64_t t_end = tile_off[tile + 1];
    const int64_t beg = tile_off[tile] + cidx * chunk;
    const int64_t end = min(t_end, beg + chunk);
    const int txi = tile % tl.nt[0], tyi = (tile / tl.nt[0]) % tl.nt[1], tzi = tile / (tl.nt[0] * tl.nt[1]);
    const int T0x = txi * TX, T0y = tyi * TY, T0z = tzi * TZ;
    const int m = g.m;
    // field tile origin (buffer indices); reads past a partial last tile are clamped
    const int fox = T0x + m - 1 - B::H, foy = T0y + m - 1 - B::H, foz = T0z + m - 1 - B::H;
    const float* fsrc[6] = {F.ex, F.ey, F.ez, F.bx, F.by, F.bz};
    auto load = [&](int c, const float* src) {
        const int DX = B::cdim(c, 0), DY = B::cdim(c, 1);
        float* dst = fld + B::coff(c);
        for (int r = threadIdx.x; r < DX * DY * B::cdim(c, 2); r += blockDim.x) {
            const int lx = r % DX, ly = (r / DX) % DY, lz = r / (DX * DY);
            const int gx = min(fox + lx, g.fe[0] - 1), gy = min(foy + ly, g.fe[1] - 1), gz = min(foz + lz, g.fe[2] - 1);
            dst[lz * B::planep(c) + ly * B::rowp(c) + lx] = src[g.fi(gz, gy, gx)];
        }
    };
    load(0, F.ex); load(1, F.ey); load(2, F.ez); load(3, F.bx); load(4, F.by); load(5, F.bz);
    for (int e = threadIdx.x; e < 4 * B::AVOL; e += blockDim.x) acc[e] = ACC(0);
    if (threadIdx.x < 12) cnt12[threadIdx.x] = 0;
    const int wv = threadIdx.x >> 6, lane = lane_id();
    if (lane < 8) rg_hist[wv][lane] = 0u;
    __syncthreads();
    const int aox = T0x + m - 2 - B::H, aoy = T0y + m - 2 - B::H, aoz = T0z + m - 2 - B::H;
    int64_t oob = 0;
    int spin_bad = 0;
    uint32_t strays = 0;  // iterations whose particle was outside the tile
    int it = 0;
    const int len = (int)(end - beg);
    // Each wave takes 64 consecutive slots (coalesced loads), then deals its
    // particles out again over its lanes so that lanes whose deposit boxes
    // start in the same LDS bank are served in different groups of 16
    // (bank_spread_lane); the results go back to the particles' own slots.
    for (int t0 = threadIdx.x & ~63; t0 < len; t0 += blockDim.x, it++) {
        const int t = t0 + lane;
        const int cell = t < len ? P.cell[beg + t] : -1;
        int cx = 0, cy = 0, cz = 0;
        bool ok = cell >= 0;
        if (ok) {
            K.cs.split(cell, cx, cy, cz);
            if ((unsigned)(cx - T0x) >= (unsigned)TX || (unsigned)(cy - T0y) >= (unsigned)TY ||
                (unsigned)(cz - T0z) >= (unsigned)TZ) {
                strays |= 1u << it;  // moved out of the tile since the last sort
                ok = false;
            }
        }
        float f[3] = {0.0f, 0.0f, 0.0f}, u[3] = {0.0f, 0.0f, 0.0f}, w = 0.0f, sp[3] = {0.0f, 0.0f, 0.0f};
        int r = 32;
        if (ok) {
            const int64_t p = beg + t;
            f[0] = P.fx[p]; f[1] = P.fy[p]; f[2] = P.fz[p];
            u[0] = P.ux[p]; u[1] = P.uy[p]; u[2] = P.uz[p];
            w = P.w[p];
            if (SPIN) { sp[0] = P.sx[p]; sp[1] = P.sy[p]; sp[2] = P.sz[p]; }
            // bank of the deposit base: the cell's box offset moved to the
            // nearest node of the old position, in 4-byte words
            const int ofs = ((cz + m - 1 - aoz) * B::AY + (cy + m - 1 - aoy)) * B::AX + (cx + m - 1 - aox) +
                            (f[0] >= 0.5f) + (f[1] >= 0.5f) * B::AX + (f[2] >= 0.5f) * (B::AX * B::AY);
            r = (ofs * (int)(sizeof(ACC) / 4)) & 31;
        }
        const int to = bank_spread_lane(r, rg_hist[wv]);
        const int okv = lane_send(to, ok ? 1 : 0);
        const int tt = lane_send(to, t);
        cx = lane_send(to, cx); cy = lane_send(to, cy); cz = lane_send(to, cz);
#pragma unroll
        for (int a = 0; a < 3; a++) { f[a] = lane_send(to, f[a]); u[a] = lane_send(to, u[a]); }
        w = lane_send(to, w);
        if (SPIN) {
#pragma unroll
            for (int a = 0; a < 3; a++) sp[a] = lane_send(to, sp[a]);
        }
        if (!okv) continue;
        const int64_t p = beg + tt;
        // field reads relative to the LDS tile; deposits relative to the
        // accumulator box (an in-tile particle's boxes lie inside both)
        auto fr = [&](int c, int k, int j, int i) -> float {
            const float* fb = fld + B::coff(c) + (cz + m - 1 - foz) * B::planep(c) + (cy + m - 1 - foy) * B::rowp(c) +
                              (cx + m - 1 - fox);
            return fb[(k - (cz + m - 1)) * B::planep(c) + (j - (cy + m - 1)) * B::rowp(c) + (i - (cx + m - 1))];
        };
        LdsSink32<B::AX, B::AY, B::AVOL, ACC> sink{acc, K.inv_ufix,
                                              ((cz + m - 1 - aoz) * B::AY + (cy + m - 1 - aoy)) * B::AX + (cx + m - 1 - aox), 0};
        int nc;
        if (!push32_core<M, SPIN>(cx, cy, cz, m, f, u, w, K, fr, sink, nc, cnt12, sp, spin_bad)) {
            oob++;
            continue;
        }
        P.cell[p] = nc;
        P.fx[p] = f[0]; P.fy[p] = f[1]; P.fz[p] = f[2];
        P.ux[p] = u[0]; P.uy[p] = u[1]; P.uz[p] = u[2];
        if (SPIN) { P.sx[p] = sp[0]; P.sy[p] = sp[1]; P.sz[p] = sp[2]; }
    }
    if (strays) {  // HBM path for the particles outside the tile
        float* jdst[4] = {F.jx, F.jy, F.jz, F.rho};
        it = 0;
        for (int t = threadIdx.x; t < len; t += blockDim.x, it++) {
            if (!(strays & (1u << it))) continue;
            const int64_t p = beg + t;
            const int cell = P.cell[p];
            int cx, cy, cz;
            K.cs.split(cell, cx, cy, cz);
            float f[3] = {P.fx[p], P.fy[p], P.fz[p]};
            float u[3] = {P.ux[p], P.uy[p], P.uz[p]};
            auto fr = [&](int c, int k, int j, int i) -> float { return fsrc[c][g.fi(k, j, i)]; };
            HbmSink32 sink{{jdst[0], jdst[1], jdst[2], jdst[3]}, g, cz + m - 1, cy + m - 1, cx + m - 1, 0, 0, 0};
            int nc;
            float sp[3];
            if (SPIN) { sp[0] = P.sx[p]; sp[1] = P.sy[p]; sp[2] = P.sz[p]; }
            if (!push32_core<M, SPIN>(cx, cy, cz, m, f, u, P.w[p], K, fr, sink, nc, cnt12, sp, spin_bad)) {
                oob++;
                continue;
            }
            P.cell[p] = nc;
            P.fx[p] = f[0]; P.fy[p] = f[1]; P.fz[p] = f[2];
            P.ux[p] = u[0]; P.uy[p] = u[1]; P.uz[p] = u[2];
            if (SPIN) { P.sx[p] = sp[0]; P.sy[p] = sp[1]; P.sz[p] = sp[2]; }
        }
    }
    __syncthreads();
    float* jdst[4] = {F.jx, F.jy, F.jz, F.rho};
    for (int e = threadIdx.x; e < 4 * B::AVOL; e += blockDim.x) {
        const ACC a = acc[e];
        if (a == ACC(0)) continue;
        const int c = e / B::AVOL, r = e % B::AVOL;
        const float v = sizeof(ACC) == 8 ? (float)a : (float)a * K.ufix[c];
        const int gx = aox + r % B::AX, gy = aoy + (r / B::AX) % B::AY, gz = aoz + r / (B::AX * B::AY);
        if (gx > g.je[0] - 2 || gy > g.je[1] - 2 || gz > g.je[2] - 2) continue;
        hbm_add(&jdst[c][g.ji(gz, gy, gx)], v);
    }
    if (threadIdx.x < 12 && cnt12[threadIdx.x])
        atomicAdd((unsigned long long*)&cnt[CNT_PEND + threadIdx.x], (unsigned long long)cnt12[threadIdx.x]);
    if (oob) atomicAdd((unsigned long long*)&cnt[CNT_OOB], (unsigned long long)oob);
    if (SPIN && spin_bad) atomicAdd((unsigned long long*)&cnt[CNT_SPIN], (unsigned long long)spin_bad);
}

// ---------------------------------------------------------------------
// Two-kernel push over cell-ordered tiles (fixed-point accumulators).
//  k_push32_gat: gather, Boris, move, spin and the write-back of a tile
//   chunk's particles in storage order, the field tile alone in LDS; the
//   lanes of a wave hold the particles of a few neighbouring cells (the
//   field reads mostly broadcast).  Per slot it writes a deposit record:
//   the old cell c (-1: nothing to deposit; -2 - c: the particle changed
//   its nearest node along some axis), old offsets, unwrapped new offsets.
//   A particle outside its tile takes the HBM path, deposit included.
//  k_push32_dep: the deposit of those records into the accumulator box
//   alone in LDS.  The chunk's particles are listed in two classes, those
//   that kept their nearest node along every axis (a box one line shorter
//   per axis, fixed at compile time) and the rest; each list is deposited
//   in passes of 256 where thread t takes entry spread_pick(t), then the
//   particles are dealt by the bank of their deposit base.  A wave runs the
//   union of its lanes' box lines, so a class of its own for the still
//   particles saves the lines only moving neighbours need.
// Each kernel holds half of the LDS and fewer registers than the one-kernel
// forms, so more waves share a CU.
// ---------------------------------------------------------------------
struct Rec32 {
    int32_t* cell;
    float *f0x, *f0y, *f0z, *f1x, *f1y, *f1z;
};
static Rec32 rec32(Species& S) {
    const int64_t n = std::max<int64_t>(S.cap, 1);
    float* b = (float*)S.rec32;
    return Rec32{(int32_t*)b, b + n, b + 2 * n, b + 3 * n, b + 4 * n, b + 5 * n, b + 6 * n};
}

template <int M, int TX, int TY, int TZ, bool SPIN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_push32_gat(
    Soa32 P, Fields32 F, Grid g, Consts32 K, Tiling tl, const int64_t* __restrict__ tile_off,
    const int64_t* __restrict__ work, const int64_t* __restrict__ nwork, int chunk, Rec32 R,
    int64_t* __restrict__ cnt) {
    using B = Box32<M, TX, TY, TZ, int>;
    extern __shared__ double lds_d[];
    float* fld = reinterpret_cast<float*>(lds_d);  // the six field boxes
    __shared__ unsigned cnt12[12];
    const int64_t wi = blockIdx.x;
    if (wi >= *nwork) return;
    const int64_t item = work[wi];
    const int tile = (int)(item >> 20);
    const int64_t cidx = item & ((1 << 20) - 1);
    const int64_t beg = tile_off[tile] + cidx * chunk;
    const int64_t end = min(tile_off[tile + 1], beg + chunk);
    const int txi = tile % tl.nt[0], tyi = (tile / tl.nt[0]) % tl.nt[1], tzi = tile / (tl.nt[0] * tl.nt[1]);
    const int T0x = txi * TX, T0y = tyi * TY, T0z = tzi * TZ;
    const int m = g.m;
    const int fox = T0x + m - 1 - B::H, foy = T0y + m - 1 - B::H, foz = T0z + m - 1 - B::H;
    const float* fsrc[6] = {F.ex, F.ey, F.ez, F.bx, F.by, F.bz};
    auto load = [&](int c, const float* src) {
        const int DX = B::cdim(c, 0), DY = B::cdim(c, 1);
        float* dst = fld + B::coff(c);
        for (int r = threadIdx.x; r < DX * DY * B::cdim(c, 2); r += blockDim.x) {
            const int lx = r % DX, ly = (r / DX) % DY, lz = r / (DX * DY);
            const int gx = min(fox + lx, g.fe[0] - 1), gy = min(foy + ly, g.fe[1] - 1), gz = min(foz + lz, g.fe[2] - 1);
            dst[lz * B::planep(c) + ly * B::rowp(c) + lx] = src[g.fi(gz, gy, gx)];
        }
    };
    load(0, F.ex); load(1, F.ey); load(2, F.ez); load(3, F.bx); load(4, F.by); load(5, F.bz);
    if (threadIdx.x < 12) cnt12[threadIdx.x] = 0;
    __syncthreads();
    int64_t oob = 0;
    int spin_bad = 0;
    uint32_t strays = 0;  // passes whose particle was outside the tile
    const int len = (int)(end - beg);
    int it = 0;
    for (int t0 = 0; t0 < len; t0 += blockDim.x, it++) {
        const int t = t0 + threadIdx.x;
        if (t >= len) continue;
        const int64_t p = beg + t;
        const int cell = P.cell[p];
        int rc = -1;
        float f0[3] = {0.0f, 0.0f, 0.0f}, f1[3] = {0.0f, 0.0f, 0.0f};
        if (cell >= 0) {
            int cx, cy, cz;
            K.cs.split(cell, cx, cy, cz);
            if ((unsigned)(cx - T0x) >= (unsigned)TX || (unsigned)(cy - T0y) >= (unsigned)TY ||
                (unsigned)(cz - T0z) >= (unsigned)TZ) {
                strays |= 1u << it;  // moved out of the tile since the last sort
            } else {
                float f[3] = {P.fx[p], P.fy[p], P.fz[p]};
                float u[3] = {P.ux[p], P.uy[p], P.uz[p]};
                float sp[3] = {0.0f, 0.0f, 0.0f};
                if (SPIN) { sp[0] = P.sx[p]; sp[1] = P.sy[p]; sp[2] = P.sz[p]; }
                auto fr = [&](int c, int k, int j, int i) -> float {
                    const float* fb = fld + B::coff(c) + (cz + m - 1 - foz) * B::planep(c) +
                                      (cy + m - 1 - foy) * B::rowp(c) + (cx + m - 1 - fox);
                    return fb[(k - (cz + m - 1)) * B::planep(c) + (j - (cy + m - 1)) * B::rowp(c) +
                              (i - (cx + m - 1))];
                };
                float E[3], Bf[3], W0[3][M + 1];
                gather32<M>(cx + m - 1, cy + m - 1, cz + m - 1, f[0], f[1], f[2], E, Bf, fr, W0[0], W0[1], W0[2]);
                boris32(u, E, Bf, K.qm);
                if constexpr (SPIN) {
                    const float gam = sqrtf(1.0f + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
                    if (!spin_precess<float>(E, Bf, K.qm, u[0], u[1], u[2], gam, K.spin, sp[0], sp[1], sp[2]))
                        spin_bad++;
                }
                const float ig = rsqrtf(1.0f + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
                const float d[3] = {u[0] * ig * K.cfx, u[1] * ig * K.cfy, u[2] * ig * K.cfz};
                // a non-finite momentum would put the deposit box outside the arrays
                if (!(fabsf(d[0]) < 1.0f && fabsf(d[1]) < 1.0f && fabsf(d[2]) < 1.0f)) {
                    oob++;
                } else {
                    bool moved = false;
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        f0[a] = f[a];
                        f1[a] = f[a] + d[a];
                        moved |= (f0[a] >= 0.5f ? 1 : 0) != (int)floorf(f1[a] + 0.5f);
                    }
                    rc = moved ? -2 - cell : cell;
                    const int nc = move32(cx, cy, cz, f, d, K, cnt12);
                    P.cell[p] = nc;
                    P.fx[p] = f[0]; P.fy[p] = f[1]; P.fz[p] = f[2];
                    P.ux[p] = u[0]; P.uy[p] = u[1]; P.uz[p] = u[2];
                    if (SPIN) { P.sx[p] = sp[0]; P.sy[p] = sp[1]; P.sz[p] = sp[2]; }
                }
            }
        }
        R.cell[p] = rc;
        R.f0x[p] = f0[0]; R.f0y[p] = f0[1]; R.f0z[p] = f0[2];
        R.f1x[p] = f1[0]; R.f1y[p] = f1[1]; R.f1z[p] = f1[2];
    }
    if (strays) {  // HBM path for the particles outside the tile
        float* jdst[4] = {F.jx, F.jy, F.jz, F.rho};
        it = 0;
        for (int t = threadIdx.x; t < len; t += blockDim.x, it++) {
            if (!(strays & (1u << it))) continue;
            const int64_t p = beg + t;
            const int cell = P.cell[p];
            int cx, cy, cz;
            K.cs.split(cell, cx, cy, cz);
            float f[3] = {P.fx[p], P.fy[p], P.fz[p]};
            float u[3] = {P.ux[p], P.uy[p], P.uz[p]};
            auto fr = [&](int c, int k, int j, int i) -> float { return fsrc[c][g.fi(k, j, i)]; };
            HbmSink32 sink{{jdst[0], jdst[1], jdst[2], jdst[3]}, g, cz + m - 1, cy + m - 1, cx + m - 1, 0, 0, 0};
            int nc;
            float sp[3];
            if (SPIN) { sp[0] = P.sx[p]; sp[1] = P.sy[p]; sp[2] = P.sz[p]; }
            if (!push32_core<M, SPIN>(cx, cy, cz, m, f, u, P.w[p], K, fr, sink, nc, cnt12, sp, spin_bad)) {
                oob++;
                continue;
            }
            P.cell[p] = nc;
            P.fx[p] = f[0]; P.fy[p] = f[1]; P.fz[p] = f[2];
            P.ux[p] = u[0]; P.uy[p] = u[1]; P.uz[p] = u[2];
            if (SPIN) { P.sx[p] = sp[0]; P.sy[p] = sp[1]; P.sz[p] = sp[2]; }
        }
    }
    __syncthreads();
    if (threadIdx.x < 12 && cnt12[threadIdx.x])
        atomicAdd((unsigned long long*)&cnt[CNT_PEND + threadIdx.x], (unsigned long long)cnt12[threadIdx.x]);
    if (oob) atomicAdd((unsigned long long*)&cnt[CNT_OOB], (unsigned long long)oob);
    if (SPIN && spin_bad) atomicAdd((unsigned long long*)&cnt[CNT_SPIN], (unsigned long long)spin_bad);
}

template <int M, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_push32_dep(
    Soa32 P, Fields32 F, Grid g, Consts32 K, Tiling tl, const int64_t* __restrict__ tile_off,
    const int64_t* __restrict__ work, const int64_t* __restrict__ nwork, int chunk, Rec32 R) {
    using B = Box32<M, TX, TY, TZ, int>;
    static_assert(PIC_PUSH_CHUNK <= 65536, "chunk slots are listed as 16-bit offsets");
    extern __shared__ double lds_d[];
    int* acc = reinterpret_cast<int*>(lds_d);                     // 4 * AVOL accumulators
    uint16_t* lst = reinterpret_cast<uint16_t*>(acc + 4 * B::AVOL);  // chunk slots by class
    __shared__ unsigned rg_hist[256 / 64][8];  // per wave: bank-spreading counters
    __shared__ int ncls[2];
    const int64_t wi = blockIdx.x;
    if (wi >= *nwork) return;
    const int64_t item = work[wi];
    const int tile = (int)(item >> 20);
    const int64_t cidx = item & ((1 << 20) - 1);
    const int64_t beg = tile_off[tile] + cidx * chunk;
    const int64_t end = min(tile_off[tile + 1], beg + chunk);
    const int txi = tile % tl.nt[0], tyi = (tile / tl.nt[0]) % tl.nt[1], tzi = tile / (tl.nt[0] * tl.nt[1]);
    const int T0x = txi * TX, T0y = tyi * TY, T0z = tzi * TZ;
    const int m = g.m;
    for (int e = threadIdx.x; e < 4 * B::AVOL; e += blockDim.x) acc[e] = 0;
    const int wv = threadIdx.x >> 6, lane = lane_id();
    if (lane < 8) rg_hist[wv][lane] = 0u;
    if (threadIdx.x < 2) ncls[threadIdx.x] = 0;
    __syncthreads();
    const int len = (int)(end - beg);
    // the two lists: still particles from the front, the others from the back
    const uint64_t below = (1ull << lane) - 1ull;
    for (int t = threadIdx.x; t - lane < len; t += blockDim.x) {
        const int c = t < len ? R.cell[beg + t] : -1;
        const uint64_t m0 = __ballot(c >= 0), m1 = __ballot(c < -1);
        int o0 = 0, o1 = 0;
        if (lane == 0) {
            o0 = m0 ? atomicAdd(&ncls[0], __popcll(m0)) : 0;
            o1 = m1 ? atomicAdd(&ncls[1], __popcll(m1)) : 0;
        }
        o0 = __shfl(o0, 0);
        o1 = __shfl(o1, 0);
        if (c >= 0) lst[o0 + __popcll(m0 & below)] = (uint16_t)t;
        if (c < -1) lst[len - 1 - (o1 + __popcll(m1 & below))] = (uint16_t)t;
    }
    __syncthreads();
    const int aox = T0x + m - 2 - B::H, aoy = T0y + m - 2 - B::H, aoz = T0z + m - 2 - B::H;
    auto run = [&](auto still) {
        constexpr bool STILL = decltype(still)::value;
        const int n = ncls[STILL ? 0 : 1];
        for (int pass = 0; 4 * pass < (n + 63) >> 6; pass++) {
            const int e = spread_row(threadIdx.x, pass, n), pk = e;
            if ((threadIdx.x >> 6) + 4 * pass >= (n + 63) >> 6) break;  // this wave's rows are done
            int bd = -1;
            float a0[3] = {0.0f, 0.0f, 0.0f}, a1[3] = {0.0f, 0.0f, 0.0f}, q = 0.0f;
            if (pk >= 0) {
                const int64_t p = beg + lst[STILL ? e : len - 1 - e];
                const int rc = R.cell[p];
                const int cell = STILL ? rc : -2 - rc;
                int cx, cy, cz;
                K.cs.split(cell, cx, cy, cz);
                bd = ((cz + m - 1 - aoz) * B::AY + (cy + m - 1 - aoy)) * B::AX + (cx + m - 1 - aox);
                a0[0] = R.f0x[p]; a0[1] = R.f0y[p]; a0[2] = R.f0z[p];
                a1[0] = R.f1x[p]; a1[1] = R.f1y[p]; a1[2] = R.f1z[p];
                q = K.q * P.w[p];
            }
            {
                // bank of the deposit base: the cell's box offset moved to
                // the nearest node of the old position
                const int r = bd < 0 ? 32
                                     : (bd + (a0[0] >= 0.5f) + (a0[1] >= 0.5f) * B::AX +
                                        (a0[2] >= 0.5f) * (B::AX * B::AY)) & 31;
                const int to = bank_spread_lane(r, rg_hist[wv]);
                bd = lane_send(to, bd);
#pragma unroll
                for (int a = 0; a < 3; a++) { a0[a] = lane_send(to, a0[a]); a1[a] = lane_send(to, a1[a]); }
                q = lane_send(to, q);
            }
            if (bd >= 0) {
                // the old position's primal weights, as the gather evaluated them
                float S0[3][M + 1];
#pragma unroll
                for (int a = 0; a < 3; a++) shape32<M>(a0[a] - (float)(a0[a] >= 0.5f ? 1 : 0), S0[a]);
                LdsSink32<B::AX, B::AY, B::AVOL, int> sink{acc, K.inv_ufix, bd, 0};
                esirkepov32<M, decltype(sink), STILL>(q, a0, a1, S0[0], S0[1], S0[2], K, sink);
            }
        }
    };
    run(std::true_type{});
    run(std::false_type{});
    __syncthreads();
    float* jdst[4] = {F.jx, F.jy, F.jz, F.rho};
    for (int e = threadIdx.x; e < 4 * B::AVOL; e += blockDim.x) {
        const int a = acc[e];
        if (a == 0) continue;
        const int c = e / B::AVOL, r = e % B::AVOL;
        const float v = (float)a * K.ufix[c];
        const int gx = aox + r % B::AX, gy = aoy + (r / B::AX) % B::AY, gz = aoz + r / (B::AX * B::AY);
        if (gx > g.je[0] - 2 || gy > g.je[1] - 2 || gz > g.je[2] - 2) continue;
        hbm_add(&jdst[c][g.ji(gz, gy, gx)], v);
    }
}

// direct push: any particle order, gather from HBM, global atomics
template <int M, bool SPIN>
__global__ __launch_bounds__(256) void k_push32_direct(Soa32 P, Fields32 F, Grid g, Consts32 K,
                                                       int64_t* __restrict__ cnt) {
    __shared__ unsigned cnt12[12];
    if (threadIdx.x < 12) cnt12[threadIdx.x] = 0;
    __syncthreads();
    const int64_t n = cnt[CNT_SLOTS];
    const float* fsrc[6] = {F.ex, F.ey, F.ez, F.bx, F.by, F.bz};
    const int m = g.m;
    int64_t oob = 0;
    int spin_bad = 0;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int cell = P.cell[p];
        if (cell < 0) continue;
        int cx, cy, cz;
        K.cs.split(cell, cx, cy, cz);
        float f[3] = {P.fx[p], P.fy[p], P.fz[p]};
        float u[3] = {P.ux[p], P.uy[p], P.uz[p]};
        auto fr = [&](int c, int k, int j, int i) -> float { return fsrc[c][g.fi(k, j, i)]; };
        HbmSink32 sink{{F.jx, F.jy, F.jz, F.rho}, g, cz + m - 1, cy + m - 1, cx + m - 1, 0, 0, 0};
        int nc;
        float sp[3];
        if (SPIN) { sp[0] = P.sx[p]; sp[1] = P.sy[p]; sp[2] = P.sz[p]; }
        if (!push32_core<M, SPIN>(cx, cy, cz, m, f, u, P.w[p], K, fr, sink, nc, cnt12, sp, spin_bad)) {
            oob++;
            continue;
        }
        P.cell[p] = nc;
        P.fx[p] = f[0]; P.fy[p] = f[1]; P.fz[p] = f[2];
        P.ux[p] = u[0]; P.uy[p] = u[1]; P.uz[p] = u[2];
        if (SPIN) { P.sx[p] = sp[0]; P.sy[p] = sp[1]; P.sz[p] = sp[2]; }
    }
    __syncthreads();
    if (threadIdx.x < 12 && cnt12[threadIdx.x])
        atomicAdd((unsigned long long*)&cnt[CNT_PEND + threadIdx.x], (unsigned long long)cnt12[threadIdx.x]);
    if (oob) atomicAdd((unsigned long long*)&cnt[CNT_OOB], (unsigned long long)oob);
    if (SPIN && spin_bad) atomicAdd((unsigned long long*)&cnt[CNT_SPIN], (unsigned long long)spin_bad);
}

// ---------------------------------------------------------------------
// tile sort (fp32 layout): drops removed particles (cell < 0)
// ---------------------------------------------------------------------
// Push chunk of the FP32 pushes: PIC_PUSH_CHUNK_Q32 for the quadratic
// one-kernel push with fixed-point accumulators outside differentiated
// runs, else PIC_PUSH_CHUNK (the two-kernel deposit lists a chunk in LDS).
static int push_chunk32(const pic_ctx* c) {
    return (c->fix32 && c->g.m == 2 && !c->diff) ? PIC_PUSH_CHUNK_Q32 : PIC_PUSH_CHUNK;
}

struct TileArgs32 {
    int n[3];
    Tiling tl;
    CellSplit cs;
    FastDiv td[3];  // tile edges
    int cpt;        // bins per tile: 1, or its cells (cell order inside tiles)
    int chunk;      // push chunk of the in-tile transpose
};

// sort bin of a cell: its tile, or (cpt > 1) the tile's first bin plus the
// cell's index inside the tile, x fastest
__device__ __forceinline__ int64_t bin_of32(int cell, const TileArgs32& A) {
    int cx, cy, cz;
    A.cs.split(cell, cx, cy, cz);
    const int tx = A.td[0].div(cx), ty = A.td[1].div(cy), tz = A.td[2].div(cz);
    const int64_t tile = (tz * A.tl.nt[1] + ty) * A.tl.nt[0] + tx;
    if (A.cpt == 1) return tile;
    const int lx = cx - tx * A.tl.t[0], ly = cy - ty * A.tl.t[1], lz = cz - tz * A.tl.t[2];
    return tile * A.cpt + (lz * A.tl.t[1] + ly) * A.tl.t[0] + lx;
}

constexpr int SORT_R32 = 4;  // rows of 64 slots per wave and block iteration

// keep_dead: removed slots counted in bin ntiles instead of dropped
__global__ __launch_bounds__(256) void k_hist32(const int32_t* __restrict__ cellv,
                                                 const int64_t* __restrict__ cnt,
                                                 int64_t* __restrict__ tile_cnt, TileArgs32 T, int keep_dead) {
    constexpr int R = SORT_R32;
    int64_t lo, hi;
    block_range<R>(cnt[CNT_SLOTS], lo, hi);
    const int64_t wseg = (threadIdx.x >> 6) * 64 * R + lane_id();
    for (int64_t b = lo; b < hi; b += (int64_t)blockDim.x * R) {
        int cell[R];
        int64_t key[R];
        bool valid[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t p = b + wseg + r * 64;
            cell[r] = p < hi ? cellv[p] : -1;
        }
        bool uni = true;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t p = b + wseg + r * 64;
            valid[r] = cell[r] >= 0 || (keep_dead && p < hi);
            key[r] = cell[r] >= 0 ? bin_of32(cell[r], T) : (valid[r] ? T.tl.nbins() : -1);
            uni = uni && valid[r] && key[r] == key[0];
        }
        const int64_t k0 = __shfl(key[0], 0);
        if (__all(uni && key[0] == k0)) {  // the wave's 64 R slots in one bin
            if (lane_id() == 0) atomicAdd((unsigned long long*)&tile_cnt[k0], (unsigned long long)(64 * R));
            continue;
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            int leader, rank, count;
            wave_groups(key[r], valid[r], leader, rank, count);
            if (valid[r] && leader == lane_id())
                atomicAdd((unsigned long long*)&tile_cnt[key[r]], (unsigned long long)count);
        }
    }
}

// perm != null: removed slots go after the last tile (bin ntiles) and
// perm[q] records the source slot of slot q.  Each wave moves 64 R
// consecutive slots per iteration; when they share one bin (the common
// case for tile-ordered input) one cursor atomic places them contiguously
// and in order.
__global__ __launch_bounds__(256) void k_scatter32(Soa32 A, Soa32 Bo, const int64_t* __restrict__ cnt,
                                                    int64_t* __restrict__ tile_cur, TileArgs32 T,
                                                    int32_t* __restrict__ perm, const int64_t* __restrict__ tile_off,
                                                    int scramble) {
    constexpr int R = SORT_R32;
    int64_t lo, hi;
    block_range<R>(cnt[CNT_N - 1], lo, hi);
    const int64_t wseg = (threadIdx.x >> 6) * 64 * R + lane_id();
    for (int64_t b = lo; b < hi; b += (int64_t)blockDim.x * R) {
        int cell[R], head[R];
        int64_t key[R], base[R], id[R];
        float v[R][7], sv[R][3];
        bool valid[R];
        // every row's loads in flight together
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t p = b + wseg + r * 64;
            cell[r] = -1;
            if (p < hi) {
                cell[r] = A.cell[p];
                v[r][0] = A.fx[p]; v[r][1] = A.fy[p]; v[r][2] = A.fz[p];
                v[r][3] = A.ux[p]; v[r][4] = A.uy[p]; v[r][5] = A.uz[p];
                v[r][6] = A.w[p];
                id[r] = A.id[p];
                if (A.sx) { sv[r][0] = A.sx[p]; sv[r][1] = A.sy[p]; sv[r][2] = A.sz[p]; }
            }
        }
        bool uni = true;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t p = b + wseg + r * 64;
            valid[r] = cell[r] >= 0 || (perm && p < hi);
            key[r] = cell[r] >= 0 ? bin_of32(cell[r], T) : (valid[r] ? T.tl.nbins() : -1);
            uni = uni && valid[r] && key[r] == key[0];
        }
        const int64_t k0 = __shfl(key[0], 0);
        if (__all(uni && key[0] == k0)) {
            int64_t b0 = 0;
            if (lane_id() == 0) b0 = (int64_t)atomicAdd((unsigned long long*)&tile_cur[k0], (unsigned long long)(64 * R));
            b0 = __shfl(b0, 0);
#pragma unroll
            for (int r = 0; r < R; r++) { head[r] = 0; base[r] = b0 + r * 64; }
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) {
                int leader, rank, count;
                wave_groups(key[r], valid[r], leader, rank, count);
                base[r] = 0;
                if (valid[r] && leader == lane_id())
                    base[r] = (int64_t)atomicAdd((unsigned long long*)&tile_cur[key[r]], (unsigned long long)count);
                base[r] = __shfl(base[r], leader) + rank;
                head[r] = lane_id();
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t p = b + wseg + r * 64;
            if (!valid[r]) continue;
            const int64_t q = placed_slot(base[r] + (lane_id() - head[r]), key[r], T.tl.ntiles, tile_off, scramble,
                                            T.chunk);
            Bo.cell[q] = cell[r];
            Bo.fx[q] = v[r][0]; Bo.fy[q] = v[r][1]; Bo.fz[q] = v[r][2];
            Bo.ux[q] = v[r][3]; Bo.uy[q] = v[r][4]; Bo.uz[q] = v[r][5];
            Bo.w[q] = v[r][6];
            Bo.id[q] = id[r];
            if (A.sx) { Bo.sx[q] = sv[r][0]; Bo.sy[q] = sv[r][1]; Bo.sz[q] = sv[r][2]; }
            if (perm) perm[q] = (int32_t)p;
        }
    }
}

int sort32(pic_ctx* c, int s, int32_t* perm) {
    Species& S = c->sp[s];
    // Cell order for the two-kernel push: every cell_order-th sort bins
    // by cell; the tile sorts in between keep each tile's relative order
    // (only particles that changed cell inside their tile, or arrived from
    // another tile, then sit out of cell order).  Differentiated runs
    // (which record and replay tile-sort permutations) keep tile bins.
    const bool cellmode = c->tl.fine && !perm && !c->diff;
    const bool fine = cellmode && (!S.cellsorted || S.sorts_since_fine + 1 >= c->cell_order);
    TileArgs32 T;
    T.cpt = fine ? c->tl.cells() : 1;
    T.chunk = push_chunk32(c);
    const int64_t nbins = c->tl.nbins() * T.cpt + (perm ? 1 : 0);
    int r = launch_fill(c, c->tile_cnt, nbins * 8);
    if (r) return r;
    for (int a = 0; a < 3; a++) { T.n[a] = c->g.n[a]; T.td[a] = FastDiv(c->tl.t[a]); }
    T.tl = c->tl;
    T.cs = CellSplit(c->g.n[0], c->g.n[1]);
    const int blocks = (int)std::min<int64_t>(blocks_for(std::max<int64_t>(S.cap, 1), 256), 256 * 32);
    k_hist32<<<blocks, 256, 0, c->stream>>>(S.cell[S.cur], S.d_cnt, c->tile_cnt, T, perm ? 1 : 0);
    PIC_LAUNCH_CHECK(c);
    PIC_HIP(c, hipMemcpyAsync(S.d_cnt + (CNT_N - 1), S.d_cnt + CNT_SLOTS, 8,
                              hipMemcpyDeviceToDevice, c->stream));
    r = launch_scan_bins(c, S, nbins, T.cpt);
    if (r) return r;
    // in cell mode no chunk is transposed (a tile sort outside it
    // transposes the chunks once again)
    const int scramble = cellmode ? 0 : !S.scrambled;
    k_scatter32<<<blocks, 256, 0, c->stream>>>(soa32(S, S.cur), soa32(S, 1 - S.cur), S.d_cnt,
                                               c->tile_cur, T, perm, S.tile_off, scramble);
    if (scramble) S.scrambled = 1;
    if (cellmode) {
        S.scrambled = 0;
        S.cellsorted = 1;
        S.sorts_since_fine = fine ? 0 : S.sorts_since_fine + 1;
    } else {
        S.cellsorted = 0;
    }
    PIC_LAUNCH_CHECK(c);
    S.cur = 1 - S.cur;
    S.sorted = 1;
    S.pushes_since_sort = 0;
    return PIC_OK;
}

// ---------------------------------------------------------------------
// boundary (fp32): the push already moved and classified every particle;
// publish its pending counts as this call's counts.
// ---------------------------------------------------------------------
__global__ void k_boundary32(int64_t* __restrict__ cnt) {
    const int i = threadIdx.x;
    if (i >= 12) return;
    const int64_t v = cnt[CNT_PEND + i];
    cnt[CNT_PEND + i] = 0;
    cnt[(i < 6 ? CNT_LOST_STEP : CNT_WRAP_STEP) + i % 6] = v;
    cnt[(i < 6 ? CNT_LOST : CNT_WRAP) + i % 6] += v;
    if (i < 6 && v) atomicAdd((unsigned long long*)&cnt[CNT_LIVE], (unsigned long long)(-v));
}

int boundary32(pic_ctx* c, int s) {
    k_boundary32<<<1, 64, 0, c->stream>>>(c->sp[s].d_cnt);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

// ---------------------------------------------------------------------
// loader (fp32 layout) from the shared seeded generator
// ---------------------------------------------------------------------
struct LoaderArgs32 {
    uint64_t seed;
    int64_t id_offset;
    double u_th, wt;
    int n[3];
};

__global__ __launch_bounds__(256) void k_loader32(Soa32 A, const int64_t* __restrict__ prefix,
                                                   int64_t n, LoaderArgs32 L) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        int32_t ix, iy, iz;
        pic_locate(p, prefix, L.n[0], L.n[1], &ix, &iy, &iz);
        const uint64_t pid = (uint64_t)(L.id_offset + p);
        double x, y, z, ux, uy, uz;
        pic_load_one(L.seed, pid, ix, iy, iz, L.u_th, &x, &y, &z, &ux, &uy, &uz);
        int cx, cy, cz;
        float fx, fy, fz;
        to_cell_offset(x, L.n[0], cx, fx);
        to_cell_offset(y, L.n[1], cy, fy);
        to_cell_offset(z, L.n[2], cz, fz);
        A.cell[p] = (cz * L.n[1] + cy) * L.n[0] + cx;
        A.fx[p] = fx; A.fy[p] = fy; A.fz[p] = fz;
        A.ux[p] = (float)ux; A.uy[p] = (float)uy; A.uz[p] = (float)uz;
        A.w[p] = (float)L.wt;
        A.id[p] = (int64_t)pid;
    }
}

int loader32(pic_ctx* c, int s, const pic_loader* l, int64_t n, const int64_t* d_prefix) {
    Species& S = c->sp[s];
    LoaderArgs32 L{l->seed, l->id_offset, l->u_th, l->density / l->ppc, {c->g.n[0], c->g.n[1], c->g.n[2]}};
    const int blocks = (int)std::min<int64_t>(blocks_for(std::max<int64_t>(n, 1), 256), 256 * 64);
    k_loader32<<<blocks, 256, 0, c->stream>>>(soa32(S, S.cur), d_prefix, n, L);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

// ---------------------------------------------------------------------
// diagnostics (fp32 particles, fp64 sums)
// ---------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_diag_p32(Soa32 A, const int64_t* __restrict__ cnt, double mass,
                                                   int s, double emax, double* __restrict__ buf) {
    __shared__ double hist[PIC_SPECTRUM_BINS];
    __shared__ double ke_s;
    for (int b = threadIdx.x; b < PIC_SPECTRUM_BINS; b += blockDim.x) hist[b] = 0;
    if (threadIdx.x == 0) ke_s = 0;
    __syncthreads();
    const int64_t n = cnt[CNT_SLOTS];
    double ke = 0;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        if (A.cell[p] < 0) continue;
        const double ux = A.ux[p], uy = A.uy[p], uz = A.uz[p], wp = A.w[p];
        const double u2 = ux * ux + uy * uy + uz * uz;
        const double gm1 = u2 / (sqrt(1.0 + u2) + 1.0);
        ke += mass * wp * gm1;
        if (s == 0) {
            const int b = (int)(gm1 / emax * PIC_SPECTRUM_BINS);
            if (b >= 0 && b < PIC_SPECTRUM_BINS) lds_add(&hist[b], wp);
        }
    }
    for (int off = 32; off > 0; off >>= 1) ke += __shfl_down(ke, off);
    if (lane_id() == 0) atomicAdd(&ke_s, ke);
    __syncthreads();
    if (threadIdx.x == 0) hbm_add(&buf[DG_KIN + s], ke_s);
    if (s == 0)
        for (int b = threadIdx.x; b < PIC_SPECTRUM_BINS; b += blockDim.x)
            if (hist[b] != 0) hbm_add(&buf[DG_SPEC + b], hist[b]);
}

int diag_p32(pic_ctx* c, int s, double emax) {
    Species& S = c->sp[s];
    const int blocks = (int)std::min<int64_t>(blocks_for(S.cap, 256), 256 * 16);
    k_diag_p32<<<blocks, 256, 0, c->stream>>>(soa32(S, S.cur), S.d_cnt, S.m, s, emax, c->diag_buf);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

// ---------------------------------------------------------------------
// push launcher
// ---------------------------------------------------------------------
static Consts32 consts32(const pic_ctx* c, const Species& S) {
    Consts32 K;
    K.qm = (float)(S.q / S.m * (c->p.dt / 2.0));
    K.q = (float)S.q;
    K.cfx = (float)(c->p.dt / c->p.dx);
    K.cfy = (float)(c->p.dt / c->p.dy);
    K.cfz = (float)(c->p.dt / c->p.dz);
    K.cjx = (float)(c->p.dx / c->p.dt);
    K.cjy = (float)(c->p.dy / c->p.dt);
    K.cjz = (float)(c->p.dz / c->p.dt);
    for (int a = 0; a < 3; a++) { K.n[a] = c->g.n[a]; K.per[a] = c->p.periodic[a]; }
    K.m = c->g.m;
    K.cs = CellSplit(c->g.n[0], c->g.n[1]);
    K.spin = c->spin;
    const double qw = std::fabs(S.q) * (S.wmax > 0 ? S.wmax : 1.0);
    const double bound[4] = {qw * c->p.dx / c->p.dt, qw * c->p.dy / c->p.dt, qw * c->p.dz / c->p.dt, qw};
    // 2^-18 of the bound for chunks of 2^12 particles, 2^-17 for 2^13: a
    // chunk's sum stays below 2^30 units
    int e = -30;
    for (int n = push_chunk32(c); n > 1; n >>= 1) e++;
    static_assert((PIC_PUSH_CHUNK & (PIC_PUSH_CHUNK - 1)) == 0 && (PIC_PUSH_CHUNK_Q32 & (PIC_PUSH_CHUNK_Q32 - 1)) == 0,
                  "chunks are powers of two");
    for (int k = 0; k < 4; k++) {
        const double u = std::ldexp(bound[k], e);
        K.ufix[k] = (float)u;
        K.inv_ufix[k] = (float)(1.0 / u);
    }
    return K;
}

template <int M, int KIND, typename ACC, bool SPIN>
static int launch_tiled32_s(pic_ctx* c, Species& S, Fields32 Fd, Consts32 K) {
    constexpr int TX = TileShape<KIND, M>::X, TY = TileShape<KIND, M>::Y, TZ = TileShape<KIND, M>::Z;
    using B = Box32<M, TX, TY, TZ, ACC>;
    const int chunk = push_chunk32(c);
    const int64_t max_items = (int64_t)c->tl.ntiles + S.cap / chunk + 1;
    PIC_HIP(c, hipFuncSetAttribute((const void*)k_push32_tiled<M, TX, TY, TZ, ACC, SPIN>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)B::LDS));
    k_push32_tiled<M, TX, TY, TZ, ACC, SPIN><<<(unsigned)max_items, Push32Threads<M, ACC, SPIN>::N, B::LDS,
                                               c->stream>>>(
        soa32(S, S.cur), Fd, c->g, K, c->tl, S.tile_off, c->work, c->nwork, chunk, S.d_cnt);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

template <int M, int KIND, typename ACC>
static int launch_tiled32(pic_ctx* c, Species& S, Fields32 Fd, Consts32 K) {
    return S.spin ? launch_tiled32_s<M, KIND, ACC, true>(c, S, Fd, K)
                  : launch_tiled32_s<M, KIND, ACC, false>(c, S, Fd, K);
}

template <int M, bool SPIN>
static int launch_split32_s(pic_ctx* c, Species& S, Fields32 Fd, Consts32 K) {
    constexpr int TX = TileShape<2, M>::X, TY = TileShape<2, M>::Y, TZ = TileShape<2, M>::Z;
    using B = Box32<M, TX, TY, TZ, int>;
    constexpr size_t lds_g = (size_t)B::FTOT * sizeof(float),
                     lds_d = (size_t)4 * B::AVOL * sizeof(int) + (size_t)PIC_PUSH_CHUNK * sizeof(uint16_t);
    if (!S.rec32 && hipMalloc(&S.rec32, (size_t)7 * std::max<int64_t>(S.cap, 1) * 4) != hipSuccess) {
        S.rec32 = nullptr;
        return set_err(c, PIC_ERR_NOMEM, "deposit records for %lld particles", (long long)S.cap);
    }
    const Rec32 R = rec32(S);
    const int64_t max_items = (int64_t)c->tl.ntiles + S.cap / PIC_PUSH_CHUNK + 1;
    PIC_HIP(c, hipFuncSetAttribute((const void*)k_push32_gat<M, TX, TY, TZ, SPIN>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_g));
    PIC_HIP(c, hipFuncSetAttribute((const void*)k_push32_dep<M, TX, TY, TZ>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_d));
    k_push32_gat<M, TX, TY, TZ, SPIN><<<(unsigned)max_items, 256, lds_g, c->stream>>>(
        soa32(S, S.cur), Fd, c->g, K, c->tl, S.tile_off, c->work, c->nwork, PIC_PUSH_CHUNK, R, S.d_cnt);
    PIC_LAUNCH_CHECK(c);
    k_push32_dep<M, TX, TY, TZ><<<(unsigned)max_items, 256, lds_d, c->stream>>>(
        soa32(S, S.cur), Fd, c->g, K, c->tl, S.tile_off, c->work, c->nwork, PIC_PUSH_CHUNK, R);
    PIC_LAUNCH_CHECK(c);
    return PIC_OK;
}

int launch_push32(pic_ctx* c, int s) {
    Species& S = c->sp[s];
    Fields32 Fd{(const float*)c->F[0], (const float*)c->F[1], (const float*)c->F[2],
                (const float*)c->F[3], (const float*)c->F[4], (const float*)c->F[5],
                (float*)c->J[0], (float*)c->J[1], (float*)c->J[2], (float*)c->rho[s]};
    Consts32 K = consts32(c, S);
    const int kernel = c->push_kernel == 0 ? 2 : c->push_kernel;
    if (kernel == 1) {
        PhaseTimer t(c, PIC_T_PUSH);
        const int blocks = (int)std::min<int64_t>(blocks_for(std::max<int64_t>(S.cap, 1), 256), 256 * 64);
        const Soa32 P = soa32(S, S.cur);
        if (S.spin) {
            if (c->g.m == 2) k_push32_direct<2, true><<<blocks, 256, 0, c->stream>>>(P, Fd, c->g, K, S.d_cnt);
            else k_push32_direct<4, true><<<blocks, 256, 0, c->stream>>>(P, Fd, c->g, K, S.d_cnt);
        } else {
            if (c->g.m == 2) k_push32_direct<2, false><<<blocks, 256, 0, c->stream>>>(P, Fd, c->g, K, S.d_cnt);
            else k_push32_direct<4, false><<<blocks, 256, 0, c->stream>>>(P, Fd, c->g, K, S.d_cnt);
        }
        PIC_LAUNCH_CHECK(c);
        return PIC_OK;
    }
    {
        PhaseTimer t(c, PIC_T_SORT);
        if (!S.sorted || S.pushes_since_sort >= c->sort_interval) {
            int r = sort32(c, s, nullptr);
            if (r) return r;
        }
        int r = build_work(c, s, push_chunk32(c));
        if (r) return r;
    }
    S.pushes_since_sort++;
    PhaseTimer t(c, PIC_T_PUSH);
    if (c->fix32 && c->g.m == 4 && S.cellsorted)
        return S.spin ? launch_split32_s<4, true>(c, S, Fd, K) : launch_split32_s<4, false>(c, S, Fd, K);
    if (c->fix32) {
        if (c->g.m == 2) return launch_tiled32<2, 2, int>(c, S, Fd, K);
        return launch_tiled32<4, 2, int>(c, S, Fd, K);
    }
    if (c->g.m == 2) return launch_tiled32<2, 1, double>(c, S, Fd, K);
    return launch_tiled32<4, 1, double>(c, S, Fd, K);
}

}  // namespace pic

// ---------------------------------------------------------------------
// host-side conversion for pic_particles_set / pic_particles_get
// ---------------------------------------------------------------------
using namespace pic;

int particles_set32(pic_ctx* c, int s, int64_t n, const double* x, const double* y, const double* z,
                    const double* ux, const double* uy, const double* uz, const double* w,
                    const int64_t* id) {
    Species& S = c->sp[s];
    std::vector<int32_t> cell(n);
    std::vector<float> a[7];
    for (auto& v : a) v.resize(n);
    for (int64_t q = 0; q < n; q++) {
        // a periodic coordinate outside [0, n) is wrapped first (the
        // reference wraps it at the next push)
        double v[3] = {x[q], y[q], z[q]};
        int cc[3];
        float ff[3];
        for (int ax = 0; ax < 3; ax++) {
            const int N = c->g.n[ax];
            if (c->p.periodic[ax]) {
                if (v[ax] < 0.0) v[ax] += N;
                if (v[ax] >= N) v[ax] -= N;
            }
            if (!(v[ax] >= 0.0 && v[ax] <= N))
                return set_err(c, PIC_ERR_ARG, "particle %lld outside the domain", (long long)q);
            to_cell_offset(v[ax], N, cc[ax], ff[ax]);
        }
        cell[q] = (cc[2] * c->g.n[1] + cc[1]) * c->g.n[0] + cc[0];
        a[0][q] = ff[0]; a[1][q] = ff[1]; a[2][q] = ff[2];
        a[3][q] = (float)ux[q]; a[4][q] = (float)uy[q]; a[5][q] = (float)uz[q];
        a[6][q] = (float)w[q];
    }
    const int b = S.cur;
    float* dst[7] = {S.fx[b], S.fy[b], S.fz[b], S.vx[b], S.vy[b], S.vz[b], S.fw[b]};
    if (n) {
        PIC_HIP(c, hipMemcpyAsync(S.cell[b], cell.data(), n * 4, hipMemcpyHostToDevice, c->stream));
        for (int f = 0; f < 7; f++)
            PIC_HIP(c, hipMemcpyAsync(dst[f], a[f].data(), n * 4, hipMemcpyHostToDevice, c->stream));
        PIC_HIP(c, hipMemcpyAsync(S.id[b], id, n * 8, hipMemcpyHostToDevice, c->stream));
    }
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int particles_get32(pic_ctx* c, int s, int64_t slots, std::vector<double>* a, std::vector<int64_t>& id) {
    Species& S = c->sp[s];
    const int b = S.cur;
    std::vector<int32_t> cell(slots);
    std::vector<float> t[7];
    float* src[7] = {S.fx[b], S.fy[b], S.fz[b], S.vx[b], S.vy[b], S.vz[b], S.fw[b]};
    if (slots) {
        PIC_HIP(c, hipMemcpyAsync(cell.data(), S.cell[b], slots * 4, hipMemcpyDeviceToHost, c->stream));
        for (int f = 0; f < 7; f++) {
            t[f].resize(slots);
            PIC_HIP(c, hipMemcpyAsync(t[f].data(), src[f], slots * 4, hipMemcpyDeviceToHost, c->stream));
        }
        PIC_HIP(c, hipMemcpyAsync(id.data(), S.id[b], slots * 8, hipMemcpyDeviceToHost, c->stream));
    }
    PIC_HIP(c, hipStreamSynchronize(c->stream));
    for (int f = 0; f < 10; f++) a[f].assign(slots, 0.0);
    const int nx = c->g.n[0], ny = c->g.n[1];
    for (int64_t p = 0; p < slots; p++) {
        if (cell[p] < 0) {  // removed
            a[6][p] = __builtin_nan("");
            continue;
        }
        const int cx = cell[p] % nx, cy = (cell[p] / nx) % ny, cz = cell[p] / (nx * ny);
        a[0][p] = cx + (double)t[0][p];
        a[1][p] = cy + (double)t[1][p];
        a[2][p] = cz + (double)t[2][p];
        for (int f = 3; f < 7; f++) a[f][p] = t[f][p];
    }
    return PIC_OK;
}
