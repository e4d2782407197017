// spin.h -- the electron spin precession of particle.cpp:318-368 with the
// g-factor table lookup of particle.cpp:603-637, for the push kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace pic {

// Table of g.txt (constants.cpp:114-136) on the device and the constants of
// the precession.  es = E_schwinger (constants.cpp:83).
struct SpinConsts {
    const double* g = nullptr;
    int ng = 0;
    double es = 1.0;
};

// g_factor, particle.cpp:603-637: 2.003219 below chi = 0.005, else the
// table entry of chi's bin (steps 1e-4, 1e-2, 1e-2, 1e-2, 0.1 over [0.005,
// 0.01), [0.01, 1), [1, 10), [10, 20), [20, 100]).  Outside [0, 100] (and
// for NaN) the reference prints an error and exits; here 'bad' is set and
// the caller leaves the spin unchanged.  For chi >= 99.9 the reference's
// index passes the last entry (it reads past its table); here it stops at
// the last entry.
__device__ __forceinline__ double g_factor_tab(double x, const SpinConsts& S, bool& bad) {
    if (x < 0.005 && x >= 0.0) return 2.003219;
    int idx;
    if (x >= 0.005 && x < 0.01) idx = (int)((x - 0.005) / 0.0001);
    else if (x >= 0.01 && x < 1.0) idx = (int)((x - 0.01) / 0.01) + 50;
    else if (x >= 1.0 && x < 10.0) idx = (int)((x - 1.0) / 0.01) + 150;
    else if (x >= 10.0 && x < 20.0) idx = (int)((x - 10.0) / 0.01) + 1050;
    else if (x >= 20.0 && x <= 100.0) idx = (int)((x - 20.0) / 0.1) + 2050;
    else {
        bad = true;
        return 2.0;
    }
    return S.g[idx < S.ng - 1 ? idx : S.ng - 1];
}

// One precession step of the spin (sx, sy, sz), particle.cpp:321-367, in
// the reference's operation order (bit-identical in FP64 when compiled
// without contraction).  E, B: the gathered fields at the old position (the
// reference divides the Boris push's E*qm, B*qm back by qm, qm = q/m dt/2);
// u, gam: the momentum and Lorentz factor after the push.  The rotation
// parameter uses q/m dt/2 again (particle.cpp:350), which is qm.  Returns
// false (spin unchanged) where g_factor has no value.
template <typename T>
__device__ __forceinline__ bool spin_precess(const T* E, const T* B, T qm, T ux, T uy, T uz, T gam,
                                             const SpinConsts& S, T& sx, T& sy, T& sz) {
    const T ex = E[0], ey = E[1], ez = E[2];
    const T bx = B[0], by = B[1], bz = B[2];
    const T vx = ux / gam, vy = uy / gam, vz = uz / gam;
    const T px0 = vx * ex, py0 = vy * ey, pz0 = vz * ez;
    const T fLx = ex + vy * bz - vz * by;
    const T fLy = ey + vz * bx - vx * bz;
    const T fLz = ez + vx * by - vy * bx;
    const T ve = px0 + py0 + pz0;
    T chi = fLx * fLx + fLy * fLy + fLz * fLz - ve * ve;
    chi = T(0) < chi ? chi : T(0);  // std::max(0.0, .) (0 for NaN as well)
    chi = sqrt(chi) / (T)S.es * gam;
    bool bad = false;
    const T g = (T)g_factor_tab((double)chi, S, bad);
    if (bad) return false;
    const T alpha = -(g / T(2) - T(1)) * gam / (T(1) + gam);
    const T beta = (g / T(2) - T(1)) + T(1) / gam;
    const T gamma = g / T(2) - gam / (T(1) + gam);
    const T vb = vx * bx + vy * by + vz * bz;
    const T ox = alpha * vb * vx + beta * bx + gamma * (vy * ez - vz * ey);
    const T oy = alpha * vb * vy + beta * by + gamma * (vz * ex - vx * ez);
    const T oz = alpha * vb * vz + beta * bz + gamma * (vx * ey - vy * ex);
    const T tx = ox * qm, ty = oy * qm, tz = oz * qm;
    const T tt = T(2) / (T(1) + tx * tx + ty * ty + tz * tz);
    const T sxr = sx + (sy * tz - sz * ty);
    const T syr = sy + (sz * tx - sx * tz);
    const T szr = sz + (sx * ty - sy * tx);
    const T nx = sx + (syr * tz - szr * ty) * tt;
    const T ny = sy + (szr * tx - sxr * tz) * tt;
    const T nz = sz + (sxr * ty - syr * tx) * tt;
    sx = nx;
    sy = ny;
    sz = nz;
    return true;
}

}  // namespace pic
