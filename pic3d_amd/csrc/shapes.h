// shapes.h -- B-spline shape functions of the reference (particle.cpp:8-48,
// coefficients constants.cpp:84-85), same expressions and association
// order, for device code.
#pragma once

namespace pic {

// quadratic: weights at nodes ic-1, ic, ic+1 with ic = (int)(x + 0.5)
template <typename T> __device__ __forceinline__ void shape2_delta(T d, T* W) {
    W[0] = T(0.5) * (T(0.5) - d) * (T(0.5) - d);
    W[1] = T(0.75) - d * d;
    W[2] = T(0.5) * (T(0.5) + d) * (T(0.5) + d);
}

// x / 3 correctly rounded, for normal double x: q = RN(x * RN(1/3)), the
// remainder x - 3q is exact in one fma, and RN(q + r * RN(1/3)) lies within
// 7 ulp * 2^-55 of x / 3, while x / 3 (never a rounding midpoint: 3 times a
// midpoint is an odd multiple of the half-ulp grid, x an even one) lies at
// least ulp / 6 from every midpoint -- so the result is RN(x / 3).  With
// the exact scalings by 1/8 and 1/2 it gives the reference's x / 24 and
// x / 6 bit for bit without a division sequence.
__device__ __forceinline__ double div3_rn(double x) {
    const double y = 1.0 / 3.0;
    const double q = x * y;
    const double r = fma(-q, 3.0, x);
    return fma(r, y, q);
}

template <typename T> struct Div6_24 {  // a / 6 and a / 24 in the scalar T
    static __device__ __forceinline__ T d6(T a) { return a / T(6.0); }
    static __device__ __forceinline__ T d24(T a) { return a / T(24.0); }
};
// FP64: the reference's rounded quotients, bit for bit (the quartic
// arguments t^4 are 0 or >= 2^-208, so always normal)
template <> struct Div6_24<double> {
    static __device__ __forceinline__ double d6(double a) { return div3_rn(a) * 0.5; }
    static __device__ __forceinline__ double d24(double a) { return div3_rn(a) * 0.125; }
};
// FP32 form: reciprocal multiplies (within an ulp; no bit contract)
template <> struct Div6_24<float> {
    static __device__ __forceinline__ float d6(float a) { return a * (1.0f / 6.0f); }
    static __device__ __forceinline__ float d24(float a) { return a * (1.0f / 24.0f); }
};

// quartic: weights at nodes ii-2 .. ii+2, d = x - (ii - 2) in [1.5, 2.5]
template <typename T> __device__ __forceinline__ void shape4_d(T d, T* W) {
    using Q = Div6_24<T>;
    const T A0 = T(55.0 / 96.0), A1 = T(5.0 / 24.0), A2 = T(5.0 / 4.0), A3 = T(5.0 / 6.0),
            B0 = T(115.0 / 192.0), B2 = T(5.0 / 8.0);
    T t, t2;
    t = d - T(2.5);
    t2 = t * t;
    W[0] = Q::d24(t2 * t2);
    t = d - T(1.0);
    t2 = t * t;
    W[1] = Q::d6(-t2 * t2) + t2 * t * A3 - t2 * A2 + t * A1 + A0;
    t = d - T(2.0);
    t2 = t * t;
    W[2] = t2 * t2 / T(4.0) - t2 * B2 + B0;
    t = d - T(3.0);
    t2 = t * t;
    W[3] = Q::d6(-t2 * t2) - t2 * t * A3 - t2 * A2 - t * A1 + A0;
    t = d - T(1.5);
    t2 = t * t;
    W[4] = Q::d24(t2 * t2);
}

// shape_function_2/4 of a buffer coordinate x (x >= 0): returns the centre
// node (int)(x + 0.5) and fills M+1 weights.
template <int M> __device__ __forceinline__ int shape_at(double x, double* W) {
    const int ic = (int)(x + 0.5);
    if (M == 2) shape2_delta<double>(x - (double)ic, W);
    else shape4_d<double>(x - (double)(ic - 2), W);
    return ic;
}

}  // namespace pic
