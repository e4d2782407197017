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

// quartic: weights at nodes ii-2 .. ii+2, d = x - (ii - 2) in [1.5, 2.5]
template <typename T> __device__ __forceinline__ void shape4_d(T d, T* W) {
    const T A0 = T(55.0 / 96.0), A1 = T(5.0 / 24.0), A2 = T(5.0 / 4.0), A3 = T(5.0 / 6.0),
            B0 = T(115.0 / 192.0), B2 = T(5.0 / 8.0);
    T t, t2;
    t = d - T(2.5);
    t2 = t * t;
    W[0] = t2 * t2 / T(24.0);
    t = d - T(1.0);
    t2 = t * t;
    W[1] = -t2 * t2 / T(6.0) + t2 * t * A3 - t2 * A2 + t * A1 + A0;
    t = d - T(2.0);
    t2 = t * t;
    W[2] = t2 * t2 / T(4.0) - t2 * B2 + B0;
    t = d - T(3.0);
    t2 = t * t;
    W[3] = -t2 * t2 / T(6.0) - t2 * t * A3 - t2 * A2 - t * A1 + A0;
    t = d - T(1.5);
    t2 = t * t;
    W[4] = t2 * t2 / T(24.0);
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
