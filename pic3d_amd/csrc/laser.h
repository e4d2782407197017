// laser.h -- the x-min/x-max laser source of boundary.cpp:8-40, shared by
// the boundary kernel (k_field.hip) and its adjoint (k_adj.hip).
#pragma once

#include <hip/hip_runtime.h>

struct pic_ctx;

namespace pic {

struct LaserArgs {
    double pi, l0, waist, xfocus, yfocus, zfocus, omega, phase, amp, dur, t_start, lx;
    double dt, dx, dy, dz, time;
    int pol, left, right;
    // when env_on, the x-min source uses env_val in place of
    // generate_laser_pulse(time, 0) (the differentiated envelope)
    int env_on;
    double env_val;
    // x-max Silver-Mueller columns: written (xw), E/Bx/J read (xre), By/Bz
    // read (xrb).  The reference: local_xcells+1 from local_xcells
    // (boundary.cpp:92-96); the face form: n+m-1 from n+m-1 and n+m-2.
    int xw, xre, xrb;
};

// generate_laser_pulse, boundary.cpp:8-27 (the sin envelope, in amplitude)
__device__ __forceinline__ double d_envelope(const LaserArgs& L, double time, double position) {
    if (L.env_on && position == 0.0) return L.env_val;
    double t = time - position - L.t_start;
    double env = 0.0;
    if (t >= 0.0 && t <= 1.0 * L.dur) env = sin(L.pi * t / L.dur);
    return L.amp * env;
}

// Gpulse, boundary.cpp:29-40
__device__ __forceinline__ double d_gpulse(const LaserArgs& L, double time, double x, double y, double z) {
    double k = 2.0 * L.pi / L.l0;
    double zr = 0.5 * k * L.waist * L.waist;
    double xe = x - L.xfocus;
    double gouy = atan(xe / zr);
    double w = L.waist * sqrt(1 + pow(xe / zr, 2.0));
    double ph = (xe / zr) * (y * y + z * z) / w / w - gouy + k * xe - L.omega * time;
    return -1.0 * d_envelope(L, time, x) * L.waist / w * exp(-(y * y + z * z) / (w * w)) *
           cos(ph + L.phase);
}

// Source parameters of step 'step' for the context (k_field.hip).
LaserArgs laser_args(const ::pic_ctx* c, int step, int left, int right);

}  // namespace pic
