// push64.h -- FP64 particle/field views and push constants shared by the
// forward push (k_push64.hip) and its adjoint (k_adj.hip).
#pragma once

#include "pic_internal.h"
#include "spin.h"

namespace pic {

// Dead slots (removed at an open boundary, not yet compacted) carry a NaN
// weight and a negative id.
struct P64 {
    double *x, *y, *z, *ux, *uy, *uz;
    const double* w;
    double *xo, *yo, *zo;  // may be null
    int64_t* cnt;
    double *sx, *sy, *sz;  // spin (species with precession only)
};

struct Fields64 {
    const double *ex, *ey, *ez, *bx, *by, *bz;
    double *jx, *jy, *jz, *rho;
};

struct PushConsts {
    double q, qm, dt, dx, dy, dz, cx, cy, cz;  // cx = q*dx/dt (times w per particle)
    double N[3];
    int per[3];
    SpinConsts spin;
    // int64 fixed-point units of the tiled push's LDS deposit (jx, jy, jz,
    // rho): 2^-48 of one particle's largest contribution, |q| w_max d/dt
    // (|q| w_max for rho), and their inverses
    double unit[4], inv_unit[4];
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

PushConsts push_consts64(const pic_ctx* c, const Species& S);

}  // namespace pic
