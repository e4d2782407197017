/*
 * Seeded initial-state generator shared by the host loader, the device
 * loader kernel and the CPU oracle, so all three see identical particles.
 *
 * The reference loads a deterministic lattice with constant momenta
 * (src/particle.cpp:50-111); the benchmark configurations need random
 * in-cell positions and Maxwellian momenta instead, so this is new code.
 *
 * RNG: splitmix64 (Steele, Lea, Flood 2014, "Fast splittable pseudorandom
 * number generators"), used as a counter-based hash of (seed, id, draw).
 * Normals: Box-Muller.  Positions use only integer -> double conversion and
 * one add, so they are bit-identical on host and device; momenta go through
 * log/cos/sin and agree to a few ulps.
 *
 * The including file defines PIC_LD_FN (function qualifiers) before
 * including this header.
 */
#ifndef PIC_LOADER_IMPL_H
#define PIC_LOADER_IMPL_H

#include <math.h>
#include <stdint.h>

#ifndef PIC_LD_FN
#define PIC_LD_FN static inline
#endif

PIC_LD_FN uint64_t pic_splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* 64 random bits for draw 'draw' of particle 'id' under 'seed'. */
PIC_LD_FN uint64_t pic_rng_bits(uint64_t seed, uint64_t id, uint32_t draw) {
    uint64_t h = pic_splitmix64(seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(draw + 1u)));
    return pic_splitmix64(h ^ id);
}

/* Uniform double in [0, 1) with 53 random bits. */
PIC_LD_FN double pic_rng_u01(uint64_t seed, uint64_t id, uint32_t draw) {
    return (double)(pic_rng_bits(seed, id, draw) >> 11) * (1.0 / 9007199254740992.0);
}

/* Density (in units of the loader's 'density') at x-cell centre xc. */
PIC_LD_FN double pic_profile_density(int profile, double density, double x0, double x1,
                                     double scale_len, double cutoff, double xc) {
    if (profile == 0) return density;
    if (profile == 1) return (xc >= x0 && xc < x1) ? density : 0.0;
    /* ramp + target */
    if (xc >= x0 && xc < x1) return density;
    if (xc < x0) {
        double n = density * exp((xc - x0) / scale_len);
        return n >= cutoff ? n : 0.0;
    }
    return 0.0;
}

/* Particles in every cell of x-column ix (the profile varies along x only). */
PIC_LD_FN int32_t pic_profile_count(int profile, int32_t ppc, double density, double x0,
                                    double x1, double scale_len, double cutoff, int32_t ix) {
    double n = pic_profile_density(profile, density, x0, x1, scale_len, cutoff, ix + 0.5);
    if (n <= 0.0) return 0;
    return (int32_t)(ppc * (n / density) + 0.5);
}

/* Position offsets and momenta of one particle with the given id inside
 * cell (ix, iy, iz).  Draws 0-2: in-cell offsets; 3-4: Box-Muller pair for
 * ux, uy; 5-6: Box-Muller for uz. */
PIC_LD_FN void pic_load_one(uint64_t seed, uint64_t id, int32_t ix, int32_t iy, int32_t iz,
                            double u_th, double* x, double* y, double* z, double* ux,
                            double* uy, double* uz) {
    const double two_pi = 6.283185307179586;
    *x = (double)ix + pic_rng_u01(seed, id, 0);
    *y = (double)iy + pic_rng_u01(seed, id, 1);
    *z = (double)iz + pic_rng_u01(seed, id, 2);
    double a = 1.0 - pic_rng_u01(seed, id, 3); /* (0, 1] */
    double b = pic_rng_u01(seed, id, 4);
    double r = sqrt(-2.0 * log(a));
    *ux = u_th * (r * cos(two_pi * b));
    *uy = u_th * (r * sin(two_pi * b));
    double c = 1.0 - pic_rng_u01(seed, id, 5);
    double d = pic_rng_u01(seed, id, 6);
    *uz = u_th * (sqrt(-2.0 * log(c)) * cos(two_pi * d));
}

/* Locate particle p: particles are ordered by z, then y, then x-column, with
 * prefix[ix] = number of particles per (y,z) row before column ix
 * (prefix has nx+1 entries, prefix[nx] = row total). */
PIC_LD_FN void pic_locate(int64_t p, const int64_t* prefix, int32_t nx, int32_t ny,
                          int32_t* ix, int32_t* iy, int32_t* iz) {
    int64_t row_total = prefix[nx];
    int64_t row = p / row_total;
    int64_t r = p - row * row_total;
    *iy = (int32_t)(row % ny);
    *iz = (int32_t)(row / ny);
    /* largest ix with prefix[ix] <= r */
    int32_t lo = 0, hi = nx - 1;
    while (lo < hi) {
        int32_t mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= r) lo = mid; else hi = mid - 1;
    }
    *ix = lo;
}

#endif /* PIC_LOADER_IMPL_H */
