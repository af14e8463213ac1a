// Device arithmetic shared by the gfx950 kernels: Philox4x32-10, the
// project's own sin/cos of 2*pi*u, Moeller-Trumbore and the cosine-weighted
// hemisphere sample.  All of it compiles with -ffp-contract=off so every
// expression rounds exactly as written; the CPU oracle restates the same
// steps and is compared bit-for-bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace prd {

// ---------------------------------------------------------------- Philox --
// Philox4x32 with 10 rounds (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC 2011): multipliers 0xD2511F53 and
// 0xCD9E8D57, Weyl key increments 0x9E3779B9 and 0xBB67AE85.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Top 24 bits as a float in [0, 1): exact.
__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// sin and cos of 2*pi*u for u in [0, 1): quadrant from 4u (exact), then
// Taylor polynomials of degree 11 (sin) and 12 (cos) on [0, pi/2) in Horner
// form; absolute error below 1e-7.
__device__ __forceinline__ void sincos2pi(float u, float &s, float &c) {
    const float x = u * 4.0f;
    const int q = (int)x;
    const float r = x - (float)q;
    const float a = r * 1.57079632679489661923f;
    const float a2 = a * a;
    float sp = -2.50521083854417188e-08f;
    sp = sp * a2 + 2.75573192239858907e-06f;
    sp = sp * a2 + -1.98412698412698413e-04f;
    sp = sp * a2 + 8.33333333333333333e-03f;
    sp = sp * a2 + -1.66666666666666667e-01f;
    const float sa = a + a * (a2 * sp);
    float cp = 2.08767569878680990e-09f;
    cp = cp * a2 + -2.75573192239858907e-07f;
    cp = cp * a2 + 2.48015873015873016e-05f;
    cp = cp * a2 + -1.38888888888888889e-03f;
    cp = cp * a2 + 4.16666666666666667e-02f;
    cp = cp * a2 + -0.5f;
    const float ca = 1.0f + a2 * cp;
    switch (q & 3) {
        case 0: s = sa; c = ca; break;
        case 1: s = ca; c = -sa; break;
        case 2: s = -sa; c = -ca; break;
        default: s = -ca; c = sa; break;
    }
}

// Reference normalisation (geometry.h:28-34).
__device__ __forceinline__ void normalize3(float &x, float &y, float &z) {
    const float n2 = x * x + y * y + z * z;
    if (n2 > 0.f) {
        const float inv = 1.0f / sqrtf(n2);
        x *= inv;
        y *= inv;
        z *= inv;
    }
}

// Moeller-Trumbore in the reference's operation order (geometry.h:241-262).
// Returns false on a cull or a barycentric miss; t is only written on true.
__device__ __forceinline__ bool mt_hit(float ox, float oy, float oz, float dx, float dy, float dz, float ax, float ay,
                                       float az, float abx, float aby, float abz, float acx, float acy, float acz,
                                       float eps, float &t) {
    const float px = dy * acz - dz * acy;
    const float py = dz * acx - dx * acz;
    const float pz = dx * acy - dy * acx;
    const float det = abx * px + aby * py + abz * pz;
    if (det < eps) return false;
    const float inv = 1.0f / det;
    const float tx = ox - ax, ty = oy - ay, tz = oz - az;
    const float u = (tx * px + ty * py + tz * pz) * inv;
    if (u < 0.f || u > 1.f) return false;
    const float qx = ty * abz - tz * aby;
    const float qy = tz * abx - tx * abz;
    const float qz = tx * aby - ty * abx;
    const float v = (dx * qx + dy * qy + dz * qz) * inv;
    if (v < 0.f || u + v > 1.f) return false;
    t = (acx * qx + acy * qy + acz * qz) * inv;
    return true;
}

// Cosine-weighted direction about the unit normal n (reference
// geometry.cpp:4-29): frame from h = n with its smallest-magnitude component
// set to 1, x = norm(h x n), z = norm(x x n); sample
// (sin(theta) cos(phi), cos(theta), sin(theta) sin(phi)) with
// cos(theta) = sqrt(1 - u1), sin(theta) = sqrt(u1), phi = 2 pi u2; the result
// is normalised, then normalised again as the Ray constructor does.
__device__ __forceinline__ void sample_hemisphere(float nx, float ny, float nz, float u1, float u2, float &dx,
                                                  float &dy, float &dz) {
    float hx = nx, hy = ny, hz = nz;
    const float fx = fabsf(nx), fy = fabsf(ny), fz = fabsf(nz);
    if (fx <= fy && fx <= fz)
        hx = 1.f;
    else if (fy <= fx && fy <= fz)
        hy = 1.f;
    else
        hz = 1.f;
    float xx = hy * nz - hz * ny, xy = hz * nx - hx * nz, xz = hx * ny - hy * nx;
    normalize3(xx, xy, xz);
    float zx = xy * nz - xz * ny, zy = xz * nx - xx * nz, zz = xx * ny - xy * nx;
    normalize3(zx, zy, zz);
    const float st = sqrtf(u1), ct = sqrtf(1.0f - u1);
    float sp, cp;
    sincos2pi(u2, sp, cp);
    const float sx = st * cp, sy = ct, sz = st * sp;
    dx = xx * sx + nx * sy + zx * sz;
    dy = xy * sx + ny * sy + zy * sz;
    dz = xz * sx + nz * sy + zz * sz;
    normalize3(dx, dy, dz);
    normalize3(dx, dy, dz);
}

}  // namespace prd
