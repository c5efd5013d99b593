// Device-side arithmetic shared by the FLIP kernels.
//
// The helpers reproduce the reference engine's float/double mixing exactly
// (jklae/FLIPEngine3D src/engine/vmath.h:38-97, vmath.cpp:33-123 for the
// float 3-vector; grid3d.h:52-109 for index <-> position conversion).  All
// kernels are compiled with -ffp-contract=off so no multiply-add is fused
// behind the source's back; float sqrt/div and double sqrt/div lower to
// correctly rounded sequences on gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flipmi {

struct F3 {
    float x, y, z;
};

__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator*(float s, F3 v) { return F3{v.x * s, v.y * s, v.z * s}; }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// vmath::length: sqrt of the float dot product, correctly rounded to float.
__device__ __forceinline__ float len3(F3 v) { return sqrtf(dot3(v, v)); }
// vmath operator/: multiply by a float reciprocal formed in double.
__device__ __forceinline__ F3 div3(F3 v, float s) {
    float inv = (float)(1.0 / (double)s);
    return F3{v.x * inv, v.y * inv, v.z * inv};
}
__device__ __forceinline__ F3 normalize3(F3 v) { return div3(v, len3(v)); }

struct I3 {
    int i, j, k;
};

// (i, j, k) of flat index q of an ni x nj x nk array, i fastest, in 32-bit
// arithmetic: two 32-bit divisions instead of three 64-bit ones, which cost
// several times the instructions.  Grids hold fewer than 2^32 elements
// (make_geom checks).
__device__ __forceinline__ void unflatten3(size_t q, int ni, int nj, int &i, int &j, int &k) {
    const unsigned u = (unsigned)q, a = (unsigned)ni, b = (unsigned)nj;
    const unsigned r = u / a, kk = r / b;
    i = (int)(u - r * a);
    j = (int)(r - kk * b);
    k = (int)kk;
}

// Grid3d::positionToGridIndex: floor(x * (1/dx)) in double.
__device__ __forceinline__ I3 pos_to_index(double x, double y, double z, double invdx) {
    return I3{(int)floor(x * invdx), (int)floor(y * invdx), (int)floor(z * invdx)};
}
__device__ __forceinline__ I3 pos_to_index(F3 p, double invdx) {
    return pos_to_index((double)p.x, (double)p.y, (double)p.z, invdx);
}
// Grid3d::GridIndexToPosition (vec3 form): ((float)i * dx) rounded to float.
__device__ __forceinline__ float idx_to_pos1(int i, double dx) { return (float)((double)(float)i * dx); }
__device__ __forceinline__ F3 idx_to_pos(int i, int j, int k, double dx) {
    return F3{idx_to_pos1(i, dx), idx_to_pos1(j, dx), idx_to_pos1(k, dx)};
}
// Grid3d::GridIndexToCellCenter (vec3 form).
__device__ __forceinline__ float idx_to_center1(int i, double dx) {
    return (float)((double)(float)i * dx + 0.5 * dx);
}

// Order-preserving map from float to uint32 (for LDS atomic minimum).
__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Dense grid dimensions and flat index (Array3d, i fastest: array3d.h:425-428).
struct Dims {
    int ni, nj, nk;
    __device__ __forceinline__ bool in(int i, int j, int k) const {
        return i >= 0 && j >= 0 && k >= 0 && i < ni && j < nj && k < nk;
    }
    __device__ __forceinline__ size_t at(int i, int j, int k) const {
        return (size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * (size_t)k);
    }
    __host__ __device__ __forceinline__ size_t count() const { return (size_t)ni * nj * nk; }
};

// Interpolation::trilinearInterpolate(p[8], x, y, z) (interpolation.cpp:62-71).
__device__ __forceinline__ double trilerp8(const double *p, double x, double y, double z) {
    return p[0] * (1 - x) * (1 - y) * (1 - z) + p[1] * x * (1 - y) * (1 - z) +
           p[2] * (1 - x) * y * (1 - z) + p[3] * (1 - x) * (1 - y) * z +
           p[4] * x * (1 - y) * z + p[5] * (1 - x) * y * z + p[6] * x * y * (1 - z) +
           p[7] * x * y * z;
}

// Collision::findClosestPointOnTriangle (collision.cpp:161-270), in the
// reference's float/double mix.
__device__ inline F3 closest_point_on_triangle(F3 p0, F3 v0, F3 v1, F3 v2) {
    F3 edge0 = v1 - v0, edge1 = v2 - v0, pv = v0 - p0;
    double a = dot3(edge0, edge0), b = dot3(edge0, edge1), c = dot3(edge1, edge1);
    double d = dot3(edge0, pv), e = dot3(edge1, pv);
    double det = a * c - b * b;
    if (fabs(det) < 1e-9) {
        float feps = 1e-6f;
        float len1 = len3(v0 - v1), len2 = len3(v1 - v2), len3v = len3(v2 - v0);
        if (len1 < feps && len2 < feps && len3v < feps) return v0;
        F3 dp, dq;
        if (len1 < feps) { dp = v0; dq = v2; }
        else if (len2 < feps) { dp = v1; dq = v0; }
        else { dp = v2; dq = v1; }
        float num = dot3(p0 - dp, dq - dp), den = dot3(dq - dp, dq - dp);
        if (fabsf(den) < feps) return v0;
        float lam = num / den;
        if (lam <= 0.0f) return dp;
        if (lam >= 1.0f) return dq;
        return dp + lam * (dq - dp);
    }
    double s = b * e - c * d, t = b * d - a * e;
    auto clamp01 = [](double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); };
    if (s + t < det) {
        if (s < 0.0) {
            if (t < 0.0) {
                if (d < 0.0) { s = clamp01(-d / a); t = 0.0; }
                else { s = 0.0; t = clamp01(-e / c); }
            } else { s = 0.0; t = clamp01(-e / c); }
        } else if (t < 0.0) { s = clamp01(-d / a); t = 0.0; }
        else { double inv = 1.0 / det; s *= inv; t *= inv; }
    } else {
        if (s < 0.0) {
            double tmp0 = b + d, tmp1 = c + e;
            if (tmp1 > tmp0) { s = clamp01((tmp1 - tmp0) / (a - 2 * b + c)); t = 1 - s; }
            else { t = clamp01(-e / c); s = 0.0; }
        } else if (t < 0.0) {
            if (a + d > b + e) { s = clamp01((c + e - b - d) / (a - 2 * b + c)); t = 1 - s; }
            else { s = clamp01(-e / c); t = 0.0; }
        } else {
            s = clamp01((c + e - b - d) / (a - 2 * b + c)); t = 1.0 - s;
        }
    }
    return v0 + (float)s * edge0 + (float)t * edge1;
}

// Interpolation::trilinearInterpolate(vec3, dx, Array3d<float>)
// (interpolation.cpp:73-112) on a nodal grid of dims N: nodes outside the
// grid read as 0.
__device__ inline double nodal_trilinear(const float *__restrict__ phi, Dims N, double dx,
                                         double invdx, F3 p) {
    I3 c = pos_to_index(p, invdx);
    F3 gp = idx_to_pos(c.i, c.j, c.k, dx);
    double ix = (double)(p.x - gp.x) * invdx, iy = (double)(p.y - gp.y) * invdx,
           iz = (double)(p.z - gp.z) * invdx;
    double pts[8];
    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                           {1, 0, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 1}};
#pragma unroll
    for (int q = 0; q < 8; q++) {
        int a = c.i + off[q][0], b = c.j + off[q][1], d = c.k + off[q][2];
        pts[q] = N.in(a, b, d) ? (double)phi[N.at(a, b, d)] : 0.0;
    }
    return (double)(float)trilerp8(pts, ix, iy, iz);
}

// SplitMix64 and a keyed uniform draw in [0, 1): the generator behind the
// inflow jitter.  The reference draws rand() (fluidsimulation.h:1501-1503)
// from several threads at once; here draw (seed, stream, item) is a pure
// function of what it jitters, so results do not depend on scheduling.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ double draw_unit(uint64_t seed, uint64_t stream, uint64_t item) {
    uint64_t h = splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + item);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace flipmi
