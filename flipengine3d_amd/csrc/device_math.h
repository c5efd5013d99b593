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
    __device__ __forceinline__ size_t count() const { return (size_t)ni * nj * nk; }
};

// Interpolation::trilinearInterpolate(p[8], x, y, z) (interpolation.cpp:62-71).
__device__ __forceinline__ double trilerp8(const double *p, double x, double y, double z) {
    return p[0] * (1 - x) * (1 - y) * (1 - z) + p[1] * x * (1 - y) * (1 - z) +
           p[2] * (1 - x) * y * (1 - z) + p[3] * (1 - x) * (1 - y) * z +
           p[4] * x * (1 - y) * z + p[5] * (1 - x) * y * z + p[6] * x * y * (1 - z) +
           p[7] * x * y * z;
}

}  // namespace flipmi
