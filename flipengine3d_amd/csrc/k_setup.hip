// Static scene set-up kernels: the domain-wall solid SDF, the weight grid
// and the near-solid grid.  All three depend only on the domain box, so a
// scene computes them once.
#include "device_math.h"
#include "flipmi_kernels.h"

namespace flipmi {

// Nodal solid SDF of the domain walls.  Each node takes the minimum distance
// to the box triangles whose exact band covers it, evaluated in the frame of
// the 10^3 node block that owns it (meshlevelset.cpp:967-1000), is negated
// where the node is inside the box (meshlevelset.cpp:1298-1317) and the
// whole field negated (meshlevelset.cpp:702-708): positive on the fluid side.
__global__ void k_domain_solid(GridGeom g, BoxMesh mesh, float *phi) {
    const int bw = 10;
    int ni = g.I + 1, nj = g.J + 1, nk = g.K + 1;
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)ni * nj * nk;
    if (tid >= total) return;
    int i = (int)(tid % ni);
    int j = (int)((tid / ni) % nj);
    int k = (int)(tid / ((size_t)ni * nj));
    float upper = (float)((double)(ni + nj + nk) * g.dx);
    int bi = i / bw, bj = j / bw, bk = k / bw;
    double bdx = bw * g.dx;
    F3 boff = idx_to_pos(bi, bj, bk, bdx);
    F3 gpos = idx_to_pos(i - bi * bw, j - bj * bw, k - bk * bw, g.dx);
    float best = upper;
    for (int t = 0; t < mesh.ntri; t++) {
        const BoxTri &T = mesh.tri[t];
        if (i < T.gmin[0] || j < T.gmin[1] || k < T.gmin[2] || i > T.gmax[0] || j > T.gmax[1] ||
            k > T.gmax[2])
            continue;
        F3 v0 = f3(T.v[0][0], T.v[0][1], T.v[0][2]) - boff;
        F3 v1 = f3(T.v[1][0], T.v[1][1], T.v[1][2]) - boff;
        F3 v2 = f3(T.v[2][0], T.v[2][1], T.v[2][2]) - boff;
        F3 cp = closest_point_on_triangle(gpos, v0, v1, v2);
        float dist = len3(cp - gpos);
        if (dist < best) best = dist;
    }
    F3 np = idx_to_pos(i, j, k, g.dx);
    bool inside = np.x >= mesh.bpos[0] && np.y >= mesh.bpos[1] && np.z >= mesh.bpos[2] &&
                  (double)np.x < (double)mesh.bpos[0] + mesh.bext[0] &&
                  (double)np.y < (double)mesh.bpos[1] + mesh.bext[1] &&
                  (double)np.z < (double)mesh.bpos[2] + mesh.bext[2];
    phi[tid] = inside ? best : -best;
}

// LevelsetUtils::fractionInside (levelsetutils.cpp:39-141).
__device__ float fraction_inside2(float l, float r) {
    if (l < 0 && r < 0) return 1;
    if (l < 0 && r >= 0) return l / (l - r);
    if (l >= 0 && r < 0) return r / (r - l);
    return 0;
}
__device__ void cycle4(float *a) {
    float t = a[0];
    a[0] = a[1]; a[1] = a[2]; a[2] = a[3]; a[3] = t;
}
__device__ float fraction_inside4(float bl, float br, float tl, float tr) {
    int n = (bl < 0 ? 1 : 0) + (tl < 0 ? 1 : 0) + (br < 0 ? 1 : 0) + (tr < 0 ? 1 : 0);
    float L[4] = {bl, br, tr, tl};
    if (n == 4) return 1;
    if (n == 3) {
        while (L[0] < 0) cycle4(L);
        float s0 = 1 - fraction_inside2(L[0], L[3]);
        float s1 = 1 - fraction_inside2(L[0], L[1]);
        return 1.0f - 0.5f * s0 * s1;
    }
    if (n == 2) {
        while (L[0] >= 0 || !(L[1] < 0 || L[2] < 0)) cycle4(L);
        if (L[1] < 0) {
            float sl = fraction_inside2(L[0], L[3]);
            float sr = fraction_inside2(L[1], L[2]);
            return 0.5f * (sl + sr);
        }
        float mid = 0.25f * (L[0] + L[1] + L[2] + L[3]);
        if (mid < 0) {
            float area = 0;
            float s1 = 1 - fraction_inside2(L[0], L[3]);
            float s3 = 1 - fraction_inside2(L[2], L[3]);
            area += 0.5f * s1 * s3;
            float s2 = 1 - fraction_inside2(L[2], L[1]);
            float s0 = 1 - fraction_inside2(L[0], L[1]);
            area += 0.5f * s0 * s2;
            return 1.0f - area;
        }
        float area = 0;
        float s0 = fraction_inside2(L[0], L[1]);
        float s1 = fraction_inside2(L[0], L[3]);
        area += 0.5f * s0 * s1;
        float s2 = fraction_inside2(L[2], L[1]);
        float s3 = fraction_inside2(L[2], L[3]);
        area += 0.5f * s2 * s3;
        return area;
    }
    if (n == 1) {
        while (L[0] >= 0) cycle4(L);
        float s0 = fraction_inside2(L[0], L[3]);
        float s1 = fraction_inside2(L[0], L[1]);
        return 0.5f * s0 * s1;
    }
    return 0;
}

// Tetrahedral volume fractions (levelsetutils.h:62-102, levelsetutils.cpp:213-259).
__device__ __forceinline__ void swapf(float &a, float &b) { float t = a; a = b; b = t; }
__device__ float volume_fraction4(float p0, float p1, float p2, float p3) {
    if (p0 > p1) swapf(p0, p1);
    if (p2 > p3) swapf(p2, p3);
    if (p0 > p2) swapf(p0, p2);
    if (p1 > p3) swapf(p1, p3);
    if (p1 > p2) swapf(p1, p2);
    if (p3 <= 0) return 1;
    if (p2 <= 0) return 1 - p3 * p3 * p3 / ((p3 - p2) * (p3 - p1) * (p3 - p0));
    if (p1 <= 0) {
        float a = p0 / (p0 - p2), b = p0 / (p0 - p3), c = p1 / (p1 - p3), d = p1 / (p1 - p2);
        return a * b * (1 - d) + b * (1 - c) * d + c * d;
    }
    if (p0 <= 0) return p0 * p0 * p0 / ((p0 - p1) * (p0 - p2) * (p0 - p3));
    return 0;
}
__device__ float volume_fraction8(float p000, float p100, float p010, float p110, float p001,
                                  float p101, float p011, float p111) {
    return (volume_fraction4(p000, p001, p101, p011) + volume_fraction4(p000, p101, p100, p110) +
            volume_fraction4(p000, p010, p011, p110) + volume_fraction4(p101, p011, p111, p110) +
            2 * volume_fraction4(p000, p011, p101, p110) +
            volume_fraction4(p100, p101, p001, p111) + volume_fraction4(p100, p001, p000, p010) +
            volume_fraction4(p100, p110, p111, p010) + volume_fraction4(p001, p111, p011, p010) +
            2 * volume_fraction4(p100, p111, p001, p010)) /
           12.0f;
}

__device__ __forceinline__ float clamp01f(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

// Weight grid (FluidSimulation::_updateWeightGridThread, fluidsimulation.cpp:
// 3688-3730): open fractions of faces and cells, 1 - solid fraction.
__global__ void k_weights(GridGeom g, const float *__restrict__ phi, float *wU, float *wV,
                          float *wW, float *wC) {
    Dims N{g.I + 1, g.J + 1, g.K + 1};
    size_t nU = (size_t)(g.I + 1) * g.J * g.K, nV = (size_t)g.I * (g.J + 1) * g.K;
    size_t nW = (size_t)g.I * g.J * (g.K + 1), nC = (size_t)g.I * g.J * g.K;
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    auto P = [&](int i, int j, int k) { return phi[N.at(i, j, k)]; };
    if (tid < nU) {
        int ni = g.I + 1;
        int i = (int)(tid % ni), j = (int)((tid / ni) % g.J), k = (int)(tid / ((size_t)ni * g.J));
        float f = fraction_inside4(P(i, j, k), P(i, j + 1, k), P(i, j, k + 1), P(i, j + 1, k + 1));
        wU[tid] = clamp01f(1.0f - f);
        return;
    }
    tid -= nU;
    if (tid < nV) {
        int nj = g.J + 1;
        int i = (int)(tid % g.I), j = (int)((tid / g.I) % nj), k = (int)(tid / ((size_t)g.I * nj));
        float f = fraction_inside4(P(i, j, k), P(i, j, k + 1), P(i + 1, j, k), P(i + 1, j, k + 1));
        wV[tid] = clamp01f(1.0f - f);
        return;
    }
    tid -= nV;
    if (tid < nW) {
        int i = (int)(tid % g.I), j = (int)((tid / g.I) % g.J), k = (int)(tid / ((size_t)g.I * g.J));
        float f = fraction_inside4(P(i, j, k), P(i, j + 1, k), P(i + 1, j, k), P(i + 1, j + 1, k));
        wW[tid] = clamp01f(1.0f - f);
        return;
    }
    tid -= nW;
    if (tid < nC) {
        int i = (int)(tid % g.I), j = (int)((tid / g.I) % g.J), k = (int)(tid / ((size_t)g.I * g.J));
        float p000 = P(i, j, k), p001 = P(i, j, k + 1), p010 = P(i, j + 1, k),
              p011 = P(i, j + 1, k + 1), p100 = P(i + 1, j, k), p101 = P(i + 1, j, k + 1),
              p110 = P(i + 1, j + 1, k), p111 = P(i + 1, j + 1, k + 1);
        float cw;
        if (p000 < 0 && p001 < 0 && p010 < 0 && p011 < 0 && p100 < 0 && p101 < 0 && p110 < 0 &&
            p111 < 0)
            cw = 1.0f;
        else if (p000 >= 0 && p001 >= 0 && p010 >= 0 && p011 >= 0 && p100 >= 0 && p101 >= 0 &&
                 p110 >= 0 && p111 >= 0)
            cw = 0.0f;
        else
            cw = volume_fraction8(p000, p100, p010, p110, p001, p101, p011, p111);
        wC[tid] = clamp01f(1.0f - cw);
    }
}

// Near-solid coarse grid (FluidSimulation::_updateNearSolidGrid,
// fluidsimulation.cpp:3081-3126): coarse cells (3 dx) holding a fine cell
// whose nodal solid phi is within 3 dx, feathered over 6-neighbours.
__global__ void k_near_solid_mark(GridGeom g, const float *__restrict__ phi, uint8_t *near,
                                  int ci, int cj, int ck) {
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nC = (size_t)g.I * g.J * g.K;
    if (tid >= nC) return;
    int i = (int)(tid % g.I), j = (int)((tid / g.I) % g.J), k = (int)(tid / ((size_t)g.I * g.J));
    Dims N{g.I + 1, g.J + 1, g.K + 1};
    float maxd = (float)(3 * g.dx);
    if (fabsf(phi[N.at(i, j, k)]) < maxd) {
        Dims C{ci, cj, ck};
        near[C.at(i / 3, j / 3, k / 3)] = 1;
    }
}

__global__ void k_feather6(const uint8_t *__restrict__ src, uint8_t *dst, int ci, int cj, int ck) {
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    Dims C{ci, cj, ck};
    if (tid >= C.count()) return;
    int i = (int)(tid % ci), j = (int)((tid / ci) % cj), k = (int)(tid / ((size_t)ci * cj));
    uint8_t v = src[tid];
    if (!v) {
        v = (C.in(i - 1, j, k) && src[C.at(i - 1, j, k)]) || (C.in(i + 1, j, k) && src[C.at(i + 1, j, k)]) ||
            (C.in(i, j - 1, k) && src[C.at(i, j - 1, k)]) || (C.in(i, j + 1, k) && src[C.at(i, j + 1, k)]) ||
            (C.in(i, j, k - 1) && src[C.at(i, j, k - 1)]) || (C.in(i, j, k + 1) && src[C.at(i, j, k + 1)]);
    }
    dst[tid] = v;
}

static inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_domain_solid(const GridGeom &g, const BoxMesh &mesh, float *phi_nodal,
                               hipStream_t s) {
    size_t n = (size_t)(g.I + 1) * (g.J + 1) * (g.K + 1);
    k_domain_solid<<<nblk(n, 256), 256, 0, s>>>(g, mesh, phi_nodal);
    return hipGetLastError();
}

hipError_t launch_weights(const GridGeom &g, const float *phi_nodal, float *wU, float *wV,
                          float *wW, float *wC, hipStream_t s) {
    size_t n = (size_t)(g.I + 1) * g.J * g.K + (size_t)g.I * (g.J + 1) * g.K +
               (size_t)g.I * g.J * (g.K + 1) + (size_t)g.I * g.J * g.K;
    k_weights<<<nblk(n, 256), 256, 0, s>>>(g, phi_nodal, wU, wV, wW, wC);
    return hipGetLastError();
}

hipError_t launch_near_solid(const GridGeom &g, const float *phi_nodal, uint8_t *near_a,
                             uint8_t *near_b, int ci, int cj, int ck, int feather_layers,
                             uint8_t **result, hipStream_t s) {
    size_t nc = (size_t)ci * cj * ck;
    hipError_t e = launch_fill_u8(near_a, nc, 0, s);
    if (e != hipSuccess) return e;
    size_t n = (size_t)g.I * g.J * g.K;
    k_near_solid_mark<<<nblk(n, 256), 256, 0, s>>>(g, phi_nodal, near_a, ci, cj, ck);
    uint8_t *src = near_a, *dst = near_b;
    for (int l = 0; l < feather_layers; l++) {
        k_feather6<<<nblk(nc, 256), 256, 0, s>>>(src, dst, ci, cj, ck);
        uint8_t *t = src; src = dst; dst = t;
    }
    *result = src;
    return hipGetLastError();
}

}  // namespace flipmi
