// Particle-to-grid transfer onto one staggered MAC component.
//
// Reference: VelocityAdvector::advect / _advectGrid / _advectionProducerThread
// (velocityadvector.cpp:38-459).  For every face f of component U, V or W
//     u_f = sum_p w(|x_f - p|) v_p / sum_p w,   w(d) = 1 - 4/9 d^6/r^6
//                                                    + 17/9 d^4/r^4 - 22/9 d^2/r^2
// over particles within r = sqrt(3)/2 dx, with the face valid when the weight
// sum exceeds 1e-6.  The reference evaluates distances in the frame of the
// 10^3 face block that owns f and assigns particles to blocks by their search
// box; both are replayed here, so every per-pair weight is bit-identical.
//
// Instead of scattering with atomics, each thread owns the U, V and W faces
// of one index and gathers the particles of the neighbouring cells from the
// cell-sorted arrays in a fixed order (cells k, j, i ascending; particles
// ascending inside a cell).  The sum is deterministic; it differs from the
// reference's ascending-particle-index summation order only by float
// rounding.
#include "device_math.h"
#include "flipmi_kernels.h"

namespace flipmi {

constexpr int kChunk = 10;  // velocityadvector.h:163 (_chunkWidth)

__device__ __forceinline__ bool p2g_point_in_block(F3 p, float sr, float blockdx,
                                                   double inv_blockdx, int bi, int bj, int bk) {
    I3 b = pos_to_index(p, inv_blockdx);
    F3 bp = idx_to_pos(b.i, b.j, b.k, (double)blockdx);
    if (p.x - sr > bp.x && p.y - sr > bp.y && p.z - sr > bp.z && p.x + sr < bp.x + blockdx &&
        p.y + sr < bp.y + blockdx && p.z + sr < bp.z + blockdx) {
        return b.i == bi && b.j == bj && b.k == bk;
    }
    I3 lo = pos_to_index((double)(p.x - sr), (double)(p.y - sr), (double)(p.z - sr), inv_blockdx);
    I3 hi = pos_to_index((double)(p.x + sr), (double)(p.y + sr), (double)(p.z + sr), inv_blockdx);
    return bi >= lo.i && bi <= hi.i && bj >= lo.j && bj <= hi.j && bk >= lo.k && bk <= hi.k;
}

// All three components in one pass: thread (i, j, k) owns the U, V and W
// faces of that index (those that exist) and walks the 3x3x3 cells around
// it once.  A particle from a cell outside a component's 2x3x3 set lies at
// least dx > r from that face, so it fails the distance test: every face
// sums exactly the particles, in exactly the order, of the per-component
// kernel.
__global__ __launch_bounds__(256) void k_p2g3(GridGeom g, const float *__restrict__ sx,
                                              const float *__restrict__ sy,
                                              const float *__restrict__ sz,
                                              const float *__restrict__ su,
                                              const float *__restrict__ sv,
                                              const float *__restrict__ sw,
                                              const int *__restrict__ offset, float radius_f,
                                              double radius, float *ou, float *ov, float *ow,
                                              uint8_t *vu, uint8_t *vv, uint8_t *vw) {
    const int ni = g.I + 1, nj = g.J + 1, nk = g.K + 1;
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (size_t)ni * nj * nk) return;
    int i = (int)(tid % ni), j = (int)((tid / ni) % nj), k = (int)(tid / ((size_t)ni * nj));
    const bool has[3] = {j < g.J && k < g.K, i < g.I && k < g.K, i < g.I && j < g.J};
    if (!has[0] && !has[1] && !has[2]) return;

    float hdx = (float)(0.5 * g.dx);  // velocityadvector.cpp:153-164
    const F3 off[3] = {f3(0.0f, hdx, hdx), f3(hdx, 0.0f, hdx), f3(hdx, hdx, 0.0f)};
    float eps = 1e-6f;
    float sr = (float)(radius + (double)eps);
    double chunkdx = g.dx * kChunk;
    float blockdx = (float)chunkdx;
    double inv_blockdx = 1.0 / (double)blockdx;
    float r = radius_f, rsq = r * r;
    float coef1 = (4.0f / 9.0f) * (1.0f / (r * r * r * r * r * r));
    float coef2 = (17.0f / 9.0f) * (1.0f / (r * r * r * r));
    float coef3 = (22.0f / 9.0f) * (1.0f / (r * r));

    int bi = i / kChunk, bj = j / kChunk, bk = k / kChunk;
    int li = i - bi * kChunk, lj = j - bj * kChunk, lk = k - bk * kChunk;
    F3 boff = idx_to_pos(bi, bj, bk, kChunk * g.dx);
    F3 gpos = idx_to_pos(li, lj, lk, g.dx);

    int ci0 = max(i - 1, 0), ci1 = min(i + 1, g.I - 1);
    int cj0 = max(j - 1, 0), cj1 = min(j + 1, g.J - 1);
    int ck0 = max(k - 1, 0), ck1 = min(k + 1, g.K - 1);
    const float *vel[3] = {su, sv, sw};

    float sum[3] = {0.0f, 0.0f, 0.0f}, wsum[3] = {0.0f, 0.0f, 0.0f};
    for (int ck = ck0; ck <= ck1; ck++) {
        for (int cj = cj0; cj <= cj1; cj++) {
            size_t row = (size_t)g.I * ((size_t)cj + (size_t)g.J * ck);
            int b = offset[row + ci0], e = offset[row + ci1 + 1];
            for (int q = b; q < e; q++) {
                const F3 p0 = f3(sx[q], sy[q], sz[q]);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    if (!has[c]) continue;
                    F3 p = p0 - off[c];
                    F3 pl = p - boff;
                    F3 v = gpos - pl;
                    float d2 = dot3(v, v);
                    if (!(d2 < rsq)) continue;
                    if (!p2g_point_in_block(p, sr, blockdx, inv_blockdx, bi, bj, bk)) continue;
                    I3 gmin = pos_to_index((double)(pl.x - sr), (double)(pl.y - sr), (double)(pl.z - sr), g.invdx);
                    I3 gmax = pos_to_index((double)(pl.x + sr), (double)(pl.y + sr), (double)(pl.z + sr), g.invdx);
                    gmin.i = max(gmin.i, 0); gmin.j = max(gmin.j, 0); gmin.k = max(gmin.k, 0);
                    gmax.i = min(gmax.i, kChunk - 1); gmax.j = min(gmax.j, kChunk - 1);
                    gmax.k = min(gmax.k, kChunk - 1);
                    if (li < gmin.i || li > gmax.i || lj < gmin.j || lj > gmax.j || lk < gmin.k ||
                        lk > gmax.k)
                        continue;
                    float wt = 1.0f - coef1 * d2 * d2 * d2 + coef2 * d2 * d2 - coef3 * d2;
                    sum[c] += wt * vel[c][q];
                    wsum[c] += wt;
                }
            }
        }
    }
    float *outs[3] = {ou, ov, ow};
    uint8_t *valids[3] = {vu, vv, vw};
    const int cni[3] = {g.I + 1, g.I, g.I}, cnj[3] = {g.J, g.J + 1, g.J};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (!has[c]) continue;
        size_t f = (size_t)i + (size_t)cni[c] * ((size_t)j + (size_t)cnj[c] * k);
        float u = sum[c];
        if (wsum[c] > eps) u = sum[c] / wsum[c];
        outs[c][f] = u;
        valids[c][f] = wsum[c] > eps ? 1 : 0;
    }
}

hipError_t launch_p2g3(const GridGeom &g, const float *sx, const float *sy, const float *sz,
                       const float *const svel[3], const int *offset, double radius, float *const out[3],
                       uint8_t *const valid[3], hipStream_t s) {
    size_t n = (size_t)(g.I + 1) * (g.J + 1) * (g.K + 1);
    k_p2g3<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, sx, sy, sz, svel[0], svel[1], svel[2], offset,
                                                       (float)radius, radius, out[0], out[1], out[2],
                                                       valid[0], valid[1], valid[2]);
    return hipGetLastError();
}

}  // namespace flipmi
