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
// Instead of scattering with atomics, each thread owns one face and gathers
// the particles of the 2x3x3 neighbouring cells from the cell-sorted arrays
// in a fixed order (cells k, j, i ascending; particles ascending inside a
// cell).  The sum is deterministic; it differs from the reference's
// ascending-particle-index summation order only by float rounding.
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

template <int COMP>
__global__ __launch_bounds__(256) void k_p2g(GridGeom g, const float *__restrict__ sx,
                                             const float *__restrict__ sy,
                                             const float *__restrict__ sz,
                                             const float *__restrict__ svel,
                                             const int *__restrict__ offset, float radius_f,
                                             double radius, float *out, uint8_t *valid) {
    const int ni = g.I + (COMP == 0), nj = g.J + (COMP == 1), nk = g.K + (COMP == 2);
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (size_t)ni * nj * nk) return;
    int i = (int)(tid % ni), j = (int)((tid / ni) % nj), k = (int)(tid / ((size_t)ni * nj));

    float hdx = (float)(0.5 * g.dx);  // velocityadvector.cpp:153-164
    F3 off = COMP == 0 ? f3(0.0f, hdx, hdx) : (COMP == 1 ? f3(hdx, 0.0f, hdx) : f3(hdx, hdx, 0.0f));
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

    // Cells whose particles can reach this face.
    int ci0 = COMP == 0 ? i - 1 : i - 1, ci1 = COMP == 0 ? i : i + 1;
    int cj0 = j - 1, cj1 = COMP == 1 ? j : j + 1;
    int ck0 = k - 1, ck1 = COMP == 2 ? k : k + 1;
    ci0 = max(ci0, 0); cj0 = max(cj0, 0); ck0 = max(ck0, 0);
    ci1 = min(ci1, g.I - 1); cj1 = min(cj1, g.J - 1); ck1 = min(ck1, g.K - 1);

    float sum = 0.0f, wsum = 0.0f;
    for (int ck = ck0; ck <= ck1; ck++) {
        for (int cj = cj0; cj <= cj1; cj++) {
            size_t row = (size_t)g.I * ((size_t)cj + (size_t)g.J * ck);
            int b = offset[row + ci0], e = offset[row + ci1 + 1];
            for (int q = b; q < e; q++) {
                F3 p = f3(sx[q], sy[q], sz[q]) - off;
                F3 pl = p - boff;
                F3 v = gpos - pl;
                float d2 = dot3(v, v);
                // A pair adds to the face only if it is within the radius and
                // the reference visits it (the particle is assigned to the
                // face's block and the face lies in the particle's clamped
                // local box).  The three tests are independent, so the cheap
                // distance test goes first; the sums are unchanged.
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
                sum += wt * svel[q];
                wsum += wt;
            }
        }
    }
    float u = sum;
    if (wsum > eps) u = sum / wsum;
    out[tid] = u;
    valid[tid] = wsum > eps ? 1 : 0;
}

hipError_t launch_p2g(const GridGeom &g, int comp, const float *sx, const float *sy,
                      const float *sz, const float *svel, const int *offset, double radius,
                      float *out, uint8_t *valid, hipStream_t s) {
    size_t n = (size_t)(g.I + (comp == 0)) * (g.J + (comp == 1)) * (g.K + (comp == 2));
    unsigned nb = (unsigned)((n + 255) / 256);
    if (comp == 0) k_p2g<0><<<nb, 256, 0, s>>>(g, sx, sy, sz, svel, offset, (float)radius, radius, out, valid);
    else if (comp == 1) k_p2g<1><<<nb, 256, 0, s>>>(g, sx, sy, sz, svel, offset, (float)radius, radius, out, valid);
    else k_p2g<2><<<nb, 256, 0, s>>>(g, sx, sy, sz, svel, offset, (float)radius, radius, out, valid);
    return hipGetLastError();
}

}  // namespace flipmi
