// Marker/cell classification: the liquid signed distance field and the
// fluid/air labels it defines (phi < 0 is fluid).
//
// Reference: ParticleLevelSet::calculateSignedDistanceField and
// postProcessSignedDistanceField (particlelevelset.cpp:161-196, 299-630).
// The reference bins particles into 10^3 cell blocks and takes, per block,
//     phi(c) = min(3 dx, min_p |c_local - p_local| - r)
// in block-local float coordinates.  Here one workgroup owns one 10^3 block,
// gathers the particles of the cells around it from the arrays sorted into
// half-cell bins,
// replays the reference's particle-to-block assignment and local arithmetic,
// and folds distances into LDS with an order-independent atomic minimum.
// The result is bit-identical to the reference.
#include "device_math.h"
#include "flipmi_kernels.h"

namespace flipmi {

constexpr int kBW = 10;            // particlelevelset.h:140 (_blockwidth)
constexpr int kHalo = 2;           // search radius 2r = sqrt(3) dx < 2 dx
constexpr int kSpan = kBW + 2 * kHalo;
constexpr int kSdfThreads = 256;

// Whether the reference assigns point p to block (bi, bj, bk)
// (particlelevelset.cpp:497-536; velocityadvector.cpp:280-320).
__device__ __forceinline__ bool point_in_block(F3 p, float sr, float blockdx, double inv_blockdx,
                                               int bi, int bj, int bk) {
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

// ParticleLevelSet::postProcessSignedDistanceField (particlelevelset.cpp:
// 172-196) for one cell: cells within half a cell of the surface whose centre
// lies inside the solid become -dx/2 (MeshLevelSet::getDistanceAtCellCenter,
// meshlevelset.cpp:152-162), and |phi| < 0.005 dx is pushed out to 0.005 dx.
__device__ __forceinline__ float postprocess_phi(const GridGeom &g, const float *__restrict__ solid,
                                                 int i, int j, int k, float v) {
    Dims N{g.I + 1, g.J + 1, g.K + 1};
    float eps = (float)(0.005 * g.dx);
    if ((double)v < 0.5 * g.dx) {
        float sc = 0.125f * (solid[N.at(i, j, k)] + solid[N.at(i + 1, j, k)] +
                             solid[N.at(i, j + 1, k)] + solid[N.at(i + 1, j + 1, k)] +
                             solid[N.at(i, j, k + 1)] + solid[N.at(i + 1, j, k + 1)] +
                             solid[N.at(i, j + 1, k + 1)] + solid[N.at(i + 1, j + 1, k + 1)]);
        if (sc < 0) v = (float)(-0.5f * g.dx);
    }
    if (fabsf(v) < eps) v = v > 0 ? eps : -eps;
    return v;
}

constexpr int kRows = 4 * kSpan * kSpan;           // half-cell rows of the candidate box
constexpr int kRowsPerThread = (kRows + kSdfThreads - 1) / kSdfThreads;

__global__ __launch_bounds__(kSdfThreads) void k_liquid_sdf(
    GridGeom g, const float4 *__restrict__ spos, const int *__restrict__ sub_offset, float radius,
    const float *__restrict__ solid, float *phi_out) {
    __shared__ uint32_t phi_ord[kBW * kBW * kBW];
    __shared__ int run_begin[kRows];
    __shared__ int run_prefix[kRows + 1];
    __shared__ int wave_total[kSdfThreads / 64];
    __shared__ float centre[kBW];  // Grid3d::GridIndexToCellCenter of local index c

    int nbi = (g.I + kBW - 1) / kBW, nbj = (g.J + kBW - 1) / kBW;
    int bid = blockIdx.x;
    int bi = bid % nbi, bj = (bid / nbi) % nbj, bk = bid / (nbi * nbj);
    float maxd = (float)(3.0 * g.dx);  // particlelevelset.cpp:295-297
    uint32_t init = f2ord(maxd);
    for (int c = threadIdx.x; c < kBW * kBW * kBW; c += kSdfThreads) phi_ord[c] = init;
    if (threadIdx.x < kBW) centre[threadIdx.x] = idx_to_center1(threadIdx.x, g.dx);

    // Candidate particles: cells [b0 - 2, b0 + 12) per axis.  In the
    // half-cell sort each (sy, sz) row of half-cells over those cells is one
    // contiguous run: 4 rows per (j, k) row of cells.
    int i0 = max(bi * kBW - kHalo, 0), i1 = min(bi * kBW + kBW + kHalo, g.I);
    int j0 = max(bj * kBW - kHalo, 0), j1 = min(bj * kBW + kBW + kHalo, g.J);
    int k0 = max(bk * kBW - kHalo, 0), k1 = min(bk * kBW + kBW + kHalo, g.K);
    int nsy = 2 * (j1 - j0), nsz = 2 * (k1 - k0), nrows = nsy * nsz;
    const size_t SI = 2 * (size_t)g.I, SJ = 2 * (size_t)g.J;
    // Each thread takes kRowsPerThread consecutive rows; a block-wide
    // exclusive scan of the row lengths gives every row's start.
    int len[kRowsPerThread];
    int mine = 0;
#pragma unroll
    for (int q = 0; q < kRowsPerThread; q++) {
        int r = threadIdx.x * kRowsPerThread + q;
        len[q] = 0;
        if (r < nrows) {
            size_t sy = 2 * (size_t)j0 + (size_t)(r % nsy), sz = 2 * (size_t)k0 + (size_t)(r / nsy);
            size_t b0 = 2 * (size_t)i0 + SI * (sy + SJ * sz);
            size_t b1 = b0 + 2 * (size_t)(i1 - i0);
            int s0 = sub_offset[b0];
            run_begin[r] = s0;
            len[q] = sub_offset[b1] - s0;
        }
        mine += len[q];
    }
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(inc, off, 64);
        if (lane >= off) inc += t;
    }
    if (lane == 63) wave_total[wave] = inc;
    __syncthreads();
    int run = inc - mine;
    for (int w = 0; w < wave; w++) run += wave_total[w];
#pragma unroll
    for (int q = 0; q < kRowsPerThread; q++) {
        int r = threadIdx.x * kRowsPerThread + q;
        if (r < nrows) run_prefix[r] = run;
        run += len[q];
    }
    if (threadIdx.x == kSdfThreads - 1) run_prefix[nrows] = run;
    __syncthreads();
    int total = run_prefix[nrows];

    float r = radius;
    float sr = 2.0f * r;  // particlelevelset.h:142 (_searchRadiusFactor)
    float blockdx = (float)(kBW * g.dx);
    double inv_blockdx = 1.0 / (double)blockdx;
    F3 boff = idx_to_pos(bi, bj, bk, kBW * g.dx);
    for (int t = threadIdx.x; t < total; t += kSdfThreads) {
        int lo = 0, hi = nrows - 1;
        while (lo < hi) {  // last row whose prefix <= t
            int mid = (lo + hi + 1) >> 1;
            if (run_prefix[mid] <= t) lo = mid; else hi = mid - 1;
        }
        int q = run_begin[lo] + (t - run_prefix[lo]);
        const float4 p4 = spos[q];
        F3 p0 = f3(p4.x, p4.y, p4.z);
        if (!point_in_block(p0, sr, blockdx, inv_blockdx, bi, bj, bk)) continue;
        F3 p = p0 - boff;
        I3 gmin = pos_to_index((double)(p.x - sr), (double)(p.y - sr), (double)(p.z - sr), g.invdx);
        I3 gmax = pos_to_index((double)(p.x + sr), (double)(p.y + sr), (double)(p.z + sr), g.invdx);
        gmin.i = max(gmin.i, 0); gmin.j = max(gmin.j, 0); gmin.k = max(gmin.k, 0);
        gmax.i = min(gmax.i, kBW - 1); gmax.j = min(gmax.j, kBW - 1); gmax.k = min(gmax.k, kBW - 1);
        for (int k = gmin.k; k <= gmax.k; k++) {
            float cz = centre[k];
            for (int j = gmin.j; j <= gmax.j; j++) {
                float cy = centre[j];
                for (int i = gmin.i; i <= gmax.i; i++) {
                    F3 c = f3(centre[i], cy, cz);
                    F3 dv = c - p;
                    float d2 = dot3(dv, dv);
                    // Values only fall, so a distance no smaller than the
                    // value already there cannot change it.  A squared
                    // distance clearly beyond (value + r)^2 (margins of 2e-5
                    // relative cover every rounding of the exact test) skips
                    // the square root; otherwise the exact distance decides.
                    uint32_t *slot = &phi_ord[i + kBW * (j + kBW * k)];
                    uint32_t cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    float lim = ord2f(cur) + r * 1.00002f;
                    if (lim > 0.0f && d2 > lim * lim * 1.00002f) continue;
                    float dist = sqrtf(d2) - r;
                    uint32_t o = f2ord(dist);
                    if (o < cur) atomicMin(slot, o);
                }
            }
        }
    }
    __syncthreads();

    // Write back, fused with the post-processing pass when a solid is given
    // (calculateSignedDistanceField alone otherwise).
    for (int c = threadIdx.x; c < kBW * kBW * kBW; c += kSdfThreads) {
        int li = c % kBW, lj = (c / kBW) % kBW, lk = c / (kBW * kBW);
        int i = bi * kBW + li, j = bj * kBW + lj, k = bk * kBW + lk;
        if (i >= g.I || j >= g.J || k >= g.K) continue;
        float v = ord2f(phi_ord[c]);
        if (solid) v = postprocess_phi(g, solid, i, j, k, v);
        phi_out[(size_t)i + (size_t)g.I * ((size_t)j + (size_t)g.J * k)] = v;
    }
}

// The post-processing pass on its own, over an existing field (one thread
// per cell).
__global__ void k_liquid_postprocess(GridGeom g, const float *__restrict__ solid, float *phi) {
    size_t n = (size_t)g.I * g.J * g.K;
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    int i = (int)(c % (size_t)g.I), j = (int)((c / (size_t)g.I) % (size_t)g.J), k = (int)(c / ((size_t)g.I * g.J));
    phi[c] = postprocess_phi(g, solid, i, j, k, phi[c]);
}

hipError_t launch_liquid_postprocess(const GridGeom &g, const float *solid_phi_nodal, float *phi,
                                     hipStream_t s) {
    size_t n = (size_t)g.I * g.J * g.K;
    if (n == 0) return hipSuccess;
    k_liquid_postprocess<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, solid_phi_nodal, phi);
    return hipGetLastError();
}

hipError_t launch_liquid_sdf(const GridGeom &g, const float4 *spos, const int *sub_offset, double radius,
                             const float *solid_phi_nodal, float *phi, hipStream_t s) {
    int nbi = (g.I + kBW - 1) / kBW, nbj = (g.J + kBW - 1) / kBW, nbk = (g.K + kBW - 1) / kBW;
    k_liquid_sdf<<<nbi * nbj * nbk, kSdfThreads, 0, s>>>(g, spos, sub_offset, (float)radius, solid_phi_nodal,
                                                         phi);
    return hipGetLastError();
}

}  // namespace flipmi
