// Surface reconstruction: the isomesh of FluidSimulation::
// _outputSurfaceMeshThread (fluidsimulation.cpp:5134-5243) for one
// polygonizer slice, bit-identical to the reference.
//
//  1. Blocks: the 10^3 node blocks of the (subdivided) node grid holding a
//     particle, feathered by one block (ParticleMesher::_initializeCompute
//     ChunkDataActiveBlocks, particlemesher.cpp:107-122), and their bounds,
//     which the host turns into the compute chunk (:124-275).
//  2. Field: one workgroup per block of the chunk.  Its candidates are the
//     particles of the half-cell bins around the block (the step's half-cell
//     sort); each replays the reference's block binning (:486-541) and
//     folds |node - p| - r into the block's nodes in LDS with an ordered-int
//     atomic minimum (:574-608), in the reference's block-local float
//     arithmetic.  The minimum does not depend on order, so the field is
//     exact.  Nodes are negated and clamped to 0 inside the solid
//     (ScalarField::setSolidSDF / getScalarFieldValue, scalarfield.cpp:
//     372-448).
//  3. Marching cubes (Polygonizer3d, polygonizer3d.cpp:349-675): per cell
//     the case, the cut edges it owns (the reference makes a vertex in the
//     first cell, in k-j-i order, that meets the edge, so ownership follows
//     from the cell's position alone) and its triangle count; two scans give
//     every cell its first vertex and triangle, which reproduces the
//     reference's numbering.  Vertices are interpolated in the owner cell,
//     clamped out of the solid along the edge, as the reference does.
//  4. Post-processing: removal near the domain (order-preserving
//     compaction), Laplacian smoothing over a vertex -> triangle adjacency in
//     ascending triangle order, inverted contact normals, domain transform.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_math.h"
#include "mc_tables.h"
#include "surface.h"

namespace flipmi {

__constant__ uint64_t c_mc_tri[256] = {FLIPMI_MC_TRIANGLES};

constexpr int kMB = 10;               // particlemesher.h (_blockwidth)
constexpr int kMeshThreads = 256;
constexpr int kMeshMaxSpan = 40;      // candidate half-cells per axis of a block
constexpr int kMeshMaxRows = kMeshMaxSpan * kMeshMaxSpan;
constexpr int kMeshRowsPerThread = (kMeshMaxRows + kMeshThreads - 1) / kMeshThreads;

struct MeshChunk {
    GridGeom g;
    int nb[3];        // global block grid
    int b0[3];        // first block of the chunk
    int cb[3];        // blocks of the chunk's block array
    int nn[3];        // chunk nodes (field dims)
    int g0[3];        // first node of the chunk (subdivided index)
    float off[3];     // positionOffset
    double subdx, inv_subdx;
    float blockdx_f;  // `float blockdx = _blockwidth * _subdx` (particlemesher.cpp:494)
    double inv_blockdx_f, blockdx_d;
    float sr, r, fill;
    float bb_pos[3];  // the chunk's AABB, expanded by 2 (r + 1e-6)
    double bb_ext[3];
    int matching, aligned2;  // solid node modes of ScalarField::setSolidSDF
};

// ---- 1. blocks --------------------------------------------------------------

__global__ void k_mesh_seed(const float4 *__restrict__ sp, int n, double inv_blockdx, int nbi, int nbj, int nbk,
                            uint8_t *seed) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    float4 p = sp[q];
    I3 b = pos_to_index(f3(p.x, p.y, p.z), inv_blockdx);
    if (b.i >= 0 && b.j >= 0 && b.k >= 0 && b.i < nbi && b.j < nbj && b.k < nbk)
        seed[(size_t)b.i + (size_t)nbi * ((size_t)b.j + (size_t)nbj * b.k)] = 1;
}

// GridUtils::featherGrid26 (gridutils.cpp:265-298) and the bounds of the
// result (bounds[0..2] min, [3..5] max).
__global__ void k_mesh_feather(const uint8_t *__restrict__ seed, int nbi, int nbj, int nbk, uint8_t *active,
                               int *bounds) {
    size_t nb = (size_t)nbi * nbj * nbk;
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nb) return;
    int i = (int)(q % nbi), j = (int)((q / nbi) % nbj), k = (int)(q / ((size_t)nbi * nbj));
    uint8_t a = 0;
    for (int dk = -1; dk <= 1; dk++)
        for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++) {
                int x = i + di, y = j + dj, z = k + dk;
                if (x < 0 || y < 0 || z < 0 || x >= nbi || y >= nbj || z >= nbk) continue;
                a |= seed[(size_t)x + (size_t)nbi * ((size_t)y + (size_t)nbj * z)];
            }
    active[q] = a;
    if (a) {
        atomicMin(&bounds[0], i); atomicMin(&bounds[1], j); atomicMin(&bounds[2], k);
        atomicMax(&bounds[3], i); atomicMax(&bounds[4], j); atomicMax(&bounds[5], k);
    }
}

__global__ void k_mesh_bounds_init(int *bounds) {
    if (threadIdx.x < 3) bounds[threadIdx.x] = 0x7fffffff;
    else if (threadIdx.x < 6) bounds[threadIdx.x] = -1;
}

// ---- 2. field -----------------------------------------------------------------

__device__ __forceinline__ bool chunk_box_inside(const MeshChunk &c, F3 p) {  // aabb.cpp:130-133
    return p.x >= c.bb_pos[0] && p.y >= c.bb_pos[1] && p.z >= c.bb_pos[2] &&
           (double)p.x < (double)c.bb_pos[0] + c.bb_ext[0] && (double)p.y < (double)c.bb_pos[1] + c.bb_ext[1] &&
           (double)p.z < (double)c.bb_pos[2] + c.bb_ext[2];
}

// Whether _computeGridCountDataThread bins chunk-local p to block b (an
// active block).
__device__ __forceinline__ bool mesh_binned_to(const MeshChunk &c, F3 p, int bi, int bj, int bk) {
    const float sr = c.sr, bdx = c.blockdx_f;
    I3 b = pos_to_index(p, c.inv_blockdx_f);
    F3 bp = idx_to_pos(b.i, b.j, b.k, (double)bdx);
    if (p.x - sr > bp.x && p.y - sr > bp.y && p.z - sr > bp.z && p.x + sr < bp.x + bdx && p.y + sr < bp.y + bdx &&
        p.z + sr < bp.z + bdx)
        return b.i == bi && b.j == bj && b.k == bk;
    I3 lo = pos_to_index((double)(p.x - sr), (double)(p.y - sr), (double)(p.z - sr), c.inv_blockdx_f);
    I3 hi = pos_to_index((double)(p.x + sr), (double)(p.y + sr), (double)(p.z + sr), c.inv_blockdx_f);
    return bi >= lo.i && bi <= hi.i && bj >= lo.j && bj <= hi.j && bk >= lo.k && bk <= hi.k;
}

__global__ __launch_bounds__(kMeshThreads) void k_mesh_field(MeshChunk c, const float4 *__restrict__ spos,
                                                             const int *__restrict__ sub_offset,
                                                             const uint8_t *__restrict__ active, float *field) {
    __shared__ uint32_t ord[kMB * kMB * kMB];
    __shared__ int run_begin[kMeshMaxRows];
    __shared__ int run_prefix[kMeshMaxRows + 1];
    __shared__ int wave_total[kMeshThreads / 64];
    __shared__ float npos[kMB];

    const int bid = blockIdx.x;
    const int bi = bid % c.cb[0], bj = (bid / c.cb[0]) % c.cb[1], bk = bid / (c.cb[0] * c.cb[1]);
    const int gbi = c.b0[0] + bi, gbj = c.b0[1] + bj, gbk = c.b0[2] + bk;
    const bool act = active[(size_t)gbi + (size_t)c.nb[0] * ((size_t)gbj + (size_t)c.nb[1] * gbk)] != 0;
    const uint32_t init = f2ord(c.fill);
    for (int q = threadIdx.x; q < kMB * kMB * kMB; q += kMeshThreads) ord[q] = init;
    if (threadIdx.x < kMB) npos[threadIdx.x] = idx_to_pos1(threadIdx.x, c.subdx);

    // Candidate half-cells: a particle reaches this block's nodes only from
    // [first node - 1 node - sr, last node + 1 node + sr]; one half-cell of
    // margin on each side covers every rounding.
    int hs0[3], hs1[3];
    const int bix[3] = {bi, bj, bk};
    const int nhalf[3] = {2 * c.g.I, 2 * c.g.J, 2 * c.g.K};
    const double hinv = 2.0 * c.g.invdx;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double node0 = (double)(c.g0[d] + kMB * bix[d]);
        const double lo = (node0 - 1.0) * c.subdx - (double)c.sr;
        const double hi = (node0 + kMB + 1.0) * c.subdx + (double)c.sr;
        hs0[d] = max((int)floor(lo * hinv) - 1, 0);
        hs1[d] = min((int)floor(hi * hinv) + 1, nhalf[d] - 1);
    }
    int nsy = hs1[1] - hs0[1] + 1, nsz = hs1[2] - hs0[2] + 1;
    int nrows = (act && hs1[0] >= hs0[0] && nsy > 0 && nsz > 0) ? nsy * nsz : 0;
    if (nrows > kMeshMaxRows) nrows = 0;  // excluded on the host (span check)
    const size_t SI = (size_t)nhalf[0], SJ = (size_t)nhalf[1];
    int len[kMeshRowsPerThread];
    int mine = 0;
#pragma unroll
    for (int q = 0; q < kMeshRowsPerThread; q++) {
        int rr = threadIdx.x * kMeshRowsPerThread + q;
        len[q] = 0;
        if (rr < nrows) {
            size_t sy = (size_t)hs0[1] + (size_t)(rr % nsy), sz = (size_t)hs0[2] + (size_t)(rr / nsy);
            size_t row = SI * (sy + SJ * sz);
            int s0 = sub_offset[row + (size_t)hs0[0]];
            run_begin[rr] = s0;
            len[q] = sub_offset[row + (size_t)hs1[0] + 1] - s0;
        }
        mine += len[q];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) wave_total[wave] = inc;
    __syncthreads();
    int run = inc - mine;
    for (int w = 0; w < wave; w++) run += wave_total[w];
#pragma unroll
    for (int q = 0; q < kMeshRowsPerThread; q++) {
        int rr = threadIdx.x * kMeshRowsPerThread + q;
        if (rr < nrows) run_prefix[rr] = run;
        run += len[q];
    }
    if (threadIdx.x == kMeshThreads - 1) run_prefix[nrows] = run;
    __syncthreads();
    const int total = run_prefix[nrows];

    const float r = c.r, sr = c.sr;
    const F3 off = f3(c.off[0], c.off[1], c.off[2]);
    const F3 boff = idx_to_pos(bi, bj, bk, c.blockdx_d);
    for (int t = threadIdx.x; t < total; t += kMeshThreads) {
        int lo = 0, hi = nrows - 1;
        while (lo < hi) {  // last row whose prefix <= t
            int mid = (lo + hi + 1) >> 1;
            if (run_prefix[mid] <= t) lo = mid; else hi = mid - 1;
        }
        const float4 p4 = spos[run_begin[lo] + (t - run_prefix[lo])];
        const F3 p0 = f3(p4.x, p4.y, p4.z);
        if (!chunk_box_inside(c, p0)) continue;
        const F3 p = p0 - off;
        if (!mesh_binned_to(c, p, bi, bj, bk)) continue;
        const F3 pb = p - boff;
        I3 g0 = pos_to_index((double)(pb.x - sr), (double)(pb.y - sr), (double)(pb.z - sr), c.inv_subdx);
        I3 g1 = pos_to_index((double)(pb.x + sr), (double)(pb.y + sr), (double)(pb.z + sr), c.inv_subdx);
        g0.i = max(g0.i, 0); g0.j = max(g0.j, 0); g0.k = max(g0.k, 0);
        g1.i = min(g1.i + 1, kMB - 1); g1.j = min(g1.j + 1, kMB - 1); g1.k = min(g1.k + 1, kMB - 1);
        for (int k = g0.k; k <= g1.k; k++) {
            const float z = npos[k];
            for (int j = g0.j; j <= g1.j; j++) {
                const float y = npos[j];
                for (int i = g0.i; i <= g1.i; i++) {
                    const F3 dv = f3(npos[i], y, z) - pb;
                    const float d2 = dot3(dv, dv);
                    // Values only fall: a distance no smaller than the value
                    // there cannot change it.  A squared distance clearly
                    // beyond (value + r)^2 (2e-5 relative margins cover the
                    // rounding of the exact test) skips the square root.
                    uint32_t *slot = &ord[i + kMB * (j + kMB * k)];
                    const uint32_t cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    const float lim = ord2f(cur) + r * 1.00002f;
                    if (lim > 0.0f && d2 > lim * lim * 1.00002f) continue;
                    const uint32_t o = f2ord(sqrtf(d2) - r);
                    if (o < cur) atomicMin(slot, o);
                }
            }
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kMB * kMB * kMB; q += kMeshThreads) {
        const int i = kMB * bi + q % kMB, j = kMB * bj + (q / kMB) % kMB, k = kMB * bk + q / (kMB * kMB);
        if (i >= c.nn[0] || j >= c.nn[1] || k >= c.nn[2]) continue;
        field[(size_t)i + (size_t)c.nn[0] * ((size_t)j + (size_t)c.nn[1] * k)] = -ord2f(ord[q]);
    }
}

// The value the polygonizer reads at each node: 0 where the node is in the
// solid and the field is above the threshold (scalarfield.cpp:372-448,
// meshlevelset.cpp:309-340).
__global__ void k_mesh_solid_clamp(MeshChunk c, const float *__restrict__ solid, float *field) {
    const size_t nn = (size_t)c.nn[0] * c.nn[1] * c.nn[2];
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nn) return;
    const int i = (int)(q % c.nn[0]), j = (int)((q / c.nn[0]) % c.nn[1]), k = (int)(q / ((size_t)c.nn[0] * c.nn[1]));
    const Dims N{c.g.I + 1, c.g.J + 1, c.g.K + 1};
    float d;
    if (c.matching) {
        d = solid[N.at(i, j, k)];
    } else if (c.aligned2 && ((i | j | k) & 1) == 0) {
        d = solid[N.at(i >> 1, j >> 1, k >> 1)];
    } else {
        F3 p = idx_to_pos(i, j, k, c.subdx) + f3(c.off[0], c.off[1], c.off[2]);
        d = (float)nodal_trilinear(solid, N, c.g.dx, c.g.invdx, p);
    }
    const float v = field[q];
    if (d < 0 && v > 0.0f) field[q] = 0.0f;
}

// ---- 3. marching cubes ---------------------------------------------------------

// Cut edges a cell owns: the vertex of an edge is made by the first cell in
// k-j-i order that contains it (polygonizer3d.cpp:619-631), which for
// edges 5, 6, 10 is always this cell and for the others depends only on
// whether the cell lies on the chunk's low i, j or k side.
__device__ __forceinline__ int mc_owned(int i, int j, int k) {
    int m = (1 << 5) | (1 << 6) | (1 << 10);
    if (j == 0) m |= (1 << 1) | (1 << 2);
    if (k == 0) m |= (1 << 4) | (1 << 9);
    if (i == 0) m |= (1 << 7) | (1 << 11);
    if (j == 0 && k == 0) m |= 1 << 0;
    if (i == 0 && j == 0) m |= 1 << 3;
    if (i == 0 && k == 0) m |= 1 << 8;
    return m;
}

__device__ __forceinline__ int mc_corner(int c) {
    const int t[8] = {FLIPMI_MC_CORNER_OFFSETS};
    return t[c];
}

// Edge e's vertex slot: the edge grids (U, V, W, polygonizer3d.h EdgeGrid)
// index an edge by its lower node; edges 2, 3, 6, 7 run from their upper
// node to their lower one.
__device__ __forceinline__ size_t mc_edge_slot(int e, int i, int j, int k, const int nn[3]) {
    const int a = kMcEdgeA[e], b = kMcEdgeB[e];
    const int lo = (e == 2 || e == 3 || e == 6 || e == 7) ? b : a;
    const int o = mc_corner(lo);
    const size_t NN = (size_t)nn[0] * nn[1] * nn[2];
    const size_t node = (size_t)(i + (o & 1)) + (size_t)nn[0] * ((size_t)(j + ((o >> 1) & 1)) +
                                                               (size_t)nn[1] * (size_t)(k + ((o >> 2) & 1)));
    const int type = (e == 0 || e == 2 || e == 4 || e == 6) ? 0 : (e >= 8 ? 1 : 2);
    return (size_t)type * NN + node;
}

__global__ void k_mc_count(MeshChunk c, const float *__restrict__ val, uint8_t *cube_out, int *vcnt, int *tcnt) {
    const int ci = c.nn[0] - 1, cj = c.nn[1] - 1, ck = c.nn[2] - 1;
    const size_t nc = (size_t)ci * cj * ck;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nc) return;
    const int i = (int)(q % ci), j = (int)((q / ci) % cj), k = (int)(q / ((size_t)ci * cj));
    int cube = 0;
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const int o = mc_corner(a);
        const size_t n = (size_t)(i + (o & 1)) + (size_t)c.nn[0] * ((size_t)(j + ((o >> 1) & 1)) +
                                                                 (size_t)c.nn[1] * (size_t)(k + ((o >> 2) & 1)));
        if (val[n] > 0.0f) cube |= 1 << a;
    }
    const int cut = mc_cut_edges(cube);
    cube_out[q] = (uint8_t)cube;
    vcnt[q] = __popc(cut & mc_owned(i, j, k));
    tcnt[q] = (int)(c_mc_tri[cube] >> 60);
}

// Polygonizer3d::_vertexInterp (polygonizer3d.cpp:409-444) on the edge from
// corner node pa (value va) to pb (vb), positions in chunk coordinates.
__device__ __forceinline__ F3 mc_interp(const MeshChunk &c, const float *__restrict__ solid, F3 pa, F3 pb,
                                        double va, double vb) {
    const double eps = 1e-10;
    double minmu = 0.0, maxmu = 1.0;
    const Dims N{c.g.I + 1, c.g.J + 1, c.g.K + 1};
    const F3 off = f3(c.off[0], c.off[1], c.off[2]);
    const double s1 = nodal_trilinear(solid, N, c.g.dx, c.g.invdx, pa + off);
    const double s2 = nodal_trilinear(solid, N, c.g.dx, c.g.invdx, pb + off);
    if ((s1 < 0.0 && s2 >= 0.0) || (s2 < 0.0 && s1 >= 0.0)) {
        const double diff = s2 - s1;
        if (fabs(diff) > eps) {
            const double su = -s1 / diff;
            if (s1 < 0.0) minmu = su; else maxmu = su;
        } else {
            maxmu = minmu;
        }
    }
    minmu = fmax(minmu, eps);
    maxmu = fmin(maxmu, 1.0 - eps);
    double mu = (0.0 - va) / (vb - va);
    if (mu < minmu) mu = minmu;
    if (mu > maxmu) mu = maxmu;
    return pa + (float)mu * (pb - pa);
}

__global__ void k_mc_vertices(MeshChunk c, const float *__restrict__ val, const float *__restrict__ solid,
                              const uint8_t *__restrict__ cube_in, const int *__restrict__ vbase, float scale,
                              float inv_scale, float *verts, int *edges) {
    const int ci = c.nn[0] - 1, cj = c.nn[1] - 1, ck = c.nn[2] - 1;
    const size_t nc = (size_t)ci * cj * ck;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nc) return;
    const int i = (int)(q % ci), j = (int)((q / ci) % cj), k = (int)(q / ((size_t)ci * cj));
    const int own = mc_cut_edges(cube_in[q]) & mc_owned(i, j, k);
    if (own == 0) return;
    const float fdx = (float)c.subdx;
    const F3 off = f3(c.off[0], c.off[1], c.off[2]);
    int id = vbase[q];
    for (int e = 0; e < 12; e++) {
        if (!((own >> e) & 1)) continue;
        const int oa = mc_corner(kMcEdgeA[e]), ob = mc_corner(kMcEdgeB[e]);
        const int ai = i + (oa & 1), aj = j + ((oa >> 1) & 1), ak = k + ((oa >> 2) & 1);
        const int bi = i + (ob & 1), bj = j + ((ob >> 1) & 1), bk = k + ((ob >> 2) & 1);
        const double va = val[(size_t)ai + (size_t)c.nn[0] * ((size_t)aj + (size_t)c.nn[1] * ak)];
        const double vb = val[(size_t)bi + (size_t)c.nn[0] * ((size_t)bj + (size_t)c.nn[1] * bk)];
        // _getVertexPosition: (float)dx * vec3(i, j, k) (polygonizer3d.cpp:354-357).
        const F3 pa = fdx * f3((float)ai, (float)aj, (float)ak);
        const F3 pb = fdx * f3((float)bi, (float)bj, (float)bk);
        F3 v = mc_interp(c, solid, pa, pb, va, vb);
        // m.translate(positionOffset), then the mesher's scale by localdx /
        // subdx and back (particlemesher.cpp:46-53, polygonizer output).
        v = v + off;
        v = F3{v.x * scale, v.y * scale, v.z * scale};
        v = F3{v.x * inv_scale, v.y * inv_scale, v.z * inv_scale};
        verts[3 * (size_t)id] = v.x;
        verts[3 * (size_t)id + 1] = v.y;
        verts[3 * (size_t)id + 2] = v.z;
        edges[mc_edge_slot(e, i, j, k, c.nn)] = id;
        id++;
    }
}

__global__ void k_mc_triangles(MeshChunk c, const uint8_t *__restrict__ cube_in, const int *__restrict__ tbase,
                               const int *__restrict__ edges, int nv, int *tris, int *bad) {
    const int ci = c.nn[0] - 1, cj = c.nn[1] - 1, ck = c.nn[2] - 1;
    const size_t nc = (size_t)ci * cj * ck;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nc) return;
    const uint64_t w = c_mc_tri[cube_in[q]];
    const int nt = (int)(w >> 60);
    if (nt == 0) return;
    const int i = (int)(q % ci), j = (int)((q / ci) % cj), k = (int)(q / ((size_t)ci * cj));
    int *out = tris + 3 * (size_t)tbase[q];
    for (int s = 0; s < 3 * nt; s++) {
        int id = edges[mc_edge_slot((int)((w >> (4 * s)) & 15), i, j, k, c.nn)];
        // Every cut edge's vertex was made by its owner cell; an id out of
        // range would mean it was not, and is reported instead of used.
        if (id < 0 || id >= nv) {
            atomicOr(bad, 1);
            id = 0;
        }
        out[s] = id;
    }
}

// ---- 4. post-processing -----------------------------------------------------------

__device__ __forceinline__ F3 load_v(const float *v, int q) {
    return f3(v[3 * (size_t)q], v[3 * (size_t)q + 1], v[3 * (size_t)q + 2]);
}

// _removeMeshNearDomain (fluidsimulation.cpp:4922-4956): triangles whose
// centroid's cell lies within `width` cells of the domain's sides go.
__global__ void k_tri_keep_interior(const float *__restrict__ verts, const int *__restrict__ tris, int nt,
                                    GridGeom g, int width, int *keep) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const F3 cen = div3(load_v(verts, tris[3 * t]) + load_v(verts, tris[3 * t + 1]) + load_v(verts, tris[3 * t + 2]),
                        3.0f);
    const I3 c = pos_to_index(cen, g.invdx);
    keep[t] = c.i >= width && c.i < g.I - width && c.j >= width && c.j < g.J - width && c.k >= width &&
              c.k < g.K - width;
}

__global__ void k_tri_compact(const int *__restrict__ tris, const int *__restrict__ keep,
                              const int *__restrict__ kpos, int nt, int *out, int *used) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt || !keep[t]) return;
    for (int q = 0; q < 3; q++) {
        const int v = tris[3 * t + q];
        out[3 * (size_t)kpos[t] + q] = v;
        used[v] = 1;
    }
}

__global__ void k_vert_compact(const float *__restrict__ verts, const int *__restrict__ used,
                               const int *__restrict__ upos, int nv, float *out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv || !used[v]) return;
    for (int d = 0; d < 3; d++) out[3 * (size_t)upos[v] + d] = verts[3 * (size_t)v + d];
}

__global__ void k_tri_remap(int *tris, int n3, const int *__restrict__ upos) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n3) tris[q] = upos[tris[q]];
}

__global__ void k_adj_count(const int *__restrict__ tris, int n3, int *cnt) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n3) atomicAdd(&cnt[tris[q]], 1);
}

__global__ void k_adj_fill(const int *__restrict__ tris, int n3, const int *__restrict__ off, int *cursor,
                           int *list) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n3) return;
    const int v = tris[q];
    list[off[v] + atomicAdd(&cursor[v], 1)] = q / 3;
}

// TriangleMesh::_updateVertexTriangles (trianglemesh.cpp:481-499) lists a
// vertex's triangles in ascending order: sort each short list.
__global__ void k_adj_sort(const int *__restrict__ off, int nv, int total, int *list) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int b = off[v], e = (v + 1 < nv) ? off[v + 1] : total;
    for (int x = b + 1; x < e; x++) {
        const int key = list[x];
        int y = x - 1;
        while (y >= b && list[y] > key) {
            list[y + 1] = list[y];
            y--;
        }
        list[y + 1] = key;
    }
}

// TriangleMesh::_smoothTriangleMesh (trianglemesh.cpp:536-570).
__global__ void k_smooth(const float *__restrict__ in, const int *__restrict__ tris, const int *__restrict__ off,
                         const int *__restrict__ list, int nv, int total, float value, float *out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int b = off[v], e = (v + 1 < nv) ? off[v + 1] : total;
    F3 avg = f3(0.0f, 0.0f, 0.0f);
    int count = 0;
    for (int x = b; x < e; x++) {
        const int t = list[x];
        for (int q = 0; q < 3; q++) {
            const int u = tris[3 * t + q];
            if (u != v) {
                avg = avg + load_v(in, u);
                count++;
            }
        }
    }
    avg = div3(avg, (float)count);
    const F3 p = load_v(in, v);
    const F3 nv3 = p + value * (avg - p);
    out[3 * (size_t)v] = nv3.x;
    out[3 * (size_t)v + 1] = nv3.y;
    out[3 * (size_t)v + 2] = nv3.z;
}

// _invertContactNormals (fluidsimulation.cpp:4898-4920).
__global__ void k_contact_flags(const float *__restrict__ verts, int nv, const float *__restrict__ solid,
                                GridGeom g, float eps, int *flag) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const Dims N{g.I + 1, g.J + 1, g.K + 1};
    flag[v] = (float)nodal_trilinear(solid, N, g.dx, g.invdx, load_v(verts, v)) < eps;
}

__global__ void k_contact_flip(int *tris, int nt, const int *__restrict__ flag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    if (flag[tris[3 * t]] || flag[tris[3 * t + 1]] || flag[tris[3 * t + 2]]) {
        const int a = tris[3 * t + 1];
        tris[3 * t + 1] = tris[3 * t + 2];
        tris[3 * t + 2] = a;
    }
}

// isomesh.scale(domainScale); isomesh.translate(domainOffset)
// (fluidsimulation.cpp:5191-5193).
__global__ void k_domain_transform(float *verts, int nv, float s, float ox, float oy, float oz) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    float *p = verts + 3 * (size_t)v;
    p[0] = p[0] * s;
    p[1] = p[1] * s;
    p[2] = p[2] * s;
    p[0] = p[0] + ox;
    p[1] = p[1] + oy;
    p[2] = p[2] + oz;
}

// _polygonizeOutputSurface's obstacle meshing offset (fluidsimulation.cpp:
// 5050-5062) on a copy of the solid SDF.
__global__ void k_solid_offset(float *solid, GridGeom g, float offset) {
    const Dims N{g.I + 1, g.J + 1, g.K + 1};
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N.count()) return;
    const int i = (int)(q % N.ni), j = (int)((q / N.ni) % N.nj), k = (int)(q / ((size_t)N.ni * N.nj));
    if (i >= 2 && i < g.I - 1 && j >= 2 && j < g.J - 1 && k >= 2 && k < g.K - 1) solid[q] = solid[q] + offset;
}

// ---- host ------------------------------------------------------------------------

double marker_particle_radius(double dx) {
    double volume = dx * dx * dx / 8.0;
    double pi = 3.141592653;
    return std::pow(3 * volume / (4 * pi), 1.0 / 3.0);
}

static inline float host_idx_to_pos(int i, double dx) { return (float)((double)(float)i * dx); }

// The widest candidate box of a block, in half-cells per axis.
static int mesh_span(const GridGeom &g, int sub, double kernel_radius) {
    const double subdx = g.dx / sub;
    const double sr = (double)(1.5f * (float)kernel_radius);
    return (int)std::ceil(((kMB + 2.0) * subdx + 2.0 * sr) * 2.0 / g.dx) + 3;
}

void validate_surface_desc(const flipmi_surface_desc &d, const GridGeom &g) {
    auto bad = [](const char *m) { throw StatusError{FLIPMI_ERR_INVALID_ARG, m}; };
    if (d.subdivisions < 1) bad("subdivision level must be greater than or equal to 1");
    if (d.polygonizer_slices < 1) bad("number of polygonizer slices must be greater than or equal to 1");
    if (d.smoothing_iterations < 0) bad("smoothing iterations must be greater than or equal to 0");
    if (d.particle_scale < 0.0) bad("marker particle scale must be greater than or equal to 0");
    if (d.min_polyhedron_triangles < 0) bad("minimum polyhedron triangle count must be greater than or equal to 0");
    if (d.remove_near_domain_distance < 0) bad("distance must be greater than or equal to 0");
    if (d.polygonizer_slices != 1) bad("more than one polygonizer slice is not supported");
    if (d.min_polyhedron_triangles != 0) bad("a minimum polyhedron triangle count above 0 is not supported");
    if (mesh_span(g, d.subdivisions, marker_particle_radius(g.dx) * d.particle_scale) > kMeshMaxSpan)
        bad("particle scale too large for the surface kernel's block search box");
    const double nodes = ((double)g.I * d.subdivisions + 1) * ((double)g.J * d.subdivisions + 1) *
                         ((double)g.K * d.subdivisions + 1);
    if (nodes > 2.0e9) bad("subdivided surface grid too large");
}

size_t SurfaceStore::device_bytes() const {
    return seed.bytes() + active.bytes() + small.bytes() + solid_copy.bytes() + field.bytes() + cube.bytes() +
           vcnt.bytes() + tcnt.bytes() + vbase.bytes() + tbase.bytes() + scan_tmp.bytes() + edges.bytes() +
           verts.bytes() + verts_alt.bytes() + tris.bytes() + tris_alt.bytes() + adj_cnt.bytes() +
           adj_off.bytes() + adj_list.bytes() + keep.bytes() + kpos.bytes();
}

template <class T>
static void ensure(DBuf<T> &b, size_t n, const char *what) {
    if (b.n >= n && b.p) return;
    hipError_t e = b.alloc(std::max<size_t>(n, 1));
    if (e != hipSuccess) throw StatusError{FLIPMI_ERR_OOM, std::string("surface buffer: ") + what};
}

static void hcheck(hipError_t e, const char *what) {
    if (e != hipSuccess) throw StatusError{FLIPMI_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

static unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

void build_surface(const SurfaceInputs &in, SurfaceStore &st, hipStream_t s) {
    const GridGeom &g = in.g;
    const flipmi_surface_desc &d = in.desc;
    validate_surface_desc(d, g);
    st.valid = true;
    st.nverts = st.ntris = 0;
    if (in.n <= 0) return;  // fluidsimulation.cpp:5047-5051
    const int sub = d.subdivisions;
    const double subdx = g.dx / (double)sub;
    const int nsn[3] = {g.I * sub + 1, g.J * sub + 1, g.K * sub + 1};
    const int nb[3] = {(nsn[0] + kMB - 1) / kMB, (nsn[1] + kMB - 1) / kMB, (nsn[2] + kMB - 1) / kMB};
    const size_t nblocks = (size_t)nb[0] * nb[1] * nb[2];
    const double radius = in.kernel_radius;

    // 1. blocks and the chunk.
    ensure(st.seed, nblocks, "blocks");
    ensure(st.active, nblocks, "blocks");
    ensure(st.small, 16, "bounds");
    hcheck(launch_fill_u8(st.seed.p, nblocks, 0, s), "fill");
    k_mesh_bounds_init<<<1, 64, 0, s>>>(st.small.p);
    k_mesh_seed<<<nblk(in.n), 256, 0, s>>>(in.spos, in.n, 1.0 / (kMB * subdx), nb[0], nb[1], nb[2], st.seed.p);
    k_mesh_feather<<<nblk(nblocks), 256, 0, s>>>(st.seed.p, nb[0], nb[1], nb[2], st.active.p, st.small.p);
    hcheck(hipGetLastError(), "surface blocks");
    int bounds[6];
    hcheck(hipMemcpyAsync(bounds, st.small.p, sizeof(bounds), hipMemcpyDeviceToHost, s), "bounds");
    hcheck(hipStreamSynchronize(s), "bounds");
    if (bounds[0] > bounds[3]) return;  // no block: empty mesh
    int lo[3] = {bounds[0], bounds[1], bounds[2]}, hi[3] = {bounds[3] + 1, bounds[4] + 1, bounds[5] + 1};
    int split = 0, width = nb[0];  // particlemesher.cpp:142-153
    if (nb[1] > width) { split = 1; width = nb[1]; }
    if (nb[2] > width) { split = 2; width = nb[2]; }
    for (int a = 0; a < 3; a++)  // particlemesher.cpp:221-237
        if (a != split) { lo[a] = std::max(lo[a] - 1, 0); hi[a] = std::min(hi[a] + 1, nb[a]); }

    MeshChunk c{};
    c.g = g;
    for (int a = 0; a < 3; a++) {
        c.nb[a] = nb[a];
        c.b0[a] = lo[a];
        c.g0[a] = kMB * lo[a];
        const int gmax = std::min(kMB * hi[a], nsn[a]);
        c.nn[a] = gmax - c.g0[a];
        c.cb[a] = (c.nn[a] + kMB - 1) / kMB;
        c.off[a] = host_idx_to_pos(c.g0[a], subdx);
        // AABB(pmin, pmax) expanded by 2 (r + eps) (particlemesher.cpp:305-310,
        // aabb.cpp:40-52, 122-128).
        const float pmin = c.off[a], pmax = host_idx_to_pos(gmax, subdx);
        const double mn = std::fmin((double)pmin, (double)pmax), mx = std::fmax((double)pmin, (double)pmax);
        const float eps = 1e-6f;
        const double v = 2.0f * (radius + eps);
        c.bb_pos[a] = (float)mn - (float)(0.5 * v);
        c.bb_ext[a] = (mx - mn) + v;
    }
    c.subdx = subdx;
    c.inv_subdx = 1.0 / subdx;
    c.blockdx_f = (float)(kMB * subdx);
    c.inv_blockdx_f = 1.0 / (double)c.blockdx_f;
    c.blockdx_d = kMB * subdx;
    c.sr = 1.5f * (float)radius;  // particlemesher.h (_searchRadiusFactor)
    c.r = (float)radius;
    c.fill = (float)(3.0 * radius);  // _getMaxDistanceValue
    c.matching = c.nn[0] == g.I + 1 && c.nn[1] == g.J + 1 && c.nn[2] == g.K + 1 && std::fabs(subdx - g.dx) < 1e-12;
    {
        const float l = (float)std::sqrt((double)(c.off[0] * c.off[0] + c.off[1] * c.off[1] + c.off[2] * c.off[2]));
        c.aligned2 = std::fabs(2 * subdx - g.dx) < 1e-6 && l < 1e-6;
    }
    if (mesh_span(g, sub, radius) > kMeshMaxSpan) throw StatusError{FLIPMI_ERR_INVALID_ARG, "surface kernel too wide"};

    // The solid the polygonizer sees (with the obstacle meshing offset).
    const size_t nsolid = (size_t)(g.I + 1) * (g.J + 1) * (g.K + 1);
    const float *solid = in.solid;
    {
        const float offset = (float)(d.obstacle_meshing_offset * g.dx);
        if (std::fabs(offset) > 1e-9f) {
            ensure(st.solid_copy, nsolid, "solid");
            hcheck(hipMemcpyAsync(st.solid_copy.p, in.solid, nsolid * sizeof(float), hipMemcpyDeviceToDevice, s),
                   "solid copy");
            k_solid_offset<<<nblk(nsolid), 256, 0, s>>>(st.solid_copy.p, g, offset);
            solid = st.solid_copy.p;
        }
    }

    // 2. field and solid clamp.
    const size_t nn = (size_t)c.nn[0] * c.nn[1] * c.nn[2];
    const size_t ncb = (size_t)c.cb[0] * c.cb[1] * c.cb[2];
    ensure(st.field, nn, "field");
    k_mesh_field<<<(unsigned)ncb, kMeshThreads, 0, s>>>(c, in.spos, in.sub_offset, st.active.p, st.field.p);
    k_mesh_solid_clamp<<<nblk(nn), 256, 0, s>>>(c, solid, st.field.p);
    hcheck(hipGetLastError(), "surface field");

    // 3. marching cubes.
    const size_t ncell = (size_t)std::max(c.nn[0] - 1, 0) * std::max(c.nn[1] - 1, 0) * std::max(c.nn[2] - 1, 0);
    if (ncell == 0) return;
    if (ncell > (size_t)0x7ffffff0) throw StatusError{FLIPMI_ERR_INVALID_ARG, "surface grid too large"};
    ensure(st.cube, ncell, "cases");
    ensure(st.vcnt, ncell, "counts");
    ensure(st.tcnt, ncell, "counts");
    ensure(st.vbase, ncell, "counts");
    ensure(st.tbase, ncell, "counts");
    ensure(st.scan_tmp, scan_scratch_elems((int)ncell), "scan");
    k_mc_count<<<nblk(ncell), 256, 0, s>>>(c, st.field.p, st.cube.p, st.vcnt.p, st.tcnt.p);
    hcheck(launch_exclusive_scan(st.vcnt.p, st.vbase.p, (int)ncell, st.scan_tmp.p, s), "scan");
    hcheck(launch_exclusive_scan(st.tcnt.p, st.tbase.p, (int)ncell, st.scan_tmp.p, s), "scan");
    int tail[4];
    hcheck(hipMemcpyAsync(&tail[0], st.vbase.p + ncell - 1, sizeof(int), hipMemcpyDeviceToHost, s), "totals");
    hcheck(hipMemcpyAsync(&tail[1], st.vcnt.p + ncell - 1, sizeof(int), hipMemcpyDeviceToHost, s), "totals");
    hcheck(hipMemcpyAsync(&tail[2], st.tbase.p + ncell - 1, sizeof(int), hipMemcpyDeviceToHost, s), "totals");
    hcheck(hipMemcpyAsync(&tail[3], st.tcnt.p + ncell - 1, sizeof(int), hipMemcpyDeviceToHost, s), "totals");
    hcheck(hipStreamSynchronize(s), "totals");
    int nv = tail[0] + tail[1], nt = tail[2] + tail[3];
    if (nt == 0 || nv == 0) return;
    ensure(st.edges, 3 * nn, "edges");
    ensure(st.verts, 3 * (size_t)nv, "vertices");
    ensure(st.verts_alt, 3 * (size_t)nv, "vertices");
    ensure(st.tris, 3 * (size_t)nt, "triangles");
    ensure(st.tris_alt, 3 * (size_t)nt, "triangles");
    const double scale = 0.1 / subdx;  // particlemesher.h (_localdx) / subdx
    hcheck(launch_fill_i32(st.edges.p, 3 * nn, -1, s), "fill");
    hcheck(launch_fill_i32(st.small.p + 8, 1, 0, s), "fill");
    k_mc_vertices<<<nblk(ncell), 256, 0, s>>>(c, st.field.p, solid, st.cube.p, st.vbase.p, (float)scale,
                                              (float)(1.0 / scale), st.verts.p, st.edges.p);
    k_mc_triangles<<<nblk(ncell), 256, 0, s>>>(c, st.cube.p, st.tbase.p, st.edges.p, nv, st.tris.p,
                                               st.small.p + 8);
    hcheck(hipGetLastError(), "marching cubes");
    int bad = 0;
    hcheck(hipMemcpyAsync(&bad, st.small.p + 8, sizeof(int), hipMemcpyDeviceToHost, s), "check");
    hcheck(hipStreamSynchronize(s), "check");
    if (bad) throw StatusError{FLIPMI_ERR_HIP, "surface: a triangle edge has no vertex (internal error)"};

    // 4. post-processing.
    float *V = st.verts.p, *Valt = st.verts_alt.p;
    int *T = st.tris.p, *Talt = st.tris_alt.p;
    ensure(st.adj_cnt, (size_t)nv, "adjacency");
    ensure(st.adj_off, (size_t)nv, "adjacency");
    ensure(st.keep, (size_t)std::max(nv, nt), "flags");
    ensure(st.kpos, (size_t)std::max(nv, nt), "flags");
    ensure(st.scan_tmp, scan_scratch_elems(std::max(nv, 3 * nt)), "scan");
    if (d.remove_near_domain) {
        const int width = 2 + d.remove_near_domain_distance;
        k_tri_keep_interior<<<nblk(nt), 256, 0, s>>>(V, T, nt, g, width, st.keep.p);
        hcheck(launch_exclusive_scan(st.keep.p, st.kpos.p, nt, st.scan_tmp.p, s), "scan");
        int kt[2];
        hcheck(hipMemcpyAsync(&kt[0], st.kpos.p + nt - 1, sizeof(int), hipMemcpyDeviceToHost, s), "totals");
        hcheck(hipMemcpyAsync(&kt[1], st.keep.p + nt - 1, sizeof(int), hipMemcpyDeviceToHost, s), "totals");
        hcheck(launch_fill_i32(st.adj_cnt.p, nv, 0, s), "fill");
        k_tri_compact<<<nblk(nt), 256, 0, s>>>(T, st.keep.p, st.kpos.p, nt, Talt, st.adj_cnt.p);
        hcheck(hipStreamSynchronize(s), "totals");
        const int nt2 = kt[0] + kt[1];
        std::swap(T, Talt);
        nt = nt2;
        if (nt == 0) {
            nv = 0;
        } else {
            hcheck(launch_exclusive_scan(st.adj_cnt.p, st.adj_off.p, nv, st.scan_tmp.p, s), "scan");
            int vt[2];
            hcheck(hipMemcpyAsync(&vt[0], st.adj_off.p + nv - 1, sizeof(int), hipMemcpyDeviceToHost, s), "totals");
            hcheck(hipMemcpyAsync(&vt[1], st.adj_cnt.p + nv - 1, sizeof(int), hipMemcpyDeviceToHost, s), "totals");
            k_vert_compact<<<nblk(nv), 256, 0, s>>>(V, st.adj_cnt.p, st.adj_off.p, nv, Valt);
            k_tri_remap<<<nblk(3 * (size_t)nt), 256, 0, s>>>(T, 3 * nt, st.adj_off.p);
            hcheck(hipStreamSynchronize(s), "totals");
            std::swap(V, Valt);
            nv = vt[0] + vt[1];
        }
    }
    if (d.smoothing_iterations > 0 && nv > 0) {
        const int n3 = 3 * nt;
        ensure(st.adj_list, (size_t)n3, "adjacency");
        hcheck(launch_fill_i32(st.adj_cnt.p, nv, 0, s), "fill");
        k_adj_count<<<nblk(n3), 256, 0, s>>>(T, n3, st.adj_cnt.p);
        hcheck(launch_exclusive_scan(st.adj_cnt.p, st.adj_off.p, nv, st.scan_tmp.p, s), "scan");
        hcheck(launch_fill_i32(st.adj_cnt.p, nv, 0, s), "fill");
        k_adj_fill<<<nblk(n3), 256, 0, s>>>(T, n3, st.adj_off.p, st.adj_cnt.p, st.adj_list.p);
        k_adj_sort<<<nblk(nv), 256, 0, s>>>(st.adj_off.p, nv, n3, st.adj_list.p);
        for (int it = 0; it < d.smoothing_iterations; it++) {
            k_smooth<<<nblk(nv), 256, 0, s>>>(V, T, st.adj_off.p, st.adj_list.p, nv, n3, (float)d.smoothing_value,
                                              Valt);
            std::swap(V, Valt);
        }
        hcheck(hipGetLastError(), "smoothing");
    }
    if (d.invert_contact_normals && nv > 0) {
        const float eps = (float)(0.08 * g.dx);  // fluidsimulation.h:1591 (_contactThresholdDistance)
        k_contact_flags<<<nblk(nv), 256, 0, s>>>(V, nv, in.solid, g, eps, st.keep.p);
        k_contact_flip<<<nblk(nt), 256, 0, s>>>(T, nt, st.keep.p);
    }
    if (nv > 0)
        k_domain_transform<<<nblk(nv), 256, 0, s>>>(V, nv, (float)in.domain_scale, in.domain_offset[0],
                                                    in.domain_offset[1], in.domain_offset[2]);
    hcheck(hipGetLastError(), "surface post-processing");
    // The result lives in verts / tris.
    if (V != st.verts.p) {
        hcheck(hipMemcpyAsync(st.verts.p, V, 3 * (size_t)nv * sizeof(float), hipMemcpyDeviceToDevice, s), "copy");
    }
    if (T != st.tris.p) {
        hcheck(hipMemcpyAsync(st.tris.p, T, 3 * (size_t)nt * sizeof(int), hipMemcpyDeviceToDevice, s), "copy");
    }
    st.nverts = nv;
    st.ntris = nt;
}

void copy_surface(const SurfaceStore &st, float *verts_xyz, int32_t *tris, hipStream_t s) {
    if (verts_xyz && st.nverts > 0)
        hcheck(hipMemcpyAsync(verts_xyz, st.verts.p, 3 * (size_t)st.nverts * sizeof(float), hipMemcpyDeviceToHost, s),
               "mesh copy");
    if (tris && st.ntris > 0)
        hcheck(hipMemcpyAsync(tris, st.tris.p, 3 * (size_t)st.ntris * sizeof(int), hipMemcpyDeviceToHost, s),
               "mesh copy");
    hcheck(hipStreamSynchronize(s), "mesh copy");
}

}  // namespace flipmi
