// Particle-to-grid transfer onto the staggered MAC grid, in the reference's
// summation order.
//
// Reference: VelocityAdvector::advect / _advectGrid / _advectionProducerThread
// (velocityadvector.cpp:38-459).  For every face f of component U, V or W
//     u_f = sum_p w(|x_f - p|) v_p / sum_p w,   w(d) = 1 - 4/9 d^6/r^6
//                                                    + 17/9 d^4/r^4 - 22/9 d^2/r^2
// over particles within r = sqrt(3)/2 dx, with the face valid when the weight
// sum exceeds 1e-6.  The reference evaluates distances in the frame of the
// 10^3 face block that owns f and assigns particles to blocks by their search
// box; both are replayed here, so every per-pair weight is bit-identical.
// Its per-block particle lists are filled in ascending particle index
// (velocityadvector.cpp:331-388) and each face accumulates in that list's
// order (:413-446), so a face's sums run over its contributing particles in
// ascending particle index.  That order is what makes the float sums, and
// the valid flags, bit-identical here.
//
// One thread owns one face.  Its candidates are the particles of the 4 x 4
// x 4 half-cells around the face (the particles sorted into half-cell bins,
// sx fastest, so 16 contiguous runs): every particle within r < 0.87 dx of
// the face lies there, with a margin of 0.13 dx.  Each hit's (w v, w) goes
// to LDS with a key (particle id, slot); the keys are sorted in registers by
// a bitonic network sized to the wave's largest hit count, and the thread
// sums in key order.  Faces with more than kCap hits are listed and summed
// by k_p2g_overflow, one wavefront per face, with the same order.
#include <stdint.h>

#include "device_math.h"
#include "flipmi_kernels.h"

namespace flipmi {

constexpr int kChunk = 10;       // velocityadvector.h:163 (_chunkWidth)
constexpr int kThreads = 64;     // one wavefront per workgroup
constexpr int kCap = 32;         // hits per face sorted in registers
constexpr int kOvfCap = 1024;    // hits per face sorted in LDS by the overflow kernel
constexpr int kOvfBlocks = 2048;
constexpr int kCompShift = 29;   // overflow list entry: comp << 29 | face

struct P2GArgs {
    GridGeom g;
    const float4 *sp;           // (x, y, z, particle id bits), half-cell order
    const float *sv[3];         // velocity components, half-cell order
    const int *sub_offset;      // [8 I J K + 1]
    float radius_f;
    double radius;
    float *out[3];
    uint8_t *valid[3];
    int *ovf_count;
    int *ovf_list;
};

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

// Everything one face needs to test a candidate particle exactly as
// _advectionProducerThread does (velocityadvector.cpp:392-446).
struct FaceCtx {
    int comp, i, j, k;
    size_t f;           // flat index in the component's grid
    F3 off;             // _getDirectionOffset
    float sr, blockdx, rsq, rsq_fast, coef1, coef2, coef3;
    double inv_blockdx;
    int bi, bj, bk, li, lj, lk;
    F3 boff, gpos;
    int lo[3], hi[3];   // candidate half-cells, inclusive
};

__device__ __forceinline__ void face_dims(const GridGeom &g, int comp, int &ni, int &nj, int &nk) {
    ni = g.I + (comp == 0);
    nj = g.J + (comp == 1);
    nk = g.K + (comp == 2);
}

__device__ __forceinline__ void face_setup(const P2GArgs &a, int comp, size_t f, FaceCtx &c) {
    const GridGeom &g = a.g;
    int ni, nj, nk;
    face_dims(g, comp, ni, nj, nk);
    c.comp = comp;
    c.f = f;
    unflatten3(f, ni, nj, c.i, c.j, c.k);
    float hdx = (float)(0.5 * g.dx);  // velocityadvector.cpp:153-164
    c.off = comp == 0 ? f3(0.0f, hdx, hdx) : comp == 1 ? f3(hdx, 0.0f, hdx) : f3(hdx, hdx, 0.0f);
    float eps = 1e-6f;
    c.sr = (float)(a.radius + (double)eps);
    double chunkdx = g.dx * kChunk;
    c.blockdx = (float)chunkdx;
    c.inv_blockdx = 1.0 / (double)c.blockdx;
    float r = a.radius_f;
    c.rsq = r * r;
    // (0.99 r)^2, below which the assignment tests are skipped (face_hit_at).
    const bool small_grid = g.I <= 4096 && g.J <= 4096 && g.K <= 4096;
    c.rsq_fast = small_grid ? 0.98f * c.rsq : 0.0f;
    c.coef1 = (4.0f / 9.0f) * (1.0f / (r * r * r * r * r * r));
    c.coef2 = (17.0f / 9.0f) * (1.0f / (r * r * r * r));
    c.coef3 = (22.0f / 9.0f) * (1.0f / (r * r));
    c.bi = c.i / kChunk; c.bj = c.j / kChunk; c.bk = c.k / kChunk;
    c.li = c.i - c.bi * kChunk; c.lj = c.j - c.bj * kChunk; c.lk = c.k - c.bk * kChunk;
    c.boff = idx_to_pos(c.bi, c.bj, c.bk, kChunk * g.dx);
    c.gpos = idx_to_pos(c.li, c.lj, c.lk, g.dx);
    // The face in half-cell units; particles within r lie in half-cells
    // [F - 2, F + 1] along each axis.
    const int F[3] = {2 * c.i + (comp != 0), 2 * c.j + (comp != 1), 2 * c.k + (comp != 2)};
    const int N[3] = {2 * g.I, 2 * g.J, 2 * g.K};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        c.lo[d] = max(F[d] - 2, 0);
        c.hi[d] = min(F[d] + 1, N[d] - 1);
    }
}

// velocityadvector.cpp:430 (Wyvill kernel), as evaluated there.
__device__ __forceinline__ float face_weight(const FaceCtx &c, float d2) {
    return 1.0f - c.coef1 * d2 * d2 * d2 + c.coef2 * d2 * d2 - c.coef3 * d2;
}

// The reference's block assignment of the (offset) particle p and the
// face's local index range over its search box (velocityadvector.cpp:
// 280-320, 397-408), for a pair already within r.
__device__ __forceinline__ bool assignment_tests(const P2GArgs &a, const FaceCtx &c, F3 p, F3 pl) {
    if (!p2g_point_in_block(p, c.sr, c.blockdx, c.inv_blockdx, c.bi, c.bj, c.bk)) return false;
    I3 gmin = pos_to_index((double)(pl.x - c.sr), (double)(pl.y - c.sr), (double)(pl.z - c.sr), a.g.invdx);
    I3 gmax = pos_to_index((double)(pl.x + c.sr), (double)(pl.y + c.sr), (double)(pl.z + c.sr), a.g.invdx);
    gmin.i = max(gmin.i, 0); gmin.j = max(gmin.j, 0); gmin.k = max(gmin.k, 0);
    gmax.i = min(gmax.i, kChunk - 1); gmax.j = min(gmax.j, kChunk - 1); gmax.k = min(gmax.k, kChunk - 1);
    return !(c.li < gmin.i || c.li > gmax.i || c.lj < gmin.j || c.lj > gmax.j || c.lk < gmin.k ||
             c.lk > gmax.k);
}

// The reference's per-pair test and weight for a particle at p0.
__device__ __forceinline__ bool face_hit_at(const P2GArgs &a, const FaceCtx &c, F3 p0, float &wt) {
    const F3 p = p0 - c.off;
    const F3 pl = p - c.boff;
    const F3 v = c.gpos - pl;
    const float d2 = dot3(v, v);
    if (!(d2 < c.rsq)) return false;
    // The block-assignment and index-range tests below pass for every pair
    // within r in exact arithmetic: the particle's search box reaches
    // r + 1e-6 > |p - f| on each axis, so it contains the face, hence covers
    // the face's block (overlapping case) or lies inside it (simple case),
    // and the face's local index lies in [floor((p - sr)/dx),
    // floor((p + sr)/dx)].  They can only fail through float rounding at a
    // distance within rounding of r.  Inside 0.99 r every axis keeps a
    // margin of 0.01 r = 0.0087 dx.  On a grid of at most 4096 cells per
    // axis a coordinate is below 2^12 dx, so one float rounding moves it by
    // at most 2^-12 dx; the tests round a coordinate at most three times
    // (p0 - off, then - sr or - boff, and the block corner), 0.00073 dx in
    // all, a twelfth of the margin.  Such pairs skip the tests; on larger
    // grids every pair within r runs them.
    if (!(d2 < c.rsq_fast) && !assignment_tests(a, c, p, pl)) return false;
    wt = face_weight(c, d2);
    return true;
}

// The distance part alone (the main kernel's inner loop): 0 = no hit,
// 1 = hit, 2 = within r but near it, needs the assignment tests (settled
// later by face_hit_at on the same particle).  d2 is the squared distance.
__device__ __forceinline__ int face_probe(const FaceCtx &c, F3 p0, float &d2) {
    const F3 p = p0 - c.off;
    const F3 pl = p - c.boff;
    const F3 v = c.gpos - pl;
    d2 = dot3(v, v);
    return !(d2 < c.rsq) ? 0 : (d2 < c.rsq_fast ? 1 : 2);
}

// The same for the particle in sorted slot q, with its w v
// (`scalar += weight * velocity`, velocityadvector.cpp:438).
__device__ __forceinline__ bool face_hit(const P2GArgs &a, const FaceCtx &c, int q, float &prod, float &wt) {
    const float4 p4 = a.sp[q];
    if (!face_hit_at(a, c, f3(p4.x, p4.y, p4.z), wt)) return false;
    prod = wt * a.sv[c.comp][q];
    return true;
}

// The particle id stored in the fourth word of sorted slot q.
__device__ __forceinline__ int sorted_id(const P2GArgs &a, int q) {
    return reinterpret_cast<const int *>(a.sp)[4 * (size_t)q + 3];
}

// Calls fn(q) for every candidate slot q of the face, lanes of a wavefront
// striding each run when `stride` is 64 (the overflow kernel) or one thread
// walking it when `stride` is 1.
template <class Fn>
__device__ __forceinline__ void for_candidates(const P2GArgs &a, const FaceCtx &c, int first, int stride, Fn &&fn) {
    const size_t SI = 2 * (size_t)a.g.I, SJ = 2 * (size_t)a.g.J;
    for (int sz = c.lo[2]; sz <= c.hi[2]; sz++) {
        for (int sy = c.lo[1]; sy <= c.hi[1]; sy++) {
            const size_t row = SI * ((size_t)sy + SJ * (size_t)sz);
            const int b = a.sub_offset[row + (size_t)c.lo[0]];
            const int e = a.sub_offset[row + (size_t)c.hi[0] + 1];
            for (int q = b + first; q < e; q += stride) fn(q);
        }
    }
}

__device__ __forceinline__ void write_face(const P2GArgs &a, const FaceCtx &c, float sum, float wsum) {
    const float eps = 1e-6f;  // velocityadvector.cpp:440-446
    float u = sum;
    if (wsum > eps) u = sum / wsum;
    a.out[c.comp][c.f] = u;
    a.valid[c.comp][c.f] = wsum > eps ? 1 : 0;
}

// Sort key of a hit: particle id above the LDS slot, so keys order by id.
template <class Key>
struct KeyTraits;
template <>
struct KeyTraits<uint32_t> {
    static constexpr int kSlotBits = 5;  // kCap = 32 slots; ids below 2^27
    __device__ static uint32_t make(int id, int slot) { return ((uint32_t)id << kSlotBits) | (uint32_t)slot; }
    __device__ static int slot(uint32_t k) { return (int)(k & ((1u << kSlotBits) - 1u)); }
    static constexpr uint32_t kMax = 0xffffffffu;
};
template <>
struct KeyTraits<uint64_t> {
    __device__ static uint64_t make(int id, int slot) { return ((uint64_t)(uint32_t)id << 32) | (uint32_t)slot; }
    __device__ static int slot(uint64_t k) { return (int)(uint32_t)k; }
    static constexpr uint64_t kMax = ~0ull;
};

// Ascending bitonic sort of k[0..N); every index is a compile-time constant
// after unrolling, so the network lives in registers.
template <int N, class Key>
__device__ __forceinline__ void bitonic_sort(Key (&k)[N]) {
#pragma unroll
    for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
            for (int i = 0; i < N; i++) {
                const int j = i ^ stride;
                if (j > i) {
                    const Key x = k[i], y = k[j];
                    const Key lo = x < y ? x : y, hi = x < y ? y : x;
                    if ((i & size) == 0) { k[i] = lo; k[j] = hi; }
                    else { k[i] = hi; k[j] = lo; }
                }
            }
        }
    }
}

// The m hits of this thread (sorted slot q and weight w in LDS): the ids
// and velocities of all hits are loaded together, w v (the reference's float
// product) replaces q in LDS, the keys (particle id, hit slot) are sorted in
// registers, and the sums run in key order.
template <int N, class Key>
__device__ __forceinline__ void sorted_sum(const P2GArgs &a, const FaceCtx &c, int comp,
                                           int (&s_q)[kCap + 1][kThreads], float (&s_w)[kCap + 1][kThreads], int m,
                                           int dropped, float &sum, float &wsum) {
    // s_w holds each hit's squared distance on entry.  Slots with q = -1
    // (pairs that failed the assignment tests) sort last and are not summed.
    const int t = threadIdx.x;
    const float *vel = a.sv[comp];
    int id[N];
    float v[N];
#pragma unroll
    for (int s = 0; s < N; s++) {
        id[s] = -1;
        v[s] = 0.0f;
        if (s < m) {
            const int q = s_q[s][t];
            if (q >= 0) {
                id[s] = sorted_id(a, q);
                v[s] = vel[q];
            }
        }
    }
    Key k[N];
#pragma unroll
    for (int s = 0; s < N; s++) {
        k[s] = KeyTraits<Key>::kMax;
        if (s < m && id[s] >= 0) {
            k[s] = KeyTraits<Key>::make(id[s], s);
            const float wt = face_weight(c, s_w[s][t]);
            s_q[s][t] = __float_as_int(wt * v[s]);
            s_w[s][t] = wt;
        }
    }
    bitonic_sort<N>(k);
    const int kept = m - dropped;
#pragma unroll
    for (int s = 0; s < N; s++) {
        if (s < kept) {
            const int slot = KeyTraits<Key>::slot(k[s]);
            sum += __int_as_float(s_q[slot][t]);
            wsum += s_w[slot][t];
        }
    }
}

// Work tiles: 32 faces along i by 2 along j, at one k, of one component.
// The U, V and W tiles of the same place are adjacent in tile order, and the
// blocks of each XCD (workgroups go to the 8 XCDs round-robin) take one
// contiguous eighth of the tiles, so neighbouring faces -- which read the
// same particles -- share an L2.
struct TileMap {
    int ntc[3];  // i-tiles per component
    int per_group, njp, ngroups;
    unsigned ntiles, nblocks;
};

inline TileMap make_tile_map(const GridGeom &g) {
    TileMap m;
    m.ntc[0] = (g.I + 1 + 31) / 32;
    m.ntc[1] = (g.I + 31) / 32;
    m.ntc[2] = m.ntc[1];
    m.per_group = m.ntc[0] + m.ntc[1] + m.ntc[2];
    m.njp = (g.J + 2) / 2;  // pairs of j covering J + 1 rows (V)
    m.ngroups = m.njp * (g.K + 1);
    m.ntiles = (unsigned)m.ngroups * (unsigned)m.per_group;
    m.nblocks = (m.ntiles + 7u) / 8u * 8u;
    return m;
}

// This block's face, or false when its lane has none.
__device__ __forceinline__ bool tile_face(const GridGeom &g, const TileMap &m, int &comp, size_t &f) {
    const unsigned b = blockIdx.x;
    const unsigned tile = (b & 7u) * (m.nblocks >> 3) + (b >> 3);
    if (tile >= m.ntiles) return false;
    const int group = (int)(tile / (unsigned)m.per_group);
    int local = (int)(tile % (unsigned)m.per_group);
    comp = local < m.ntc[0] ? 0 : (local < m.ntc[0] + m.ntc[1] ? 1 : 2);
    local -= comp == 0 ? 0 : (comp == 1 ? m.ntc[0] : m.ntc[0] + m.ntc[1]);
    const int jp = group % m.njp, k = group / m.njp;
    const int i = local * 32 + (int)(threadIdx.x & 31), j = 2 * jp + (int)(threadIdx.x >> 5);
    int ni, nj, nk;
    face_dims(g, comp, ni, nj, nk);
    if (i >= ni || j >= nj || k >= nk) return false;
    f = (size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * (size_t)k);
    return true;
}

template <class Key>
__global__ __launch_bounds__(kThreads) void k_p2g(P2GArgs a, TileMap tm) {
    // Per thread: the hits' sorted slots and squared distances; slot kCap
    // takes the stores of non-hits and of hits beyond the cap.
    __shared__ int s_q[kCap + 1][kThreads];
    __shared__ float s_w[kCap + 1][kThreads];
    int comp = 0;
    size_t f = 0;
    const bool active = tile_face(a.g, tm, comp, f);
    const int t = threadIdx.x;
    FaceCtx c;
    int nh = 0, dropped = 0;
    if (active) {
        face_setup(a, comp, f, c);
        // Candidate ranges of the 4 rows (sy) at half-cell plane sz.
        const size_t SI = 2 * (size_t)a.g.I, SJ = 2 * (size_t)a.g.J;
        auto plane_ranges = [&](int sz, int (&b)[4], int (&e)[4]) {
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const int sy = c.lo[1] + rr;
                b[rr] = e[rr] = 0;
                if (sy <= c.hi[1] && sz <= c.hi[2]) {
                    const size_t row = SI * ((size_t)sy + SJ * (size_t)sz);
                    b[rr] = a.sub_offset[row + (size_t)c.lo[0]];
                    e[rr] = a.sub_offset[row + (size_t)c.hi[0] + 1];
                }
            }
        };
        // Every candidate stores (q, d2) at slot min(nh, kCap) and nh counts
        // only hits, so no branch depends on the outcome.  Pairs near r go in
        // as -(q + 2) and are settled after the loop, by one copy of the
        // assignment tests.
        bool pending = false;
        auto consider = [&](float4 p4, int q, bool valid) {
            float d2;
            int h = face_probe(c, f3(p4.x, p4.y, p4.z), d2);
            h = valid ? h : 0;
            const int slot = nh < kCap ? nh : kCap;
            s_q[slot][t] = h == 1 ? q : -(q + 2);
            s_w[slot][t] = d2;
            pending |= h == 2;
            nh += h != 0;
        };
        // One half-cell plane (4 rows) per iteration: the first four
        // candidates of each row are loaded together (16 loads in flight)
        // while the next plane's ranges are fetched, then tested; rows with
        // more candidates (denser than one particle per half-cell) finish one
        // by one.  A plane no lane of the wavefront has candidates in is
        // skipped.
        int cb[4], ce[4], nb[4], ne[4];
        plane_ranges(c.lo[2], cb, ce);
#pragma unroll 1
        for (int pl = 0; pl < 4; pl++) {
            const bool mine = cb[0] < ce[0] || cb[1] < ce[1] || cb[2] < ce[2] || cb[3] < ce[3];
            plane_ranges(c.lo[2] + pl + 1, nb, ne);
            if (__ballot(mine) != 0ull) {
                float4 p[4][4];
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        p[rr][u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        if (cb[rr] + u < ce[rr]) p[rr][u] = a.sp[cb[rr] + u];
                    }
                }
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
#pragma unroll
                    for (int u = 0; u < 4; u++) consider(p[rr][u], cb[rr] + u, cb[rr] + u < ce[rr]);
                }
#pragma unroll
                for (int rr = 0; rr < 4; rr++)
                    for (int q = cb[rr] + 4; q < ce[rr]; q++) consider(a.sp[q], q, true);
            }
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                cb[rr] = nb[rr];
                ce[rr] = ne[rr];
            }
        }
        if (nh > kCap) {
            const int e = atomicAdd(a.ovf_count, 1);
            a.ovf_list[e] = (comp << kCompShift) | (int)f;
        } else if (pending) {
            // Settle the pairs near r; the ones that fail become -1 and are
            // left out of the sums.
            for (int s = 0; s < nh; s++) {
                const int qe = s_q[s][t];
                if (qe >= -1) continue;
                const int q = -qe - 2;
                float prod, wt;
                if (face_hit(a, c, q, prod, wt)) {
                    s_q[s][t] = q;  // its squared distance stays in s_w
                } else {
                    s_q[s][t] = -1;
                    dropped++;
                }
            }
        }
    }
    const int m = nh <= kCap ? nh : 0;
    int wm = m;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) wm = max(wm, __shfl_xor(wm, off, 64));
    float sum = 0.0f, wsum = 0.0f;
    if (wm > 16) sorted_sum<32, Key>(a, c, comp, s_q, s_w, m, dropped, sum, wsum);
    else if (wm > 8) sorted_sum<16, Key>(a, c, comp, s_q, s_w, m, dropped, sum, wsum);
    else if (wm > 4) sorted_sum<8, Key>(a, c, comp, s_q, s_w, m, dropped, sum, wsum);
    else if (wm > 0) sorted_sum<4, Key>(a, c, comp, s_q, s_w, m, dropped, sum, wsum);
    if (active && nh <= kCap) write_face(a, c, sum, wsum);
}

// One wavefront per listed face: the hits go to LDS keyed (id, slot), are
// sorted there, and lane 0 sums them in order.  Beyond kOvfCap hits the
// wavefront instead selects the next-larger id one at a time, rescanning the
// candidates (only reachable with hundreds of particles around one face).
__global__ __launch_bounds__(64) void k_p2g_overflow(P2GArgs a) {
    __shared__ uint64_t s_k[kOvfCap];
    __shared__ float2 s_v[kOvfCap];
    __shared__ int s_n;
    const int lane = threadIdx.x;
    const int count = *a.ovf_count;
    for (int e = blockIdx.x; e < count; e += gridDim.x) {
        const int code = a.ovf_list[e];
        FaceCtx c;
        face_setup(a, code >> kCompShift, (size_t)(code & ((1 << kCompShift) - 1)), c);
        if (lane == 0) s_n = 0;
        __syncthreads();
        for_candidates(a, c, lane, 64, [&](int q) {
            float prod, wt;
            if (!face_hit(a, c, q, prod, wt)) return;
            const int pos = atomicAdd(&s_n, 1);
            if (pos < kOvfCap) {
                s_k[pos] = ((uint64_t)(uint32_t)sorted_id(a, q) << 32) | (uint32_t)pos;
                s_v[pos] = make_float2(prod, wt);
            }
        });
        __syncthreads();
        const int n = s_n;
        float sum = 0.0f, wsum = 0.0f;
        if (n <= kOvfCap) {
            int P = 1;
            while (P < n) P <<= 1;
            for (int x = n + lane; x < P; x += 64) s_k[x] = ~0ull;
            __syncthreads();
            for (int size = 2; size <= P; size <<= 1) {
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    for (int x = lane; x < P; x += 64) {
                        const int y = x ^ stride;
                        if (y > x) {
                            const uint64_t u = s_k[x], w = s_k[y];
                            const bool up = (x & size) == 0;
                            if ((u > w) == up) { s_k[x] = w; s_k[y] = u; }
                        }
                    }
                    __syncthreads();
                }
            }
            if (lane == 0) {
                for (int x = 0; x < n; x++) {
                    const float2 v = s_v[(uint32_t)s_k[x]];
                    sum += v.x;
                    wsum += v.y;
                }
            }
        } else {
            long long last = -1;
            for (;;) {
                long long best = 0x7fffffffffffffffll;
                float bp = 0.0f, bw = 0.0f;
                for_candidates(a, c, lane, 64, [&](int q) {
                    const long long id = sorted_id(a, q);
                    if (id <= last || id >= best) return;
                    float prod, wt;
                    if (!face_hit(a, c, q, prod, wt)) return;
                    best = id; bp = prod; bw = wt;
                });
                long long wbest = best;
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) {
                    const long long o = __shfl_xor(wbest, off, 64);
                    wbest = o < wbest ? o : wbest;
                }
                if (wbest == 0x7fffffffffffffffll) break;
                const unsigned long long who = __ballot(best == wbest);
                const int src = __ffsll((long long)who) - 1;
                const float p = __shfl(bp, src, 64), w = __shfl(bw, src, 64);
                sum += p;  // every lane keeps the same running sums
                wsum += w;
                last = wbest;
            }
        }
        if (lane == 0) write_face(a, c, sum, wsum);
        __syncthreads();
    }
}

hipError_t launch_p2g(const GridGeom &g, const float4 *spos, const float *const svel[3], const int *sub_offset,
                      int n, double radius, float *const out[3], uint8_t *const valid[3], int *ovf_count,
                      int *ovf_list, hipStream_t s) {
    size_t nmax = 0;
    for (int comp = 0; comp < 3; comp++) {
        const size_t nf = (size_t)(g.I + (comp == 0)) * (g.J + (comp == 1)) * (g.K + (comp == 2));
        nmax = nf > nmax ? nf : nmax;
    }
    if (nmax >= ((size_t)1 << kCompShift)) return hipErrorInvalidValue;  // overflow list encoding
    const TileMap tm = make_tile_map(g);
    hipError_t e = launch_fill_i32(ovf_count, 1, 0, s);
    if (e != hipSuccess) return e;
    P2GArgs a{};
    a.g = g;
    a.sp = spos;
    for (int c = 0; c < 3; c++) {
        a.sv[c] = svel[c];
        a.out[c] = out[c];
        a.valid[c] = valid[c];
    }
    a.sub_offset = sub_offset;
    a.radius_f = (float)radius;
    a.radius = radius;
    a.ovf_count = ovf_count;
    a.ovf_list = ovf_list;
    if (n < (1 << (32 - KeyTraits<uint32_t>::kSlotBits)))
        k_p2g<uint32_t><<<tm.nblocks, kThreads, 0, s>>>(a, tm);
    else
        k_p2g<uint64_t><<<tm.nblocks, kThreads, 0, s>>>(a, tm);
    k_p2g_overflow<<<kOvfBlocks, 64, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace flipmi
