// Counting sort of particles by cell and an exclusive prefix scan.
//
// The sorted order groups particle ids by cell (i fastest, as Array3d) and
// keeps ids ascending inside a cell, so every later gather over a cell's
// particles visits them in the same order on every run: the step is
// deterministic although slots are claimed with atomics.
#include "device_math.h"
#include "flipmi_kernels.h"

namespace flipmi {

constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 4;
constexpr int kScanChunk = kScanThreads * kScanPerThread;

// Inclusive scan of one value per lane across a 64-wide wavefront.
__device__ __forceinline__ int wave_inclusive_scan(int v) {
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Block-wide exclusive scan of kScanPerThread consecutive elements per thread.
__global__ void k_scan_chunk(const int *__restrict__ in, int *out, int n, int *block_sums) {
    __shared__ int wave_tot[kScanThreads / 64];
    size_t base = (size_t)blockIdx.x * kScanChunk + (size_t)threadIdx.x * kScanPerThread;
    int v[kScanPerThread];
    int local = 0;
#pragma unroll
    for (int q = 0; q < kScanPerThread; q++) {
        size_t idx = base + q;
        v[q] = idx < (size_t)n ? in[idx] : 0;
        local += v[q];
    }
    int inc = wave_inclusive_scan(local);
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) wave_tot[wave] = inc;
    __syncthreads();
    int wave_off = 0;
    for (int w = 0; w < wave; w++) wave_off += wave_tot[w];
    int run = wave_off + inc - local;
#pragma unroll
    for (int q = 0; q < kScanPerThread; q++) {
        size_t idx = base + q;
        if (idx < (size_t)n) out[idx] = run;
        run += v[q];
    }
    if (threadIdx.x == kScanThreads - 1 && block_sums) block_sums[blockIdx.x] = wave_off + inc;
}

__global__ void k_scan_add(int *out, int n, const int *__restrict__ block_offsets) {
    size_t idx = (size_t)blockIdx.x * kScanChunk + threadIdx.x;
    int add = block_offsets[blockIdx.x];
#pragma unroll
    for (int q = 0; q < kScanPerThread; q++) {
        size_t e = idx + (size_t)q * kScanThreads;
        if (e < (size_t)n) out[e] += add;
    }
}

size_t scan_scratch_elems(int n) {
    size_t total = 0;
    size_t m = (size_t)n;
    while (m > 1) {
        m = (m + kScanChunk - 1) / kScanChunk;
        total += 2 * m + 1;
    }
    return total + 8;
}

// Exclusive scan in[0..n) -> out[0..n); out[n] is not written.
hipError_t launch_exclusive_scan(const int *in, int *out, int n, int *tmp, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int nb = (n + kScanChunk - 1) / kScanChunk;
    if (nb == 1) {
        k_scan_chunk<<<1, kScanThreads, 0, s>>>(in, out, n, nullptr);
        return hipGetLastError();
    }
    int *sums = tmp;
    int *sums_scanned = tmp + nb;
    k_scan_chunk<<<nb, kScanThreads, 0, s>>>(in, out, n, sums);
    hipError_t e = launch_exclusive_scan(sums, sums_scanned, nb, tmp + 2 * (size_t)nb + 1, s);
    if (e != hipSuccess) return e;
    k_scan_add<<<nb, kScanThreads, 0, s>>>(out, n, sums_scanned);
    return hipGetLastError();
}

// Cell of each particle and its arrival slot in that cell.  Particles that
// are neighbours in memory mostly share cells, so the lanes of a wave that
// hold the same cell take their slots with one atomic (the leader adds the
// group's size, each lane its rank in the group): a cell's count is then
// bumped once per wave instead of once per particle.  The groups are found
// first (ballots only), then every leader takes its atomic in the same
// instruction, so a wave waits for one atomic round trip, not one per cell.
// Slots only need to be distinct; the within-cell order is fixed afterwards
// by particle id.
__global__ __launch_bounds__(256) void k_cell_key(GridGeom g, const float *__restrict__ px,
                                                  const float *__restrict__ py, const float *__restrict__ pz,
                                                  int n, int *cell, int *slot, int *count, int *over) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int id = -1;
    if (i < n) {
        I3 c = pos_to_index(f3(px[i], py[i], pz[i]), g.invdx);
        if (c.i >= 0 && c.j >= 0 && c.k >= 0 && c.i < g.I && c.j < g.J && c.k < g.K)
            id = c.i + g.I * (c.j + g.J * c.k);
    }
    int my_leader = lane, rank = 0, size = 0;  // size: the group's, on its leader
    bool pending = id >= 0;
    for (;;) {
        const unsigned long long m = __ballot(pending);
        if (m == 0ull) break;
        const int leader = __ffsll((long long)m) - 1;
        const int key = __shfl(id, leader, 64);
        const bool mine = pending && id == key;
        const unsigned long long peers = __ballot(mine);
        if (mine) {
            my_leader = leader;
            rank = __popcll(peers & ((1ull << lane) - 1ull));
            if (lane == leader) size = __popcll(peers);
            pending = false;
        }
    }
    int base = 0;
    if (size > 0) {
        base = atomicAdd(&count[id], size);
        if (over && base + size > 250) atomicOr(over, 1);
    }
    const int sl = __shfl(base, my_leader, 64) + rank;
    if (i < n) {
        cell[i] = id;
        slot[i] = sl;
    }
}

__global__ void k_scatter(const int *__restrict__ cell, const int *__restrict__ slot,
                          const int *__restrict__ offset, int n, const int *over, int *order) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (over && *over == 0)) return;
    int c = cell[i];
    if (c < 0) return;
    order[offset[c] + slot[i]] = i;
}

// Ascending ids inside each cell (cells hold a handful of particles).
__global__ void k_sort_within_cell(const int *__restrict__ offset, int ncell, const int *over, int *order) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell || (over && *over == 0)) return;
    int b = offset[c], e = offset[c + 1];
    for (int p = b + 1; p < e; p++) {
        int v = order[p];
        int q = p - 1;
        while (q >= b && order[q] > v) {
            order[q + 1] = order[q];
            q--;
        }
        order[q + 1] = v;
    }
}

__global__ void k_set_last(int *offset, int ncell, const int *__restrict__ count) {
    offset[ncell] = offset[ncell - 1] + count[ncell - 1];
}

hipError_t launch_sort_by_cell(const GridGeom &g, const float *px, const float *py,
                               const float *pz, int n, SortBuffers &sb, hipStream_t s) {
    int ncell = g.I * g.J * g.K;
    hipError_t e = launch_fill_i32(sb.count, (size_t)ncell, 0, s);
    if (e != hipSuccess) return e;
    if (sb.over && (e = launch_fill_i32(sb.over, 1, 0, s)) != hipSuccess) return e;
    if (n > 0)
        k_cell_key<<<(n + 255) / 256, 256, 0, s>>>(g, px, py, pz, n, sb.cell, sb.slot, sb.count, sb.over);
    e = launch_exclusive_scan(sb.count, sb.offset, ncell, sb.scan_tmp, s);
    if (e != hipSuccess) return e;
    k_set_last<<<1, 1, 0, s>>>(sb.offset, ncell, sb.count);
    if (n > 0) k_scatter<<<(n + 255) / 256, 256, 0, s>>>(sb.cell, sb.slot, sb.offset, n, sb.over, sb.order);
    k_sort_within_cell<<<(ncell + 255) / 256, 256, 0, s>>>(sb.offset, ncell, sb.over, sb.order);
    return hipGetLastError();
}

// Half-cell ("sub-cell") bin of a position: bins are 2I x 2J x 2K, sx
// fastest.  The cell part is Grid3d::positionToGridIndex in the same double
// value the cell sort uses, and the half along each axis is whether that
// value's fraction is at least 1/2, so a bin's cell is exactly (sx>>1,
// sy>>1, sz>>1).  -1 outside the grid.
__device__ __forceinline__ int subcell_key(const GridGeom &g, float x, float y, float z) {
    const double tx = (double)x * g.invdx, ty = (double)y * g.invdx, tz = (double)z * g.invdx;
    const double fx = floor(tx), fy = floor(ty), fz = floor(tz);
    const int cx = (int)fx, cy = (int)fy, cz = (int)fz;
    if (cx < 0 || cy < 0 || cz < 0 || cx >= g.I || cy >= g.J || cz >= g.K) return -1;
    const int sx = 2 * cx + (tx - fx >= 0.5 ? 1 : 0);
    const int sy = 2 * cy + (ty - fy >= 0.5 ? 1 : 0);
    const int sz = 2 * cz + (tz - fz >= 0.5 ? 1 : 0);
    return sx + 2 * g.I * (sy + 2 * g.J * sz);
}

// Bin and arrival slot per particle.  Neighbouring particles rarely share a
// half-cell, so each lane takes its slot with its own atomic.
__global__ __launch_bounds__(256) void k_subcell_key(GridGeom g, const float *__restrict__ px,
                                                     const float *__restrict__ py,
                                                     const float *__restrict__ pz, int n, int *key, int *slot,
                                                     int *count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = subcell_key(g, px[i], py[i], pz[i]);
    key[i] = b;
    slot[i] = b >= 0 ? atomicAdd(&count[b], 1) : 0;
}

// Counting sort of particles into half-cell bins (order within a bin is the
// arrival order of the atomics; every consumer of this sort -- the liquid
// SDF's minimum and P2G's per-face sort by particle id -- is independent of
// it).  sb.count / sb.offset hold 8 I J K (+1) entries.
hipError_t launch_sort_by_subcell(const GridGeom &g, const float *px, const float *py, const float *pz, int n,
                                  SortBuffers &sb, hipStream_t s) {
    const size_t nbins = subcell_bins(g);
    if (nbins >= (size_t)INT32_MAX) return hipErrorInvalidValue;
    hipError_t e = launch_fill_i32(sb.count, nbins, 0, s);
    if (e != hipSuccess) return e;
    if (n > 0) k_subcell_key<<<(n + 255) / 256, 256, 0, s>>>(g, px, py, pz, n, sb.cell, sb.slot, sb.count);
    e = launch_exclusive_scan(sb.count, sb.offset, (int)nbins, sb.scan_tmp, s);
    if (e != hipSuccess) return e;
    k_set_last<<<1, 1, 0, s>>>(sb.offset, (int)nbins, sb.count);
    if (n > 0) k_scatter<<<(n + 255) / 256, 256, 0, s>>>(sb.cell, sb.slot, sb.offset, n, nullptr, sb.order);
    return hipGetLastError();
}

// The count grid of FluidSimulation::_removeMarkerParticles
// (fluidsimulation.cpp:4325-4343): particles are counted in order, skipping
// those removed by the solid test, and a cell stops counting at the cap
// (later particles of that cell are removed before they are added).  So a
// cell's entry is min(particles in the cell - solid-removed ones, cap),
// whatever the order.  `count` is the cell histogram of all particles from
// the sort; the solid-removed particles are subtracted per cell.
__global__ void k_removed_in_cell(GridGeom g, const float *__restrict__ px, const float *__restrict__ py,
                                  const float *__restrict__ pz, const uint8_t *__restrict__ removed, int n,
                                  int *removed_in_cell) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !removed[i]) return;
    I3 c = pos_to_index(f3(px[i], py[i], pz[i]), g.invdx);
    if (c.i < 0 || c.j < 0 || c.k < 0 || c.i >= g.I || c.j >= g.J || c.k >= g.K) return;
    atomicAdd(&removed_in_cell[c.i + g.I * (c.j + g.J * c.k)], 1);
}

__global__ void k_removal_counts(const int *__restrict__ count, const int *__restrict__ removed_in_cell,
                                 size_t ncell, int cap, int *out) {
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int m = count[c] - removed_in_cell[c];
    out[c] = m < cap ? m : cap;
}

hipError_t launch_removed_in_cell(const GridGeom &g, const float *px, const float *py, const float *pz,
                                  const uint8_t *removed, int n, int *removed_in_cell, hipStream_t s) {
    hipError_t e = launch_fill_i32(removed_in_cell, (size_t)g.I * g.J * g.K, 0, s);
    if (e != hipSuccess || n <= 0) return e;
    k_removed_in_cell<<<(n + 255) / 256, 256, 0, s>>>(g, px, py, pz, removed, n, removed_in_cell);
    return hipGetLastError();
}

hipError_t launch_removal_counts(const GridGeom &g, const int *count, const int *removed_in_cell, int cap,
                                 int *out, hipStream_t s) {
    size_t nc = (size_t)g.I * g.J * g.K;
    k_removal_counts<<<(unsigned)((nc + 255) / 256), 256, 0, s>>>(count, removed_in_cell, nc, cap, out);
    return hipGetLastError();
}

// Sorted copies: the position with the particle's id in the fourth word
// (one 16-byte load per candidate for the gathers), and the velocity.
__global__ void k_gather_sorted(const int *__restrict__ order, const int *__restrict__ nsorted,
                                const float *__restrict__ px, const float *__restrict__ py,
                                const float *__restrict__ pz, const float *__restrict__ vx,
                                const float *__restrict__ vy, const float *__restrict__ vz, float4 *spos,
                                float *su, float *sv, float *sw) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nsorted) return;  // particles outside the grid are not in the order
    int o = order[i];
    spos[i] = make_float4(px[o], py[o], pz[o], __int_as_float(o));
    su[i] = vx[o]; sv[i] = vy[o]; sw[i] = vz[o];
}

hipError_t launch_gather_sorted(const int *order, const int *nsorted, int n, const float *const src[6],
                                float4 *spos, float *const svel[3], hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_gather_sorted<<<(n + 255) / 256, 256, 0, s>>>(order, nsorted, src[0], src[1], src[2], src[3], src[4],
                                                    src[5], spos, svel[0], svel[1], svel[2]);
    return hipGetLastError();
}

}  // namespace flipmi
