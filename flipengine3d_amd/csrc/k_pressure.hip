// Pressure projection: assembly of the variational 7-point Poisson system,
// MIC(0)-preconditioned conjugate gradients and the pressure-gradient update.
//
// References: PressureSolver (pressuresolver.cpp:44-1047) and
// PCGSolver / factorModifiedIncompleteColesky0 / solveLower /
// solveLowerTransposeInPlace (pcgsolver/pcgsolver.h:69-302).
//
// Layout: every vector and matrix array is stored in slot order (SlotMap in
// flipmi_kernels.h): 8 x 8 (j, k) column tiles, and within a tile one row of
// 64 slots per wavefront step t = i + (j % 8) + (k % 8).  A cell is a row of
// the system when it is interior and its liquid phi is negative; every other
// slot holds zeros, so stencils need no row lookup.
//
// The MIC(0) factorisation and both triangular solves run as a pipelined
// wavefront: one 64-lane wave owns a tile and marches over its steps, so each
// step reads one contiguous row of every array, and every cell's -x, -y and
// -z dependencies were produced one step earlier by itself or a neighbouring
// lane (exchanged with lane shuffles).  Values a tile needs from its upstream
// neighbour tiles arrive as tagged records (value + launch epoch), so a
// consumer polls the data itself and a producer never waits for its stores.
// Each cell performs exactly the reference's operations in the reference's
// order, so the preconditioner is the reference's bit for bit.
#include <algorithm>

#include <hip/hip_ext.h>

#include "device_math.h"
#include "flipmi_kernels.h"

namespace flipmi {

constexpr int kTile = 8;
constexpr int kRedBlocks = 1024;
// The axpy's grid: fewer, longer-lived blocks stream its four arrays faster
// (mixed axpy on the 256x128x128 double dam break: 20.6 us per launch with
// 1024 blocks, 15.7 with 512, 15.0 with 384, 16.1 with 256).  Its reduction
// is a maximum, so the grid does not change any result; the SpMV's is a sum
// whose order follows the grid, so it keeps kRedBlocks.  A multiple of kXcds.
constexpr int kAxpyBlocks = 384;
constexpr int kRedThreads = 256;
constexpr double kMaxTheta = 25.0;  // pressuresolver.h:163
constexpr unsigned kSpinLimit = 1u << 20;  // ~1 s of polling: a legitimate wait is microseconds
constexpr int kRecWords = 2;       // 64-bit words per tagged FP64 record

// Records per launch: for every tile and march step, 16 slots -- the values
// of the tile's 8 cells on its j edge and of the 7 further cells on its k
// edge (the corner cell's one value serves both neighbours).
constexpr int kRecSlots = 16;

size_t edge_record_count(const SlotMap &sm) {
    return (size_t)sm.ntj * sm.ntk * sm.S * kRecSlots * kRecWords;
}

// Cell of a slot (may lie outside the grid: skew corners and padding), in
// 32-bit arithmetic: two 32-bit divisions instead of 64-bit ones, which cost
// several times the instructions (slot indices stay below 2^32).
__device__ __forceinline__ void decode_slot(const SlotMap &sm, size_t p, int &i, int &j, int &k) {
    const unsigned q = (unsigned)p;
    const unsigned grp = q >> 8, S4 = (unsigned)sm.S >> 2, ntj = (unsigned)sm.ntj;
    const unsigned tile = grp / S4, kt = tile / ntj;
    const int t = (int)((grp - tile * S4) * 4u + (q & 3u));
    const int lane = (int)((q >> 2) & 63u), a = lane & 7, b = lane >> 3;
    j = (int)(tile - kt * ntj) * kTile + a;
    k = (int)kt * kTile + b;
    i = t - a - b;
}

// A coupling entry of the matrix from the face weight it is formed from:
// -weight * dt / dx^2, the product the assembly takes (pressuresolver.cpp:
// 725, 753-798), so the FP64 entry comes out bit for bit.  A slot without
// the coupling holds -0.0f, which gives the +0.0 it stands for.
__device__ __forceinline__ double coupling(float w, double mf) { return -(double)w * mf; }

__device__ __forceinline__ bool interior(const GridGeom &g, int i, int j, int k) {
    return i >= 1 && j >= 1 && k >= 1 && i <= g.I - 2 && j <= g.J - 2 && k <= g.K - 2;
}

// Cell of slot p for a stencil.  A slot off the interior (skew corner,
// padding, wall layer) is moved to cell (1, 1, 1), so that every neighbour
// slot it reads is in range; the caller discards its result.
__device__ __forceinline__ bool stencil_cell(const GridGeom &g, const SlotMap &sm, size_t p, int &i, int &j,
                                             int &k) {
    decode_slot(sm, p, i, j, k);
    const bool in = interior(g, i, j, k);
    if (!in) {
        i = 1;
        j = 1;
        k = 1;
    }
    return in;
}

// Slot groups of this block in the per-group kernels.  Blocks are
// dispatched to the eight XCDs round-robin (block b on XCD b % 8); the
// blocks of one XCD cover one contiguous eighth of the groups, so the
// neighbour groups a stencil reads (the adjacent steps of a tile, the
// adjacent tiles) sit in the same L2 instead of being fetched by every
// XCD.  Any other placement only costs locality.  gridDim.x is a multiple
// of kXcds.
constexpr unsigned kXcds = 8;
struct GroupRange {
    size_t begin, end, step;
};
__device__ __forceinline__ GroupRange xcd_groups(size_t ngroups) {
    const unsigned x = blockIdx.x % kXcds, l = blockIdx.x / kXcds;
    const size_t part = (ngroups + kXcds - 1) / kXcds;
    const size_t b = (size_t)x * part, e = b + part < ngroups ? b + part : ngroups;
    return GroupRange{b + l, e, gridDim.x / kXcds};
}

// Whether slot group grp (256 consecutive slots) holds a row.
__device__ __forceinline__ bool group_active(const PressureBuffers &pb, size_t grp) {
    if (!pb.group_mask) return true;
    const size_t S4 = (size_t)(pb.sm.S >> 2);
    const size_t tile = grp / S4;
    const int gi = (int)(grp % S4);
    return (pb.group_mask[tile * pb.mask_words + (gi >> 6)] >> (gi & 63)) & 1ull;
}

// Entries of the active group list for this block: as in xcd_groups, the
// blocks of one XCD take one contiguous eighth of the list.
struct ListRange {
    int begin, end, step;
};
__device__ __forceinline__ ListRange xcd_list(int n) {
    const int x = (int)(blockIdx.x % kXcds), l = (int)(blockIdx.x / kXcds);
    const int part = (n + (int)kXcds - 1) / (int)kXcds;
    const int b = x * part, e = b + part < n ? b + part : n;
    return ListRange{b + l, e, (int)(gridDim.x / kXcds)};
}
// This thread's slots in the U list entries of one round (entries e0,
// e0 + step, ...).  An entry past the block's end repeats the round's first
// entry: its values are loaded like the others and never stored, so every
// load of the round is issued before any result is used.
template <int U>
__device__ __forceinline__ void list_slots(const PressureBuffers &pb, const ListRange &lr, int e0,
                                           size_t (&p)[U], bool (&ok)[U]) {
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int e = e0 + u * lr.step;
        ok[u] = e < lr.end;
        p[u] = ((size_t)pb.active_list[ok[u] ? e : e0] << 8) + threadIdx.x;
    }
}

// ---------------------------------------------------------------------------
// Assembly
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool is_row_cell(const GridGeom &g, const float *liquid, int i, int j,
                                            int k) {
    if (!interior(g, i, j, k)) return false;
    return liquid[(size_t)i + (size_t)g.I * ((size_t)j + (size_t)g.J * k)] < 0;
}

// Right-hand side (pressuresolver.cpp:580-669) and matrix coefficients
// (pressuresolver.cpp:723-808) of the row at slot p; couplings to the
// +x/+y/+z rows are stored on this row, couplings to -x/-y/-z on the
// neighbour.  Every slot (rows, other cells, padding) is written.  A row
// raises the thread's max |rhs| (IEEE bits of a non-negative double order as
// integers) and its row count.
struct AssembleInputs {
    const float *liquid, *wU, *wV, *wW, *wC, *u, *v, *w, *sU, *sV, *sW;
};
__device__ __forceinline__ void assemble_slot(const GridGeom &g, double dt, const AssembleInputs &in,
                                              PressureBuffers &pb, size_t p, unsigned long long &bmax,
                                              unsigned &rows) {
    const float *__restrict__ liquid = in.liquid;
    const float *__restrict__ wU = in.wU, *__restrict__ wV = in.wV, *__restrict__ wW = in.wW;
    const float *__restrict__ wC = in.wC;
    const float *__restrict__ u = in.u, *__restrict__ v = in.v, *__restrict__ w = in.w;
    const float *__restrict__ sU = in.sU, *__restrict__ sV = in.sV, *__restrict__ sW = in.sW;
    int i, j, k;
    decode_slot(pb.sm, p, i, j, k);
    double rhs = 0, diag = 0;
    // Couplings to the +x/+y/+z rows as the face weight they are formed from
    // (entry = -weight * mf); -0.0f stands for "no coupling" (entry +0.0).
    float wx = -0.0f, wy = -0.0f, wz = -0.0f;
    const double mf = dt / (g.dx * g.dx);
    if (p == 0) pb.sc->mf = mf;
    if (is_row_cell(g, liquid, i, j, k)) {
        size_t c = (size_t)i + (size_t)g.I * ((size_t)j + (size_t)g.J * k);
        Dims DU{g.I + 1, g.J, g.K}, DV{g.I, g.J + 1, g.K}, DW{g.I, g.J, g.K + 1};
        double volCenter = wC[c];
        double volRight = wU[DU.at(i + 1, j, k)], volLeft = wU[DU.at(i, j, k)];
        double volTop = wV[DV.at(i, j + 1, k)], volBottom = wV[DV.at(i, j, k)];
        double volFront = wW[DW.at(i, j, k + 1)], volBack = wW[DW.at(i, j, k)];
        double factor = 1.0 / g.dx;
        double div = 0.0;
        div += -factor * volRight * u[DU.at(i + 1, j, k)];
        div += factor * volLeft * u[DU.at(i, j, k)];
        div += -factor * volTop * v[DV.at(i, j + 1, k)];
        div += factor * volBottom * v[DV.at(i, j, k)];
        div += -factor * volFront * w[DW.at(i, j, k + 1)];
        div += factor * volBack * w[DW.at(i, j, k)];
        div += factor * (volRight - volCenter) * sU[DU.at(i + 1, j, k)];
        div += -factor * (volLeft - volCenter) * sU[DU.at(i, j, k)];
        div += factor * (volTop - volCenter) * sV[DV.at(i, j + 1, k)];
        div += -factor * (volBottom - volCenter) * sV[DV.at(i, j, k)];
        div += factor * (volFront - volCenter) * sW[DW.at(i, j, k + 1)];
        div += -factor * (volBack - volCenter) * sW[DW.at(i, j, k)];
        rhs = div;

        double eps = 1e-9;
        double pc = liquid[c];
        diag = (volRight + volLeft + volTop + volBottom + volFront + volBack) * mf;
        auto side = [&](double vol, int a, int b, int d, float *off) {
            double pn = liquid[(size_t)a + (size_t)g.I * ((size_t)b + (size_t)g.J * d)];
            if (pn < 0.0) {
                // vol is a float weight widened: (float)vol is exact.
                if (off && is_row_cell(g, liquid, a, b, d)) *off = (float)vol;
            } else {
                double theta = pn / (pc + eps);
                theta = theta < -kMaxTheta ? -kMaxTheta : (theta > kMaxTheta ? kMaxTheta : theta);
                diag -= vol * mf * theta;
            }
        };
        side(volRight, i + 1, j, k, &wx);
        side(volLeft, i - 1, j, k, nullptr);
        side(volTop, i, j + 1, k, &wy);
        side(volBottom, i, j - 1, k, nullptr);
        side(volFront, i, j, k + 1, &wz);
        side(volBack, i, j, k - 1, nullptr);
        diag = diag > 0.0 ? diag : 0.0;
        const unsigned long long bits = (unsigned long long)__double_as_longlong(fabs(rhs));
        bmax = bits > bmax ? bits : bmax;
        rows++;
    }
    pb.rhs[p] = rhs;
    pb.diag[p] = diag;
    pb.wx[p] = wx;
    pb.wy[p] = wy;
    pb.wz[p] = wz;
}

// The assembly on a fixed grid of blocks (grid-stride over the slots).
// max|rhs| and the row count are order-independent; each block leaves its
// share in the partials (bits of the maximum in the first kAssembleBlocks
// words, the count after them) and k_assemble_total combines them, so no
// two blocks contend for one address (an atomic per row-holding wave cost
// 460 us per assembly of 1.2 M rows).
constexpr int kAssembleBlocks = 2048;
__global__ __launch_bounds__(256) void k_assemble(GridGeom g, double dt, AssembleInputs in, PressureBuffers pb) {
    __shared__ unsigned long long wbmax[4];
    __shared__ unsigned wrows[4];
    unsigned long long bmax = 0;
    unsigned rows = 0;
    const size_t n = pb.sm.count(), stride = (size_t)gridDim.x * blockDim.x;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride)
        assemble_slot(g, dt, in, pb, p, bmax, rows);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(bmax, off, 64);
        bmax = o > bmax ? o : bmax;
        rows += __shfl_xor(rows, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        wbmax[threadIdx.x >> 6] = bmax;
        wrows[threadIdx.x >> 6] = rows;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 4; q++) {
            bmax = wbmax[q] > bmax ? wbmax[q] : bmax;
            rows += wrows[q];
        }
        reinterpret_cast<unsigned long long *>(pb.partials)[blockIdx.x] = bmax;
        pb.partials[kAssembleBlocks + blockIdx.x] = (double)rows;
    }
}

__global__ __launch_bounds__(1024) void k_assemble_total(PressureBuffers pb, int nblocks) {
    __shared__ unsigned long long wbmax[16];
    __shared__ unsigned wrows[16];
    unsigned long long bmax = 0;
    unsigned rows = 0;
    for (int q = (int)threadIdx.x; q < nblocks; q += 1024) {
        const unsigned long long b = reinterpret_cast<const unsigned long long *>(pb.partials)[q];
        bmax = b > bmax ? b : bmax;
        rows += (unsigned)pb.partials[kAssembleBlocks + q];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(bmax, off, 64);
        bmax = o > bmax ? o : bmax;
        rows += __shfl_xor(rows, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        wbmax[threadIdx.x >> 6] = bmax;
        wrows[threadIdx.x >> 6] = rows;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 16; q++) {
            bmax = wbmax[q] > bmax ? wbmax[q] : bmax;
            rows += wrows[q];
        }
        pb.sc->absmax_bits = bmax;
        pb.sc->rows = rows;
    }
}

hipError_t launch_pressure_assemble(const GridGeom &g, double dt, const float *liquid,
                                    const float *wU, const float *wV, const float *wW,
                                    const float *wC, const float *u, const float *v,
                                    const float *w, const float *sU, const float *sV,
                                    const float *sW, PressureBuffers &pb, hipStream_t s) {
    // The partials hold kAssembleBlocks maxima and as many counts.
    if (pb.npartials < 2 * kAssembleBlocks) return hipErrorInvalidValue;
    const size_t n = pb.sm.count();
    const int blocks = (int)std::min<size_t>((n + 255) / 256, (size_t)kAssembleBlocks);
    const AssembleInputs in{liquid, wU, wV, wW, wC, u, v, w, sU, sV, sW};
    k_assemble<<<blocks, 256, 0, s>>>(g, dt, in, pb);
    k_assemble_total<<<1, 1024, 0, s>>>(pb, blocks);
    return hipGetLastError();
}

// One 256-thread block per slot group: a group is active when any of its
// slots has a nonzero matrix entry or right-hand side.
__global__ __launch_bounds__(256) void k_group_mask(PressureBuffers pb) {
    const size_t ngroups = pb.sm.count() >> 8;
    const size_t S4 = (size_t)(pb.sm.S >> 2);
    for (size_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const size_t p = (grp << 8) + threadIdx.x;
        const int nz = (pb.diag[p] != 0.0) | (pb.wx[p] != 0.0f) | (pb.wy[p] != 0.0f) |
                       (pb.wz[p] != 0.0f) | (pb.rhs[p] != 0.0);
        if (__syncthreads_or(nz) && threadIdx.x == 0) {
            const size_t tile = grp / S4;
            const int gi = (int)(grp % S4);
            atomicOr(&pb.group_mask[tile * pb.mask_words + (gi >> 6)], 1ull << (gi & 63));
        }
    }
}

// The active group list: the groups whose mask bit is set, ascending.  One
// block: per tile the number of active groups (popcounts of its mask
// words), an exclusive scan over the tiles, then each tile writes its group
// ids from its offset.  Groups of the last mask word past the tile's length
// are never set (the mask is cleared before it is built).
__global__ __launch_bounds__(1024) void k_active_list(PressureBuffers pb) {
    __shared__ int wsum[16];
    __shared__ int carry;
    const int ntiles = pb.sm.ntj * pb.sm.ntk, S4 = pb.sm.S >> 2, W = pb.mask_words;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int t0 = 0; t0 < ntiles; t0 += 1024) {
        const int t = t0 + (int)threadIdx.x;
        int c = 0;
        if (t < ntiles)
            for (int w = 0; w < W; w++) c += __popcll(pb.group_mask[(size_t)t * W + w]);
        int incl = c;  // inclusive scan within the wave
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int wbase = 0, total = 0;
        for (int q = 0; q < 16; q++) {
            const int v = wsum[q];
            wbase += q < wave ? v : 0;
            total += v;
        }
        int pos = carry + wbase + incl - c;
        if (t < ntiles) {
            for (int w = 0; w < W; w++) {
                unsigned long long m = pb.group_mask[(size_t)t * W + w];
                while (m) {
                    const int bit = __ffsll((long long)m) - 1;
                    m &= m - 1ull;
                    pb.active_list[pos++] = t * S4 + w * 64 + bit;
                }
            }
        }
        __syncthreads();  // every read of wsum and carry is done
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *pb.active_count = carry;
}

hipError_t launch_group_mask(PressureBuffers &pb, hipStream_t s) {
    // The PCG kernels walk the list built here: without a mask they would
    // see no rows.
    if (!pb.group_mask || !pb.active_list) return hipErrorInvalidValue;
    const size_t words = (size_t)pb.sm.ntj * pb.sm.ntk * pb.mask_words;
    hipError_t e = launch_fill_u8((uint8_t *)pb.group_mask, words * sizeof(unsigned long long), 0, s);
    if (e != hipSuccess) return e;
    k_group_mask<<<kRedBlocks * 4, 256, 0, s>>>(pb);
    k_active_list<<<1, 1024, 0, s>>>(pb);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Cross-workgroup exchange
// ---------------------------------------------------------------------------

enum { kFactor = 0, kForward = 1, kBackward = 2 };
// A sweep launch's mode: what it reduces or decides at its end.  Nothing;
// the z.r dot of the first or a later iteration (backward); the residual
// norm of the axpy the forward sweep did; or (forward, a launch that runs
// only when a true residual is due) the norm of the true residual it
// computes before sweeping.
enum { kDotNone = 0, kDotInit = 1, kDotIter = 2, kAxpyNorm = 3, kTrueResidual = 4 };

// Values another workgroup reads within the same kernel are stored and
// loaded as relaxed device-scope atomics, which are coherent across XCDs by
// themselves, so no acquire/release fence (and no L2 writeback/invalidate)
// is ever issued.  Ordering between data and a flag is obtained either by
// draining the wave's memory operations (s_waitcnt) before the flag store,
// or -- in the solves -- by tagging every value with the launch epoch.
__device__ __forceinline__ double ld_dev(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned ld_dev(const unsigned *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(unsigned *p, unsigned v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_dev(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_dev(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Waits until every memory operation this wave issued has completed.
__device__ __forceinline__ void drain_memory() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }

// Buffer views (raw buffer resources) over [base, base + bytes).  A load
// through an offset beyond the view returns zero and a store through one is
// dropped, so an offset of kOutside stands for "no such slot" at no cost.
constexpr unsigned kOutside = 0x80000000u;
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_view(const void *base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}

// Tagged records.  A record carries a value and the 32-bit epoch of the
// launch that wrote it, in 64-bit words written and read as single-copy-
// atomic accesses, so a reader that sees the current epoch in every word has
// the whole value.  FP32: one word (epoch | bits); FP64: two words holding
// 32 bits of the value each.
template <class T>
struct Record;
template <>
struct Record<double> {
    static constexpr int W = 2;
    __device__ static bool ok(const unsigned long long (&w)[2], unsigned tag) {
        return (int)((unsigned)(w[0] >> 32) == tag) & (int)((unsigned)(w[1] >> 32) == tag);
    }
    __device__ static double value(const unsigned long long (&w)[2]) {
        return __longlong_as_double((long long)((w[1] << 32) | (w[0] & 0xffffffffull)));
    }
    // Device-scope atomic stores: published where they are issued (the
    // compiler does not sink them towards the end of the chunk).
    __device__ static void put(unsigned long long *rec, double v, unsigned tag) {
        const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
        const unsigned long long t = (unsigned long long)tag << 32;
        st_dev(rec, t | (bits & 0xffffffffull));
        st_dev(rec + 1, t | (bits >> 32));
    }
};
template <>
struct Record<float> {
    static constexpr int W = 1;
    __device__ static bool ok(const unsigned long long (&w)[2], unsigned tag) {
        return (unsigned)(w[0] >> 32) == tag;
    }
    __device__ static float value(const unsigned long long (&w)[2]) {
        return __uint_as_float((unsigned)(w[0] & 0xffffffffull));
    }
    __device__ static void put(unsigned long long *rec, float v, unsigned tag) {
        st_dev(rec, ((unsigned long long)tag << 32) | (unsigned long long)__float_as_uint(v));
    }
};

// Waits until the upstream tiles (index < 0: none) have completed `need`
// steps; both flags are polled in one round trip.  Returns false if the
// bounded spin ran out or another wave aborted.
__device__ bool wait_progress2(PressureBuffers &pb, int up0, int up1, unsigned need) {
    unsigned spins = 0;
    for (;;) {
        unsigned p0 = up0 >= 0 ? ld_dev(&pb.progress[up0]) : need;
        unsigned p1 = up1 >= 0 ? ld_dev(&pb.progress[up1]) : need;
        if (p0 >= need && p1 >= need) break;
        if (ld_dev(&pb.sc->abort) != 0) return false;
        if (++spins > kSpinLimit) {
            st_dev(&pb.sc->abort, 1);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

__device__ __forceinline__ double wave_sum_fixed(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// Maximum over the wave; the operands are never NaN.
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// Tile order of the dispatch: an atomic ticket hands tiles out in
// dependency order, so every tile's producers were dispatched before it.
__device__ __forceinline__ int take_tile(PressureBuffers &pb, bool reverse, int &jt, int &kt) {
    const int ntj = pb.sm.ntj, ntk = pb.sm.ntk;
    int ticket = 0;
    if (threadIdx.x == 0) ticket = (int)atomicAdd(pb.ticket, 1u);
    // Wave-uniform from here on (kept in scalar registers, so every loop and
    // branch that depends on the tile is uniform).
    ticket = __builtin_amdgcn_readfirstlane(__shfl(ticket, 0, 64));
    if (reverse) {
        jt = ntj - 1 - ticket % ntj;
        kt = ntk - 1 - ticket / ntj;
    } else {
        jt = ticket % ntj;
        kt = ticket / ntj;
    }
    return jt + ntj * kt;
}

// The tile of a ticket (the order of take_tile), for workgroups whose
// ticket was taken by one thread and shared through LDS.
__device__ __forceinline__ int tile_of_ticket(const PressureBuffers &pb, bool reverse, int ticket, int &jt,
                                              int &kt) {
    const int ntj = pb.sm.ntj, ntk = pb.sm.ntk;
    if (reverse) {
        jt = ntj - 1 - ticket % ntj;
        kt = ntk - 1 - ticket / ntj;
    } else {
        jt = ticket % ntj;
        kt = ticket / ntj;
    }
    return jt + ntj * kt;
}

// Processing range [ub, ue) of a tile in its march order u (forward: u = t;
// backward: u = S - 1 - t): the steps that hold at least one of its rows.
__device__ __forceinline__ void tile_urange(const PressureBuffers &pb, const GridGeom &g, int tile,
                                            bool reverse, int &ub, int &ue) {
    int imin = __builtin_amdgcn_readfirstlane(pb.tile_irange[2 * tile]);
    int imax = __builtin_amdgcn_readfirstlane(pb.tile_irange[2 * tile + 1]);
    if (imax < 0) {
        ub = ue = 0;
        return;
    }
    // Cell (i, j, k) sits at step t = i + a + b with a + b in [0, 14].
    if (reverse) {
        ub = pb.sm.S - 15 - imax;
        ue = pb.sm.S - imin;
    } else {
        ub = imin;
        ue = imax + 15;
    }
    if (ue > pb.sm.S) ue = pb.sm.S;
    // Whole groups of 32 steps (S is a multiple of 32), so a sweep's register
    // ring turns over an exact number of times; the extra steps hold no rows
    // and compute zeros.
    ub &= ~31;
    ue = (ue + 31) & ~31;
}

// The last wave of a sweep launch resets the shared counters for the next
// launch and finishes the launch's reduction (mode): the backward sweep's
// dot product, or the max norm of the residual a forward sweep updated or
// recomputed.  tile_acc is the calling wave's per-lane share of it.
__device__ void finish_sweep(PressureBuffers &pb, int mode, double tile_acc, int tile,
                             bool reset_progress, unsigned next_epoch = 0) {
    const int lane = threadIdx.x & 63;
    const int ntiles = pb.sm.ntj * pb.sm.ntk;
    const bool dot = mode == kDotInit || mode == kDotIter;
    const bool norm = mode == kAxpyNorm || mode == kTrueResidual;
    if (dot || norm) {
        double part = dot ? wave_sum_fixed(tile_acc) : wave_max(tile_acc);
        if (lane == 0) st_dev(&pb.partials[tile], part);
    }
    drain_memory();
    unsigned prev = 0;
    if (lane == 0) prev = atomicAdd(&pb.sc->counter[0], 1u);
    prev = __shfl(prev, 0, 64);
    if (prev != (unsigned)ntiles - 1) return;
    if (reset_progress)
        for (int q = lane; q < ntiles; q += 64) pb.progress[q] = 0;
    if (lane == 0) {
        *pb.ticket = 0;
        pb.sc->counter[0] = 0;
        if (next_epoch) pb.sc->sweep_epoch = next_epoch;
    }
    if (norm) {
        double m = 0.0;
        for (int q = lane; q < ntiles; q += 64) {
            const double v = ld_dev(&pb.partials[q]);
            m = v > m ? v : m;
        }
        m = wave_max(m);
        if (lane == 0) {
            PcgScalars *sc = pb.sc;
            if (mode == kAxpyNorm) {
                // |r32|_inf after the axpy: a true residual is due when it
                // meets the tolerance, has fallen by delta since the last
                // one, or the iteration cap is reached.
                sc->iters += 1;
                if (m <= sc->tol || m < sc->delta * sc->rmax || sc->iters >= sc->maxiter) sc->refresh = 1;
                else sc->resid = m;
            } else {
                // |b - A x|_inf in FP64 decides convergence.
                sc->resid = m;
                sc->refinements += 1;
                sc->refresh = 0;
                sc->rmax = m;
                if (m <= sc->tol) {
                    sc->converged = 1;
                    sc->done = 1;
                } else if (sc->iters >= sc->maxiter) {
                    sc->done = 1;
                }
            }
        }
    }
    if (dot) {
        // Fixed-order sum of the per-tile partials.
        double acc = 0.0;
        for (int q = lane; q < ntiles; q += 64) acc += ld_dev(&pb.partials[q]);
        double rho = wave_sum_fixed(acc);
        if (lane == 0) {
            PcgScalars *sc = pb.sc;
            if (mode == kDotInit) {
                sc->rho = rho;
                if (rho == 0 || rho != rho) { sc->failed = 1; sc->done = 1; }
            } else {
                double beta = rho / sc->rho;
                sc->beta = beta;
                sc->beta_f = (float)beta;
                sc->rho = rho;
            }
        }
    }
}

// Extent of each 8 x 8 column tile's rows along i, so sweeps can skip the
// leading and trailing steps where a tile has nothing to do.
__global__ __launch_bounds__(64) void k_tile_irange(GridGeom g, const float *__restrict__ liquid,
                                                    PressureBuffers pb) {
    const int tile = blockIdx.x, lane = threadIdx.x;
    const int jt = tile % pb.sm.ntj, kt = tile / pb.sm.ntj;
    const int j = jt * kTile + (lane & 7), k = kt * kTile + (lane >> 3);
    int lo = 1 << 30, hi = -1;
    if (j >= 1 && k >= 1 && j <= g.J - 2 && k <= g.K - 2) {
        const float *row = liquid + (size_t)g.I * ((size_t)j + (size_t)g.J * k);
        for (int i = 1; i <= g.I - 2; i++) {
            if (row[i] < 0) {
                lo = min(lo, i);
                hi = max(hi, i);
            }
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off, 64));
        hi = max(hi, __shfl_xor(hi, off, 64));
    }
    if (lane == 0) {
        pb.tile_irange[2 * tile] = lo;
        pb.tile_irange[2 * tile + 1] = hi;
    }
}

// Active-group bits of a tile (up to 256 groups; longer tiles treat every
// group as active).  Four scalar words, selected with uniform conditionals
// (an array here would be indexed through scratch memory).
struct GroupBits {
    unsigned long long w0, w1, w2, w3;
    // Without a usable mask all four words are all-ones, so the bit test
    // alone decides (no second condition for the optimiser to thread).
    __device__ __forceinline__ bool test(int gi) const {
        const int w = gi >> 6;
        const unsigned long long lo = (w & 1) ? w1 : w0, hi = (w & 1) ? w3 : w2;
        const unsigned long long word = (w & 2) ? hi : lo;
        return (word >> (gi & 63)) & 1ull;
    }
};
__device__ __forceinline__ GroupBits tile_group_bits(const PressureBuffers &pb, int tile) {
    GroupBits gb{~0ull, ~0ull, ~0ull, ~0ull};
    if (pb.group_mask && pb.mask_words <= 4) {
        const unsigned long long *mw = pb.group_mask + (size_t)tile * pb.mask_words;
        auto rd = [&](int w) -> unsigned long long {
            if (w >= pb.mask_words) return ~0ull;
            const unsigned long long v = mw[w];
            // readfirstlane yields an int: widen through unsigned (no sign extension).
            const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)(v & 0xffffffffull));
            const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
            return (unsigned long long)lo | ((unsigned long long)hi << 32);
        };
        gb.w0 = rd(0); gb.w1 = rd(1); gb.w2 = rd(2); gb.w3 = rd(3);
    }
    return gb;
}
// Whether the chunk at march step u0 (u0 % 8 == 0) holds a row of the tile:
// its two slot groups (forward: steps u0 .. u0 + 7; backward: t = S - 1 - u).
template <bool FWD>
__device__ __forceinline__ int chunk_group(int S, int u0) { return FWD ? (u0 >> 2) : ((S - 8 - u0) >> 2); }
template <bool FWD>
__device__ __forceinline__ bool chunk_has_rows(const GroupBits &gb, int S, int u0) {
    const int gl = chunk_group<FWD>(S, u0);
    return gb.test(gl) || gb.test(gl + 1);
}

// ---------------------------------------------------------------------------
// MIC(0) factorisation (factorModifiedIncompleteColesky0, pcgsolver.h:69-185)
// ---------------------------------------------------------------------------
// Computes invdiag and the column entries Lx, Ly, Lz (the factor arrays must
// be zero on entry).  Runs once per solve; tiles exchange three values per
// edge cell through the slot arrays, ordered by per-chunk progress flags.
constexpr int kFactorChunk = 8;

__global__ __launch_bounds__(64) void k_factor(GridGeom g, PressureBuffers pb) {
    constexpr int C = kFactorChunk;
    const int lane = threadIdx.x;
    const double mf = pb.sc->mf;
    const SlotMap sm = pb.sm;
    const int ntj = sm.ntj;
    int jt, kt;
    const int tile = take_tile(pb, false, jt, kt);
    const int a = lane & 7, b = lane >> 3;
    const int j = jt * kTile + a, k = kt * kTile + b;
    const bool col_ok = j >= 1 && k >= 1 && j <= g.J - 2 && k <= g.K - 2;
    const int up_j = jt > 0 ? tile - 1 : -1;
    const int up_k = kt > 0 ? tile - ntj : -1;
    const int src_j = (lane - 1) & 63, src_k = (lane - 8) & 63;
    const bool edge_j = a == 0, edge_k = b == 0;
    const bool pub = a == 7 || b == 7;
    const unsigned total = (unsigned)sm.S;
    const size_t grp0 = (size_t)tile * (sm.S >> 2);
    int tb, te;
    tile_urange(pb, g, tile, false, tb, te);
    const GroupBits gb = tile_group_bits(pb, tile);

    bool ok = true;
    if (te > tb) {
        if (lane == 0) st_dev(&pb.progress[tile], (unsigned)tb);
        auto slot_at = [&](int t) -> size_t {
            return (((grp0 + (size_t)(t >> 2)) * 64 + (size_t)lane) << 2) + (size_t)(t & 3);
        };
        auto live_at = [&](int t) -> bool {
            int i = t - a - b;
            return col_ok && i >= 1 && i <= g.I - 2;
        };
        double curLx = 0.0, curLy = 0.0, curLz = 0.0;
        for (int t0 = tb; t0 < te; t0 += C) {
            const int t1 = t0 + C < te ? t0 + C : te;
            if (C == 8 && !chunk_has_rows<true>(gb, sm.S, t0)) {
                // No rows in the chunk: its entries are the zeros the arrays
                // hold on entry, whatever its upstream neighbours, so it
                // neither waits nor computes.  (The previous chunk's stores
                // were drained before its progress was published.)
                curLx = curLy = curLz = 0.0;
                if (lane == 0) st_dev(&pb.progress[tile], (unsigned)t1);
                continue;
            }
            unsigned need = (unsigned)t1 + 7;
            if (need > total) need = total;
            // The chunk's operands, loaded before the wait; the couplings
            // stay face weights until used, so nothing here waits on them.
            double v0[C];
            float vw[3][C];
#pragma unroll
            for (int u = 0; u < C; u++) {
                bool live = t0 + u < t1 && live_at(t0 + u);
                size_t p = slot_at(t0 + u);
                v0[u] = live ? pb.diag[p] : 0.0;
                vw[0][u] = live ? pb.wx[p] : -0.0f;
                vw[1][u] = live ? pb.wy[p] : -0.0f;
                vw[2][u] = live ? pb.wz[p] : -0.0f;
            }
            if (up_j >= 0 || up_k >= 0) ok = wait_progress2(pb, up_j, up_k, need);
            if (!ok) break;
            double ey[3][C], ez[3][C];
#pragma unroll
            for (int u = 0; u < C; u++) {
                int t = t0 + u;
                bool live = t < t1 && live_at(t);
                int i = t - a - b;
#pragma unroll
                for (int q = 0; q < 3; q++) { ey[q][u] = 0.0; ez[q][u] = 0.0; }
                if (live && edge_j) {
                    size_t pn = sm.slot32(i, j - 1, k);
                    ey[0][u] = ld_dev(&pb.Lx[pn]); ey[1][u] = ld_dev(&pb.Ly[pn]);
                    ey[2][u] = ld_dev(&pb.Lz[pn]);
                }
                if (live && edge_k) {
                    size_t pn = sm.slot32(i, j, k - 1);
                    ez[0][u] = ld_dev(&pb.Lx[pn]); ez[1][u] = ld_dev(&pb.Ly[pn]);
                    ez[2][u] = ld_dev(&pb.Lz[pn]);
                }
            }
#pragma unroll
            for (int u = 0; u < C; u++) {
                const int t = t0 + u;
                double nj_Lx = shfl_d(curLx, src_j), nj_Ly = shfl_d(curLy, src_j);
                double nj_Lz = shfl_d(curLz, src_j), nk_Lx = shfl_d(curLx, src_k);
                double nk_Ly = shfl_d(curLy, src_k), nk_Lz = shfl_d(curLz, src_k);
                if (t >= t1) continue;  // uniform across the wave
                if (!live_at(t)) {
                    curLx = curLy = curLz = 0.0;
                    continue;
                }
                double yLx = edge_j ? ey[0][u] : nj_Lx, yLy = edge_j ? ey[1][u] : nj_Ly;
                double yLz = edge_j ? ey[2][u] : nj_Lz;
                double zLx = edge_k ? ez[0][u] : nk_Lx, zLy = edge_k ? ez[1][u] : nk_Ly;
                double zLz = edge_k ? ez[2][u] : nk_Lz;
                const double tau = 0.97, sigma = 0.25;  // pcgsolver.h:72-73
                double A = v0[u];
                double d = A;
                double L, missing;
                // Columns in increasing index order: -z, -y, -x neighbour.
                L = zLz; missing = 0; missing += zLx; missing += zLy;
                d -= L * L; d -= tau * L * missing;
                L = yLy; missing = 0; missing += yLx; missing += yLz;
                d -= L * L; d -= tau * L * missing;
                L = curLx; missing = 0; missing += curLy; missing += curLz;
                d -= L * L; d -= tau * L * missing;
                double inv;
                double fx = coupling(vw[0][u], mf), fy = coupling(vw[1][u], mf), fz = coupling(vw[2][u], mf);
                if (A == 0) {
                    inv = d;
                } else {
                    inv = 1 / sqrt(d < sigma * A ? A : d);
                    fx = fx * inv; fy = fy * inv; fz = fz * inv;
                }
                size_t p = slot_at(t);
                if (pub) {
                    st_dev(&pb.invdiag[p], inv); st_dev(&pb.Lx[p], fx);
                    st_dev(&pb.Ly[p], fy); st_dev(&pb.Lz[p], fz);
                } else {
                    pb.invdiag[p] = inv; pb.Lx[p] = fx; pb.Ly[p] = fy; pb.Lz[p] = fz;
                }
                curLx = fx; curLy = fy; curLz = fz;
            }
            // Publish: this chunk's stores have completed before the flag moves.
            drain_memory();
            if (lane == 0) st_dev(&pb.progress[tile], (unsigned)t1);
        }
    }
    drain_memory();
    if (lane == 0) st_dev(&pb.progress[tile], total);
    finish_sweep(pb, kDotNone, 0.0, tile, true);
}

// The factor's entries seen from each row: Lxm(c) = Lx(c - x) etc., so the
// forward solve reads all its operands from the row's own slot.
// The slots a sweep reads from the factor arrays: those of the active groups
// while each tile's group mask is in use (at most 256 groups per tile), every
// slot otherwise (the sweeps then read every group of the march).  One
// thread per slot of a 256-slot group: blockDim.x is 256.
template <class F>
__device__ __forceinline__ void for_swept_slots(const PressureBuffers &pb, F f) {
    if (pb.group_mask && pb.active_list && pb.mask_words <= 4) {
        const ListRange lr = xcd_list(*pb.active_count);
        for (int e = lr.begin; e < lr.end; e += lr.step) f(((size_t)pb.active_list[e] << 8) + threadIdx.x);
    } else {
        const size_t n = pb.sm.count();
        for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x)
            f(p);
    }
}

__global__ __launch_bounds__(kRedThreads) void k_lower_couplings(GridGeom g, PressureBuffers pb) {
    for_swept_slots(pb, [&](size_t p) {
        int i, j, k;
        decode_slot(pb.sm, p, i, j, k);
        double lx = 0.0, ly = 0.0, lz = 0.0;
        if (interior(g, i, j, k)) {
            lx = pb.Lx[pb.sm.slot32(i - 1, j, k)];
            ly = pb.Ly[pb.sm.slot32(i, j - 1, k)];
            lz = pb.Lz[pb.sm.slot32(i, j, k - 1)];
        }
        pb.Lxm[p] = lx;
        pb.Lym[p] = ly;
        pb.Lzm[p] = lz;
    });
}

// ---------------------------------------------------------------------------
// Triangular solves
// ---------------------------------------------------------------------------
// DIR kForward : out = L^-1 in              (solveLower, pcgsolver.h:191-204)
// DIR kBackward: out = L^-T out, in place, optional dot(out, dotvec)
//                                           (solveLowerTransposeInPlace,
//                                            pcgsolver.h:207-221)
// T = double is the reference's arithmetic; T = float is the mixed-precision
// inner solve (FP32 values, FP64 dot accumulation).
//
// One workgroup of two waves per tile.  The compute wave marches over u
// (forward: step t = u; backward: t = S - 1 - u) in chunks of kChunk = 8
// steps (two groups of four steps of the slot layout).  The operands of D
// chunks are in flight at once in a ring of register buffers, so the march
// is not held up by memory latency.  A cell at march step u needs its
// upstream neighbour's value from that tile's march step u + 7, published
// there as record (tile, u, lane-within-edge).  The poller wave fetches the
// records of each chunk (lane L: step u0 + 7 + L / 8, edge lane L % 8) until
// they carry the launch epoch and hands the values over in LDS, laid out so
// that each edge lane of the compute wave reads its eight values with two
// 16-byte LDS loads (FP32).  The compute wave thus never waits on global memory
// for its neighbours: a chunk that is not there yet costs an LDS poll, not
// a drain of its own operand loads and record stores.
constexpr int kChunk = 8;  // 8 steps x 8 edge lanes = one record per poller lane
constexpr int kMailChunks = 4;  // chunks the poller may deliver ahead

// LDS mailbox of a sweep workgroup.
template <class T>
struct SweepMailbox {
    alignas(16) T val[kMailChunks][2][kTile][kChunk];  // [ring][j / k edge][edge lane][step]
    int ready;     // chunks delivered by the poller
    int consumed;  // chunks read by the compute wave
    int tile;      // the workgroup's ticket
    int skip;      // the solve has finished: nothing to do
    unsigned epoch;  // this launch's record tag
    int stored;    // chunks whose outputs the compute wave has stored (the dot wave follows it)
};

__device__ __forceinline__ int lds_acquire(int *p) {
    return __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
}
__device__ __forceinline__ void lds_release(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Four consecutive mailbox values (steps h * 4 .. h * 4 + 3) of one edge lane.
__device__ __forceinline__ void read_mail(const float *p, int h, float (&e)[kChunk]) {
    const float4 v = *reinterpret_cast<const float4 *>(p + 4 * h);
    e[4 * h] = v.x; e[4 * h + 1] = v.y; e[4 * h + 2] = v.z; e[4 * h + 3] = v.w;
}
__device__ __forceinline__ void read_mail(const double *p, int h, double (&e)[kChunk]) {
    const double2 v0 = *reinterpret_cast<const double2 *>(p + 4 * h);
    const double2 v1 = *reinterpret_cast<const double2 *>(p + 4 * h + 2);
    e[4 * h] = v0.x; e[4 * h + 1] = v0.y; e[4 * h + 2] = v1.x; e[4 * h + 3] = v1.y;
}

template <class T>
struct SweepOps {
    const T *in;       // forward: r; backward: unused (in place on out)
    T *out;
    const T *dotvec;   // backward: vector of the fused dot product, or null
    const T *inv;      // invdiag
    const T *c1, *c2, *c3;  // forward: Lzm, Lym, Lxm; backward: Lx, Ly, Lz
};

template <class T>
__device__ __forceinline__ T shfl_t(T v, int src) { return __shfl(v, src, 64); }

// Neighbour values within a column tile without the LDS crossbar.  Lane
// L = a + 8 b of a wave holds column (a, b); a 16-lane DPP row holds b and
// b + 1.  The j neighbour (L - 1 forward, L + 1 backward) is a row shift by
// one.  The k neighbour (L - 8 / L + 8) is a rotation by 8 within the row
// of a register whose other half holds the adjacent row: gfx950's
// permlane16/32 swaps build the rows shifted by one (forward: rows
// [., r0, r1, r2]; backward: [r1, r2, r3, .]), and a per-lane select picks,
// by source half, the lane's own value or the shifted row.  Lanes on the
// tile's upstream edge take their value from the edge records instead, so
// the values these produce there are never used.
// One substitution step, q = (b - c1 x1 - c2 x2 - c3 x3) * inv.
// ref_step keeps the reference's neighbour order (solveLower /
// solveLowerTransposeInPlace, pcgsolver.h:191-221) and rounds each product
// and difference as it does (the FP64 path).  fused_step is for the FP32
// iteration of the mixed path, which need not match bits: fused terms, taken
// in the order the values arrive (own previous step, j neighbour, then the k
// neighbour, whose exchange is longest), for a shorter dependency chain.
__device__ __forceinline__ double ref_step(double b, double c1, double x1, double c2, double x2, double c3,
                                           double x3, double inv) {
    double q = b;
    q -= c1 * x1;
    q -= c2 * x2;
    q -= c3 * x3;
    return q * inv;
}
__device__ __forceinline__ float fused_step(float b, float c_own, float own, float c_j, float xj, float c_k,
                                            float xk, float inv) {
    float q = __builtin_fmaf(-c_own, own, b);
    q = __builtin_fmaf(-c_j, xj, q);
    q = __builtin_fmaf(-c_k, xk, q);
    return q * inv;
}
template <bool FWD>
__device__ __forceinline__ unsigned lane_j(unsigned v) {
    return (unsigned)__builtin_amdgcn_mov_dpp((int)v, FWD ? 0x111 : 0x101, 0xf, 0xf, true);
}
template <bool FWD>
__device__ __forceinline__ unsigned lane_k(unsigned v, bool low_half, bool row_sel) {
    auto a = __builtin_amdgcn_permlane16_swap(v, v, false, false);  // [r0 r0 r2 r2], [r1 r1 r3 r3]
    unsigned shifted, t;
    if (FWD) {
        auto b = __builtin_amdgcn_permlane32_swap(a[1], a[1], false, false);  // [r1 r1 r1 r1]
        shifted = row_sel ? b[0] : a[0];  // row 2 takes r1
        t = low_half ? v : shifted;
    } else {
        auto c = __builtin_amdgcn_permlane32_swap(a[0], a[0], false, false);  // [r2 r2 r2 r2]
        shifted = row_sel ? c[1] : a[1];  // row 1 takes r2
        t = low_half ? shifted : v;
    }
    return (unsigned)__builtin_amdgcn_mov_dpp((int)t, 0x128, 0xf, 0xf, true);  // row_ror:8
}
template <bool FWD>
__device__ __forceinline__ float lane_j(float v) {
    return __builtin_bit_cast(float, lane_j<FWD>(__builtin_bit_cast(unsigned, v)));
}
template <bool FWD>
__device__ __forceinline__ float lane_k(float v, bool low_half, bool row_sel) {
    return __builtin_bit_cast(float, lane_k<FWD>(__builtin_bit_cast(unsigned, v), low_half, row_sel));
}
template <bool FWD>
__device__ __forceinline__ double lane_j(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = lane_j<FWD>((unsigned)u), hi = lane_j<FWD>((unsigned)(u >> 32));
    return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}
template <bool FWD>
__device__ __forceinline__ double lane_k(double v, bool low_half, bool row_sel) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = lane_k<FWD>((unsigned)u, low_half, row_sel);
    const unsigned hi = lane_k<FWD>((unsigned)(u >> 32), low_half, row_sel);
    return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}

// The four steps of one slot group of this lane (16 bytes per access for
// FP32) through a buffer view.
__device__ __forceinline__ void load_group(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, float (&e)[4]) {
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 0);
    e[0] = __uint_as_float(v.x); e[1] = __uint_as_float(v.y);
    e[2] = __uint_as_float(v.z); e[3] = __uint_as_float(v.w);
}
__device__ __forceinline__ void load_group(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, double (&e)[4]) {
    const u32x4 v0 = __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 0);
    const u32x4 v1 = __builtin_amdgcn_raw_buffer_load_b128(r, vo + 16u, so, 0);
    e[0] = __hiloint2double((int)v0.y, (int)v0.x); e[1] = __hiloint2double((int)v0.w, (int)v0.z);
    e[2] = __hiloint2double((int)v1.y, (int)v1.x); e[3] = __hiloint2double((int)v1.w, (int)v1.z);
}
__device__ __forceinline__ void store_group(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, const float (&e)[4]) {
    u32x4 v;
    v.x = __float_as_uint(e[0]); v.y = __float_as_uint(e[1]);
    v.z = __float_as_uint(e[2]); v.w = __float_as_uint(e[3]);
    __builtin_amdgcn_raw_buffer_store_b128(v, r, vo, so, 0);
}
__device__ __forceinline__ void store_group(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, const double (&e)[4]) {
    u32x4 v0, v1;
    v0.x = (unsigned)__double2loint(e[0]); v0.y = (unsigned)__double2hiint(e[0]);
    v0.z = (unsigned)__double2loint(e[1]); v0.w = (unsigned)__double2hiint(e[1]);
    v1.x = (unsigned)__double2loint(e[2]); v1.y = (unsigned)__double2hiint(e[2]);
    v1.z = (unsigned)__double2loint(e[3]); v1.w = (unsigned)__double2hiint(e[3]);
    __builtin_amdgcn_raw_buffer_store_b128(v0, r, vo, so, 0);
    __builtin_amdgcn_raw_buffer_store_b128(v1, r, vo + 16u, so, 0);
}

// The poller wave of a sweep workgroup: for each chunk of the tile's march,
// polls the upstream records until every needed one carries this launch's
// epoch, then delivers the values to the mailbox.  Records outside an
// upstream tile's march range are never written and read as zero.  On a
// timeout it raises the abort flag and releases the compute wave.
template <class T, bool FWD>
__device__ void sweep_poller(PressureBuffers &pb, SweepMailbox<T> &mb, int tile, int up_j, int up_k, int tb,
                             int nchunks, int ubj, int uej, int ubk, int uek, int lane, const GroupBits &gb) {
    constexpr int C = kChunk, RW = Record<T>::W;
    const int S = pb.sm.S;
    const unsigned epoch = mb.epoch;
    const int rs = lane >> 3, re = lane & 7;  // step within the chunk, edge lane
    const int aj = FWD ? 7 : 0, bk = FWD ? 7 : 0;
    const int rslot_j = re, rslot_k = re == aj ? bk : 8 + re;
    // Tiles without an upstream neighbour read their own (never needed) records.
    const int rj_tile = up_j >= 0 ? up_j : tile, rk_tile = up_k >= 0 ? up_k : tile;
    auto rec_ptr = [&](int rt, int slot, int uu) -> const unsigned long long * {
        uu = uu < S ? uu : S - 1;
        return pb.edge + ((((size_t)rt * S + (size_t)uu) * kRecSlots + slot) * RW);
    };
    auto fetch = [&](const unsigned long long *q, unsigned long long (&w)[2]) {
        w[0] = ld_dev(q);
        w[1] = RW == 2 ? ld_dev(q + 1) : 0ull;
    };
    auto give_up = [&]() {
        if (lane == 0) {
            st_dev(&pb.sc->abort, 1);
            lds_release(&mb.ready, 1 << 29);
        }
    };
    for (int ci = 0; ci < nchunks; ci++) {
        const int slot = ci % kMailChunks;
        // Chunk ci overwrites chunk ci - kMailChunks of the ring.
        unsigned spins = 0;
        while (ci - lds_acquire(&mb.consumed) >= kMailChunks) {
            if (++spins > kSpinLimit) return give_up();
            __builtin_amdgcn_s_sleep(1);
        }
        const int uu = tb + ci * C + 7 + rs;
        // A chunk without rows computes zeros whatever its upstream values:
        // it is delivered at once (its compute wave does not read it).
        const bool rows = chunk_has_rows<FWD>(gb, S, tb + ci * C);
        const bool nj = rows & (up_j >= 0) & (uu >= ubj) & (uu < uej);
        const bool nk = rows & (up_k >= 0) & (uu >= ubk) & (uu < uek);
        T vj = T(0), vk = T(0);
        if (__ballot(nj | nk) != 0ull) {
            // The records of the chunk in one round trip per poll.  Every
            // 64th poll also reads the abort flag (one address that every
            // poller would otherwise hammer).
            unsigned long long rj[2], rk[2];
            spins = 0;
            for (;;) {
                fetch(rec_ptr(rj_tile, rslot_j, uu), rj);
                fetch(rec_ptr(rk_tile, rslot_k, uu), rk);
                const int aborted = (spins & 63u) == 63u ? ld_dev(&pb.sc->abort) : 0;
                drain_memory();
                const int bad = ((int)nj & (int)!Record<T>::ok(rj, epoch)) |
                                ((int)nk & (int)!Record<T>::ok(rk, epoch));
                if (__ballot(bad) == 0ull) break;
                if (aborted != 0 || ++spins > kSpinLimit) return give_up();
                __builtin_amdgcn_s_sleep(1);
            }
            vj = nj ? Record<T>::value(rj) : T(0);
            vk = nk ? Record<T>::value(rk) : T(0);
        }
        mb.val[slot][0][re][rs] = vj;
        mb.val[slot][1][re][rs] = vk;
        // The release orders the value stores of every lane before the count.
        if (lane == 0) lds_release(&mb.ready, ci + 1);
    }
}

// y = A x at slot p of cell (i, j, k) (sparsematrix.h:187-216: row entries
// in increasing column order).
__device__ __forceinline__ double spmv_slot(const PressureBuffers &pb, double mf, const double *__restrict__ x,
                                            size_t p, int i, int j, int k) {
    const SlotMap &sm = pb.sm;
    const unsigned pzm = sm.slot32(i, j, k - 1), pym = sm.slot32(i, j - 1, k), pxm = sm.slot32(i - 1, j, k);
    const unsigned pxp = sm.slot32(i + 1, j, k), pyp = sm.slot32(i, j + 1, k), pzp = sm.slot32(i, j, k + 1);
    double s = 0;
    s += coupling(pb.wz[pzm], mf) * x[pzm];
    s += coupling(pb.wy[pym], mf) * x[pym];
    s += coupling(pb.wx[pxm], mf) * x[pxm];
    s += pb.diag[p] * x[p];
    s += coupling(pb.wx[p], mf) * x[pxp];
    s += coupling(pb.wy[p], mf) * x[pyp];
    s += coupling(pb.wz[p], mf) * x[pzp];
    return s;
}

// The true residual of one tile for the forward sweep that follows it
// (kTrueResidual): r = b - A x in FP64 at the slots of the tile's active
// groups, r32 = (float) r, as a separate true-residual pass over the active
// slot list computes them; returns max |r| over the tile to every thread.
// The stores are visible to the workgroup's loads after it returns.
template <int NT>
__device__ double tile_true_residual(const GridGeom &g, const PressureBuffers &pb, int tile, int tb, int te,
                                     const GroupBits &gb) {
    __shared__ double wm[NT / 64];
    const double mf = pb.sc->mf;  // coupling scale of the assembled system
    const size_t tile0 = (size_t)tile * (size_t)(pb.sm.S >> 2) * 256;
    const int g0 = tb >> 2, total = ((te >> 2) - g0) * 256;
    double mx = 0.0;
    constexpr int U = 4;
    for (int e0 = (int)threadIdx.x; e0 < total; e0 += U * NT) {
        size_t p[U];
        bool ok[U], in[U];
        double ax[U], bs[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int e = e0 + u * NT, gi = g0 + (e >> 8);
            ok[u] = e < total && gb.test(gi);
            p[u] = ok[u] ? tile0 + (size_t)gi * 256 + (size_t)(e & 255) : tile0;
            int i, j, k;
            in[u] = stencil_cell(g, pb.sm, p[u], i, j, k);
            ax[u] = spmv_slot(pb, mf, pb.x, p[u], i, j, k);
            bs[u] = pb.rhs[p[u]];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (!ok[u]) continue;
            const double r = in[u] ? bs[u] - ax[u] : 0.0;
            pb.r32[p[u]] = (float)r;
            const double a = fabs(r);
            mx = a > mx ? a : mx;
        }
    }
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = mx;
    __syncthreads();
    double m = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) m = wm[w] > m ? wm[w] : m;
    return m;
}

// The x half of the axpy that an AX forward sweep does (pcgsolver.h:290),
// by the workgroup's third wave: x += alpha s32 (FP64) over the tile's
// active slot groups, four chunks (eight groups) at a time once the compute
// wave has reached them, so this traffic follows the march instead of
// meeting the first tiles' polls all at once.  All active groups lie in the
// march range [tb, te); slots of an active group that are not rows hold
// s32 = 0 and x = 0, which stays 0.
template <class T>
__device__ void sweep_x_wave(const PressureBuffers &pb, SweepMailbox<T> &mb, int tile, int tb, int te, int lane,
                             const GroupBits &gb) {
    constexpr int U = 8;
    const double alpha = pb.sc->alpha;
    const int S = pb.sm.S;
    const size_t tile0 = (size_t)tile * (size_t)(S >> 2) * 256;
    const auto r_x = buffer_view(pb.x + tile0, (unsigned)S * 64u * 8u);
    const auto r_s = buffer_view(pb.s32 + tile0, (unsigned)S * 64u * 4u);
    const unsigned vx = (unsigned)lane * 32u, vs = (unsigned)lane * 16u;
    const int gfirst = tb >> 2, gend = te >> 2;
    int seen = 0;  // chunks the compute wave is known to have reached
    for (int g0 = gfirst; g0 < gend; g0 += U) {
        const int need = (g0 - gfirst) / 2 + 1;  // chunk holding group g0, + 1
        unsigned spins = 0;
        while (seen < need) {
            seen = lds_acquire(&mb.consumed);
            if (seen >= need) break;
            if (++spins > kSpinLimit) break;  // the compute wave gave up: finish anyway
            __builtin_amdgcn_s_sleep(2);
        }
        double xv[U][4];
        float sv[U][4];
        bool act[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int gi = g0 + u;
            act[u] = gi < gend && gb.test(gi);
            load_group(r_x, act[u] ? vx : kOutside, (unsigned)gi * 2048u, xv[u]);
            load_group(r_s, act[u] ? vs : kOutside, (unsigned)gi * 1024u, sv[u]);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
#pragma unroll
            for (int m = 0; m < 4; m++) xv[u][m] = xv[u][m] + alpha * (double)sv[u][m];
            store_group(r_x, act[u] ? vx : kOutside, (unsigned)(g0 + u) * 2048u, xv[u]);
        }
    }
}

// The fused z.r of a backward sweep of the mixed path (the dot of
// pcgsolver.h:296), by the workgroup's third wave: it follows the compute
// wave, four chunks at a time once their z values are stored, reading z and
// r for them, and accumulates fma((double) z, (double) r, acc) per lane in
// march and step order, as the compute wave would.  Chunks without rows
// hold z = 0 and add nothing.  It finishes the launch (finish_sweep).
__device__ void sweep_dot_wave(PressureBuffers &pb, SweepMailbox<float> &mb, const SweepOps<float> &op, int tile,
                               int tb, int nchunks, int lane, const GroupBits &gb, int mode) {
    constexpr int C = kChunk, K = 4;
    const int S = pb.sm.S;
    const size_t tile0 = (size_t)tile * (size_t)(S >> 2) * 256;
    const unsigned fbytes = (unsigned)S * 64u * 4u;
    const auto r_z = buffer_view(op.out + tile0, fbytes);
    const auto r_r = buffer_view(op.dotvec + tile0, fbytes);
    const unsigned vf = (unsigned)lane * 16u;
    double acc = 0.0;
    int seen = 0;
    for (int c0 = 0; c0 < nchunks; c0 += K) {
        const int c1 = min(c0 + K, nchunks);
        unsigned spins = 0;
        while (seen < c1) {
            seen = lds_acquire(&mb.stored);
            if (seen >= c1) break;
            if (++spins > kSpinLimit) break;  // the compute wave gave up: finish anyway
            __builtin_amdgcn_s_sleep(1);
        }
        float zv[K][2][4], rv[K][2][4];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int gl = chunk_group<false>(S, tb + (c0 + k) * C);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const bool act = c0 + k < c1 && gb.test(gl + h);
                const unsigned so = (unsigned)(gl + h) * 1024u;
                load_group(r_z, act ? vf : kOutside, so, zv[k][h]);
                load_group(r_r, act ? vf : kOutside, so, rv[k][h]);
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int s = 0; s < C; s++) {
                // Step s of a backward chunk is element (7 - s) % 4 of group (7 - s) / 4.
                const int h = (7 - s) >> 2, m = (7 - s) & 3;
                acc = __builtin_fma((double)zv[k][h][m], (double)rv[k][h][m], acc);
            }
    }
    finish_sweep(pb, mode, acc, tile, false, mb.epoch + 1);
}

// AX (FP32 forward sweep of the mixed path only): the sweep also does the
// iteration's axpy (pcgsolver.h:289-290).  The compute wave reads
// r32 + (-alpha) q32 in place of r32 (off the step's dependency chain),
// stores it back to r32 and takes its max norm; the third wave updates x
// (sweep_x_wave).  Each value is rounded as a separate axpy pass would
// round it, and the max is order-independent.
// AUX (FP32 sweeps of the mixed path): a third wave helps, the axpy's x
// update on the forward sweep (AX), the z.r dot on the backward one (DW).
template <class T, int DIR, int D, bool AUX = false>
__global__ __launch_bounds__(256) void k_sweep(GridGeom g, PressureBuffers pb, SweepOps<T> op,
                                                          int mode) {
    constexpr int C = kChunk;
    constexpr bool FWD = DIR == kForward;
    static_assert(!AUX || sizeof(T) == 4, "helper waves serve the FP32 sweeps only");
    constexpr bool AX = AUX && FWD, DW = AUX && !FWD;
    constexpr int NV = FWD ? (AX ? 6 : 5) : (DW ? 5 : 6);
    constexpr int RW = Record<T>::W;
    __shared__ SweepMailbox<T> mb;
    if (threadIdx.x == 0) {
        const int skip = ld_dev(&pb.sc->done) != 0 || (mode == kTrueResidual && ld_dev(&pb.sc->refresh) == 0);
        mb.skip = skip;
        mb.tile = skip ? 0 : (int)atomicAdd(pb.ticket, 1u);
        mb.epoch = ld_dev(&pb.sc->sweep_epoch);
        mb.ready = 0;
        mb.consumed = 0;
        mb.stored = 0;
    }
    __syncthreads();
    if (mb.skip) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const SlotMap sm = pb.sm;
    const int S = sm.S, ntj = sm.ntj, ntk = sm.ntk;
    int jt, kt;
    const int tile = tile_of_ticket(pb, !FWD, __builtin_amdgcn_readfirstlane(mb.tile), jt, kt);
    const int a = lane & 7, b = lane >> 3;
    const int up_j = FWD ? (jt > 0 ? tile - 1 : -1) : (jt + 1 < ntj ? tile + 1 : -1);
    const int up_k = FWD ? (kt > 0 ? tile - ntj : -1) : (kt + 1 < ntk ? tile + ntj : -1);
    int tb, te, ubj = 0, uej = 0, ubk = 0, uek = 0;
    tile_urange(pb, g, tile, !FWD, tb, te);
    if (up_j >= 0) tile_urange(pb, g, up_j, !FWD, ubj, uej);
    if (up_k >= 0) tile_urange(pb, g, up_k, !FWD, ubk, uek);
    const int nchunks = te > tb ? (te - tb) / C : 0;
    const GroupBits gb = tile_group_bits(pb, tile);
    // kTrueResidual: the tile's true residual first (every wave), then the sweep over it.
    double tres_max = 0.0;
    if constexpr (FWD && !AX && sizeof(T) == 4) {
        if (mode == kTrueResidual) {
            tres_max = tile_true_residual<256>(g, pb, tile, tb, te, gb);
            if (wave >= 2) return;  // those waves only help with the residual
        }
    }
    if (wave == 1)
        return sweep_poller<T, FWD>(pb, mb, tile, up_j, up_k, tb, nchunks, ubj, uej, ubk, uek, lane, gb);
    if constexpr (AX)
        if (wave == 2) return sweep_x_wave(pb, mb, tile, tb, te, lane, gb);
    if constexpr (DW)
        if (wave == 2) return sweep_dot_wave(pb, mb, op, tile, tb, nchunks, lane, gb, mode);

    const int src_j = (FWD ? lane - 1 : lane + 1) & 63;  // neighbour lane along j
    const int src_k = (FWD ? lane - 8 : lane + 8) & 63;  // neighbour lane along k
    // Lane position in its 16-lane DPP row and the row that takes the
    // permlane32 result (lane_k).
    const bool low_half = (lane & 15) < 8;
    const bool row_sel = (lane >> 4) == (FWD ? 2 : 1);
    const bool edge_j = FWD ? a == 0 : a == 7;            // reads the upstream j tile
    const bool edge_k = FWD ? b == 0 : b == 7;
    // Edge cells (read by the downstream neighbours) publish one record per
    // step: slot b for the j edge, slot 8 + a for the rest of the k edge.
    const int aj = FWD ? 7 : 0, bk = FWD ? 7 : 0;
    const bool pub = a == aj || b == bk;
    const int pub_slot = a == aj ? b : 8 + a;
    const unsigned epoch = mb.epoch;
    const size_t grp0 = (size_t)tile * (size_t)(S >> 2);
    const bool dot = !FWD && !DW && mode != kDotNone;
    // Without a dot product the backward sweep still streams a vector (its
    // own output) so the hot loop has no data-dependent branches.
    const T *dotsrc = dot ? op.dotvec : op.out;


    auto active = [&](int gi) -> bool { return gb.test(gi); };
    // Buffer views of the tile's slots in each array (group gi of the tile
    // at byte gi * 256 * sizeof(T), this lane's four slots at lane * 4 *
    // sizeof(T) within it).  A group without rows is read and written
    // through the offset kOutside: it reads as zeros and its stores (zeros)
    // are dropped.  The group offset is scalar, so addressing a group costs
    // one select for all arrays.
    const unsigned tile_bytes = (unsigned)S * 64u * (unsigned)sizeof(T);
    const size_t tile0 = grp0 * 256;
    const auto r_b = buffer_view((FWD ? op.in : op.out) + tile0, tile_bytes);
    const auto r_c1 = buffer_view(op.c1 + tile0, tile_bytes);
    const auto r_c2 = buffer_view(op.c2 + tile0, tile_bytes);
    const auto r_c3 = buffer_view(op.c3 + tile0, tile_bytes);
    const auto r_inv = buffer_view(op.inv + tile0, tile_bytes);
    const auto r_dot = buffer_view(dotsrc + tile0, tile_bytes);
    const auto r_out = buffer_view(op.out + tile0, tile_bytes);
    // AX: the SpMV result q32 read beside r32, and r32 written back.
    const auto r_q = buffer_view((AX ? (const T *)pb.q32 : op.c1) + tile0, tile_bytes);
    const auto r_rw = buffer_view((AX ? (const T *)pb.r32 : op.out) + tile0, tile_bytes);
    const T naf = AX ? T(-pb.sc->alpha_f) : T(0);
    T rmax = T(0);  // AX: max |r| over the tile's updated residual
    const unsigned vlane = (unsigned)lane * 4u * (unsigned)sizeof(T);
    // Every lane stores a record each step (no branch around the store):
    // lanes that publish nothing write slot 8 + aj, which the corner lane
    // leaves unused and no consumer reads.
    const int rec_slot = pub ? pub_slot : 8 + aj;

    double dot_acc = 0.0;
    if (te > tb) {
        // Lower slot group of the chunk at march step u0 (u0 % 8 == 0).
        auto chunk_group = [&](int u0) -> int { return flipmi::chunk_group<FWD>(S, u0); };
        // Step s of a chunk is element m of group h: forward s = 4h + m,
        // backward s = 7 - (4h + m).
        // act[h] keeps each group's mask bit for the store of the same chunk.
        auto load_chunk = [&](int u0, T (&v)[NV][C], bool (&act)[2]) {
            const int gl = chunk_group(u0);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                act[h] = active(gl + h);
                const unsigned vo = act[h] ? vlane : kOutside;
                const unsigned so = (unsigned)(gl + h) * 256u * (unsigned)sizeof(T);
                T e[NV][4];
                load_group(r_b, vo, so, e[0]);
                load_group(r_c1, vo, so, e[1]);
                load_group(r_c2, vo, so, e[2]);
                load_group(r_c3, vo, so, e[3]);
                load_group(r_inv, vo, so, e[4]);
                if (!FWD && !DW) load_group(r_dot, vo, so, e[NV - 1]);
                if (AX) load_group(r_q, vo, so, e[NV - 1]);
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    const int s = FWD ? 4 * h + m : 7 - (4 * h + m);
#pragma unroll
                    for (int q = 0; q < NV; q++) v[q][s] = e[q][m];
                }
            }
        };
        auto store_chunk = [&](__amdgpu_buffer_rsrc_t dst, int u0, const T (&qv)[C], const bool (&act)[2]) {
            const int gl = chunk_group(u0);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                T e[4];
#pragma unroll
                for (int m = 0; m < 4; m++) e[m] = qv[FWD ? 4 * h + m : 7 - (4 * h + m)];
                // The output of a group without rows stays zero: its values
                // (all zero) are dropped.
                store_group(dst, act[h] ? vlane : kOutside, (unsigned)(gl + h) * 256u * (unsigned)sizeof(T), e);
            }
        };
        // The operands of the first D chunks.
        T buf[D][NV][C];
        bool actr[D][2];
        const int last = te - C;  // last chunk; loads past the end repeat it
#pragma unroll
        for (int d = 0; d < D; d++) {
            int uf = tb + d * C;
            load_chunk(uf < last ? uf : last, buf[d], actr[d]);
            __builtin_amdgcn_sched_barrier(0);
        }
        T cur = T(0);
        int rdy = 0;  // chunks known to be in the mailbox
        // An LDS poll while the poller has not delivered `need` chunks.
        auto wait_mail = [&](int need) {
            if (rdy >= need) return;
            unsigned spins = 0;
            while ((rdy = lds_acquire(&mb.ready)) < need) {
                if (++spins > kSpinLimit) {
                    if (lane == 0) st_dev(&pb.sc->abort, 1);
                    rdy = 1 << 29;
                    return;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        };
        for (int u0 = tb; u0 < te; u0 += D * C) {
#pragma unroll
            for (int d = 0; d < D; d++) {
                // Ranges are whole multiples of D chunks: no exit inside the
                // ring (an exit here costs the compiler its load accounting).
                const int uc = u0 + d * C;
                const int ci = (uc - tb) / C;
                unsigned long long *const recp =
                    pb.edge + ((((size_t)tile * S + (size_t)uc) * kRecSlots + rec_slot) * RW);
                if (!chunk_has_rows<FWD>(gb, S, uc)) {
                    // No rows in the chunk: each of its values is zero whatever
                    // the upstream values (every slot's invdiag is zero), so it
                    // publishes zero records without waiting for its mail
                    // (which its poller does not fetch) or running its steps.
                    // Its outputs are zero already.
                    if (lane == 0) lds_release(&mb.consumed, ci + 1);
#pragma unroll
                    for (int s = 0; s < C; s++) Record<T>::put(recp + (size_t)s * kRecSlots * RW, T(0), epoch);
                    cur = T(0);
                } else {
                    const T *const mj = mb.val[ci % kMailChunks][0][b];
                    const T *const mk = mb.val[ci % kMailChunks][1][a];
                    // The chunk's upstream values: an LDS poll while the poller
                    // has not delivered them.
                    wait_mail(ci + 1);
                    T evj[C], evk[C];
                    read_mail(mj, 0, evj);
                    read_mail(mk, 0, evk);
                    read_mail(mj, 1, evj);
                    read_mail(mk, 1, evk);
                    // The release waits for the reads above, so the poller may
                    // then reuse the ring slot.
                    if (lane == 0) lds_release(&mb.consumed, ci + 1);
                    const T(&v)[NV][C] = buf[d];
                    T qv[C], bv[C];
    #pragma unroll
                    for (int s = 0; s < C; s++) {
                        // Values produced one step earlier by the neighbouring lanes.
                        // FP32: DPP / permlane exchange; FP64: the LDS crossbar
                        // (two dwords per value make the permlane route longer).
                        T nj_cur, nk_cur;
                        if constexpr (sizeof(T) == 4) {
                            nj_cur = lane_j<FWD>(cur);
                            nk_cur = lane_k<FWD>(cur, low_half, row_sel);
                        } else {
                            nj_cur = shfl_t(cur, src_j);
                            nk_cur = shfl_t(cur, src_k);
                        }
                        const T qy = edge_j ? evj[s] : nj_cur;
                        const T qz = edge_k ? evk[s] : nk_cur;
                        // AX: the updated residual r + (-alpha) q, off the
                        // step's dependency chain; fmaxf(m, |r|) is the
                        // compare-and-keep maximum here (m is never NaN).
                        T bs = v[0][s];
                        if constexpr (AX) {
                            bs = v[0][s] + naf * v[NV - 1][s];
                            bv[s] = bs;
                            rmax = fmaxf(rmax, fabsf(bs));
                        }
                        // c1..c3 multiply the -z, -y, -x neighbours (forward) or
                        // the +x, +y, +z ones (backward).
                        T q;
                        if constexpr (sizeof(T) == 8) {
                            q = FWD ? ref_step(v[0][s], v[1][s], qz, v[2][s], qy, v[3][s], cur, v[4][s])
                                    : ref_step(v[0][s], v[1][s], cur, v[2][s], qy, v[3][s], qz, v[4][s]);
                        } else {
                            q = FWD ? fused_step(bs, v[3][s], cur, v[2][s], qy, v[1][s], qz, v[4][s])
                                    : fused_step(bs, v[1][s], cur, v[2][s], qy, v[3][s], qz, v[4][s]);
                        }
                        if (!FWD && !DW) dot_acc = __builtin_fma((double)q, (double)v[NV - 1][s], dot_acc);
                        qv[s] = q;
                        Record<T>::put(recp + (size_t)s * kRecSlots * RW, q, epoch);
                        // Keep each step's record where it is computed: no
                        // memory operation is scheduled across this point (ALU
                        // and LDS instructions may be).
                        __atomic_signal_fence(__ATOMIC_SEQ_CST);
                        __builtin_amdgcn_sched_barrier(0);
                        cur = q;
                    }
                    store_chunk(r_out, uc, qv, actr[d]);
                    if constexpr (AX) store_chunk(r_rw, uc, bv, actr[d]);
                }
                // The dot wave may read the chunk's z: the release orders the
                // stores above before the count (one CU, in issue order).
                if (DW && lane == 0) lds_release(&mb.stored, ci + 1);
                const int uf = uc + D * C;
                load_chunk(uf < last ? uf : last, buf[d], actr[d]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    if constexpr (DW) return;  // the dot wave finishes the launch
    const int fin = FWD ? (AX ? kAxpyNorm : (mode == kTrueResidual ? kTrueResidual : kDotNone))
                        : (dot ? mode : kDotNone);
    finish_sweep(pb, fin, AX ? (double)rmax : (FWD ? tres_max : dot_acc), tile, false, epoch + 1);
}

// Launches a kernel on s.  With a start event the dispatch itself records
// the kernel's start and end (hipExtLaunchKernel), so timing a kernel adds
// no marker packets to the queue between kernels.
template <typename... KArgs, typename... Args>
static hipError_t launch_on(void (*kern)(KArgs...), dim3 grid, dim3 block, hipStream_t s,
                            const KernelTiming &kt, Args... args) {
    if (kt.start) hipExtLaunchKernelGGL(kern, grid, block, 0, s, kt.start, kt.stop, 0, args...);
    else kern<<<grid, block, 0, s>>>(args...);
    return hipGetLastError();
}

// Register ring depth per value type: FP32 operands are half the size.
template <class T>
struct SweepDepth;
template <>
struct SweepDepth<double> {
    static constexpr int D = 2;
};
template <>
struct SweepDepth<float> {
    static constexpr int D = 4;
};

template <class T>
static hipError_t launch_sweep(int dir, const GridGeom &g, PressureBuffers &pb, SweepOps<T> op,
                               int mode, hipStream_t s, const KernelTiming &kt = KernelTiming{}) {
    // A fresh epoch per launch (the device's sweep_epoch, advanced by each
    // launch's last tile); records of older launches never match it.  The
    // host counts launches, an upper bound of the device epoch, and clears
    // the records before the tag could wrap.
    if (++pb.epoch >= 0xfffffff0u) {
        size_t n = edge_record_count(pb.sm);
        hipError_t e = launch_fill_u8((uint8_t *)pb.edge, n * 8, 0, s);
        if (e != hipSuccess) return e;
        e = launch_fill_i32((int *)&pb.sc->sweep_epoch, 1, 1, s);
        if (e != hipSuccess) return e;
        pb.epoch = 1;
    }
    int ntiles = pb.sm.ntj * pb.sm.ntk;
    constexpr int D = SweepDepth<T>::D;
    if (mode == kAxpyNorm) {
        if constexpr (sizeof(T) == 4) {
            if (dir == kForward) return launch_on(k_sweep<T, kForward, D, true>, ntiles, 192, s, kt, g, pb, op, mode);
        }
        return hipErrorInvalidValue;
    }
    // The forward sweep over a fresh true residual has four waves for the
    // residual; two of them then leave.
    if (dir == kForward)
        return launch_on(k_sweep<T, kForward, D>, ntiles, mode == kTrueResidual ? 256 : 128, s, kt, g, pb, op,
                         mode);
    // The FP32 backward sweep's z.r is taken by a third wave.
    if constexpr (sizeof(T) == 4)
        if (mode == kDotInit || mode == kDotIter)
            return launch_on(k_sweep<T, kBackward, D, true>, ntiles, 192, s, kt, g, pb, op, mode);
    return launch_on(k_sweep<T, kBackward, D>, ntiles, 128, s, kt, g, pb, op, mode);
}

static SweepOps<double> fwd_ops64(PressureBuffers &pb, const double *r, double *z) {
    return SweepOps<double>{r, z, nullptr, pb.invdiag, pb.Lzm, pb.Lym, pb.Lxm};
}
static SweepOps<double> bwd_ops64(PressureBuffers &pb, double *z, const double *dotvec) {
    return SweepOps<double>{nullptr, z, dotvec, pb.invdiag, pb.Lx, pb.Ly, pb.Lz};
}
static SweepOps<float> fwd_ops32(PressureBuffers &pb) {
    return SweepOps<float>{pb.r32, pb.z32, nullptr, pb.invdiag32, pb.Lzm32, pb.Lym32, pb.Lxm32};
}
static SweepOps<float> bwd_ops32(PressureBuffers &pb) {
    return SweepOps<float>{nullptr, pb.z32, pb.r32, pb.invdiag32, pb.Lx32, pb.Ly32, pb.Lz32};
}

hipError_t launch_mic0_factor(const GridGeom &g, const float *liquid, PressureBuffers &pb,
                              hipStream_t s) {
    size_t n = pb.sm.count();
    hipError_t e;
    if ((e = launch_fill_f64(pb.invdiag, n, 0.0, s)) != hipSuccess) return e;
    if ((e = launch_fill_f64(pb.Lx, n, 0.0, s)) != hipSuccess) return e;
    if ((e = launch_fill_f64(pb.Ly, n, 0.0, s)) != hipSuccess) return e;
    if ((e = launch_fill_f64(pb.Lz, n, 0.0, s)) != hipSuccess) return e;
    int ntiles = pb.sm.ntj * pb.sm.ntk;
    k_tile_irange<<<ntiles, 64, 0, s>>>(g, liquid, pb);
    k_factor<<<ntiles, 64, 0, s>>>(g, pb);
    k_lower_couplings<<<kRedBlocks, kRedThreads, 0, s>>>(g, pb);
    return hipGetLastError();
}

hipError_t launch_precond_apply(const GridGeom &g, PressureBuffers &pb, const double *r,
                                double *z, hipStream_t s) {
    hipError_t e = launch_sweep(kForward, g, pb, fwd_ops64(pb, r, z), kDotNone, s);
    if (e != hipSuccess) return e;
    return launch_sweep(kBackward, g, pb, bwd_ops64(pb, z, nullptr), kDotNone, s);
}

// ---------------------------------------------------------------------------
// Vector kernels with fused, deterministic reductions
// ---------------------------------------------------------------------------

// Block reduction with a fixed tree; op 0 = sum, 1 = max.
template <int OP>
__device__ double block_reduce(double v, double *sh) {
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = OP == 0 ? v + o : (o > v ? o : v);
    }
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double r = sh[0];
    for (int w = 1; w < kRedThreads / 64; w++) r = OP == 0 ? r + sh[w] : (sh[w] > r ? sh[w] : r);
    __syncthreads();
    return r;
}

// Returns true in the one block that finished last.  Each block stores its
// partial with a coherent store and drains it before taking a ticket, so the
// last block reads every partial with coherent loads.  Tickets are taken in
// two levels so that no cache line takes an atomic from every block (such
// atomics serialise): each block counts in at its XCD's group counter
// (blocks go to XCDs round-robin; one 128-byte line per counter), the last
// block of a group resets that counter and counts the group in at
// `counter`, and the last group's last block is the last block.  The
// caller's last block resets `counter`.
__device__ bool last_block(unsigned *counter, unsigned (*groups)[32]) {
    __shared__ bool is_last;
    drain_memory();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned nb = gridDim.x, grp = blockIdx.x % kXcds;
        const unsigned gsize = nb / kXcds + (grp < nb % kXcds ? 1u : 0u);
        const unsigned ngroups = nb < kXcds ? nb : kXcds;
        bool last = false;
        if (atomicAdd(&groups[grp][0], 1u) == gsize - 1) {
            st_dev(&groups[grp][0], 0u);  // every block of the group has counted in
            last = atomicAdd(counter, 1u) == ngroups - 1;
        }
        is_last = last;
    }
    __syncthreads();
    return is_last;
}

// z = A s and alpha = rho / (s . z), alpha rounded to float as the
// reference's BLAS::addScaled takes it (blaswrapper.h:135-158).
__global__ __launch_bounds__(kRedThreads) void k_spmv_dot(GridGeom g, PressureBuffers pb) {
    __shared__ double sh[kRedThreads / 64];
    if (pb.sc->done) return;
    double acc = 0.0;
    const ListRange lr = xcd_list(*pb.active_count);
    const double mf = pb.sc->mf;  // coupling scale of the assembled system
    constexpr int U = 2;
    for (int e0 = lr.begin; e0 < lr.end; e0 += U * lr.step) {
        size_t p[U];
        bool ok[U], in[U];
        double zc[U], sp[U];
        list_slots<U>(pb, lr, e0, p, ok);
#pragma unroll
        for (int u = 0; u < U; u++) {
            int i, j, k;
            in[u] = stencil_cell(g, pb.sm, p[u], i, j, k);
            zc[u] = spmv_slot(pb, mf, pb.s, p[u], i, j, k);
            sp[u] = pb.s[p[u]];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (!ok[u]) continue;  // uniform
            const double z = in[u] ? zc[u] : 0.0;
            pb.z[p[u]] = z;
            acc += sp[u] * z;
        }
    }
    double bs = block_reduce<0>(acc, sh);
    if (threadIdx.x == 0) st_dev(&pb.partials[blockIdx.x], bs);
    if (!last_block(&pb.sc->counter[1], pb.sc->group_counter[0])) return;
    double tot = 0.0;
    for (int q = threadIdx.x; q < (int)gridDim.x; q += kRedThreads) tot += ld_dev(&pb.partials[q]);
    tot = block_reduce<0>(tot, sh);
    if (threadIdx.x == 0) {
        PcgScalars *sc = pb.sc;
        sc->sz = tot;
        double alpha = sc->rho / tot;
        sc->alpha = alpha;
        sc->alpha_f = (float)alpha;
        sc->counter[1] = 0;
    }
}

// x += alpha s, r -= alpha z, resid = |r|_inf; convergence test
// (pcgsolver.h:280-291).
__global__ __launch_bounds__(kRedThreads) void k_axpy_norm(GridGeom g, PressureBuffers pb) {
    __shared__ double sh[kRedThreads / 64];
    if (pb.sc->done) return;
    float af = pb.sc->alpha_f;
    float naf = (float)(-pb.sc->alpha);
    double mx = 0.0;
    const ListRange lr = xcd_list(*pb.active_count);
    constexpr int U = 4;
    for (int e0 = lr.begin; e0 < lr.end; e0 += U * lr.step) {
        size_t p[U];
        bool ok[U];
        double xs[U], ss[U], rs[U], zs[U];
        list_slots<U>(pb, lr, e0, p, ok);
#pragma unroll
        for (int u = 0; u < U; u++) {
            xs[u] = pb.x[p[u]]; ss[u] = pb.s[p[u]];
            rs[u] = pb.r[p[u]]; zs[u] = pb.z[p[u]];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (!ok[u]) continue;  // uniform
            pb.x[p[u]] = xs[u] + af * ss[u];
            const double r = rs[u] + naf * zs[u];
            pb.r[p[u]] = r;
            const double a = fabs(r);
            mx = a > mx ? a : mx;
        }
    }
    double bm = block_reduce<1>(mx, sh);
    if (threadIdx.x == 0) st_dev(&pb.partials[kRedBlocks + blockIdx.x], bm);
    if (!last_block(&pb.sc->counter[2], pb.sc->group_counter[1])) return;
    double m = 0.0;
    for (int q = threadIdx.x; q < (int)gridDim.x; q += kRedThreads) {
        double v = ld_dev(&pb.partials[kRedBlocks + q]);
        m = v > m ? v : m;
    }
    m = block_reduce<1>(m, sh);
    if (threadIdx.x == 0) {
        PcgScalars *sc = pb.sc;
        sc->resid = m;
        sc->iters += 1;
        if (m <= sc->tol) {
            sc->converged = 1;
            sc->done = 1;
        } else if (sc->iters >= sc->maxiter) {
            sc->done = 1;
        }
        sc->counter[2] = 0;
    }
}

// s = z + beta s (pcgsolver.h:295-296), or s = z on the first application.
__global__ __launch_bounds__(kRedThreads) void k_update_s(GridGeom g, PressureBuffers pb, int first) {
    if (pb.sc->done) return;
    float bf = first ? 0.0f : pb.sc->beta_f;
    const ListRange lr = xcd_list(*pb.active_count);
    constexpr int U = 4;
    for (int e0 = lr.begin; e0 < lr.end; e0 += U * lr.step) {
        size_t p[U];
        bool ok[U];
        double zs[U], ss[U];
        list_slots<U>(pb, lr, e0, p, ok);
#pragma unroll
        for (int u = 0; u < U; u++) {
            zs[u] = pb.z[p[u]];
            ss[u] = pb.s[p[u]];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (!ok[u]) continue;  // uniform
            pb.s[p[u]] = first ? zs[u] : zs[u] + bf * ss[u];
        }
    }
}

// x = 0, r = b; scalars for a new solve.
__global__ void k_pcg_init_vectors(GridGeom g, PressureBuffers pb) {
    size_t n = pb.sm.count();
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x) {
        pb.x[p] = 0.0;
        pb.r[p] = pb.rhs[p];
        pb.z[p] = 0.0;
        pb.s[p] = 0.0;
    }
}

__global__ void k_pcg_init_scalars(PressureBuffers pb, double tolf, int maxiter) {
    PcgScalars *sc = pb.sc;
    double bnorm = __longlong_as_double((long long)sc->absmax_bits);
    sc->bnorm = bnorm;
    sc->resid = bnorm;
    if (tolf < 1e-30) tolf = 1e-30;
    sc->tol = tolf * bnorm;
    sc->iters = 0;
    sc->done = 0;
    sc->converged = 0;
    sc->failed = 0;
    sc->maxiter = maxiter;
    sc->rho = 0;
}

__global__ void k_reset_scalars(PressureBuffers pb) {
    PcgScalars *sc = pb.sc;
    sc->absmax_bits = 0ull;
    sc->rows = 0u;
    sc->done = 0;
    sc->abort = 0;
    sc->failed = 0;
    sc->converged = 0;
    sc->iters = 0;
}

// ---------------------------------------------------------------------------
// Isolated fluid pockets (PressureSolver::_conditionSolidVelocityField,
// pressuresolver.cpp:124-268).  A region of fluid cells joined through faces
// of weight >= 1e-6 that has more than one cell, contains an interior cell
// and borders air through no open face gets zero solid velocities on all six
// faces of each of its cells, which makes its Neumann-only block consistent.
// The reference grows each region by a serial flood fill; here regions are
// the connected components of a lock-free union-find (roots are linked under
// the smaller index with a compare-and-swap; finds halve their paths), then
// one pass collects per-region flags at the roots and one applies them.
// ---------------------------------------------------------------------------

constexpr unsigned kPocketInterior = 1u, kPocketMulti = 2u, kPocketAir = 4u;
constexpr float kOpenFace = 1e-6f;  // "0.0f + eps", pressuresolver.cpp:149

__device__ int uf_find(int *parent, int x) {
    for (;;) {
        const int p = ld_dev(parent + x);
        if (p == x) return x;
        const int gp = ld_dev(parent + p);
        // Parents only ever point to smaller indices, so x (not a root) keeps
        // an ancestor; x is never the target of a link (links need a root).
        if (gp != p) st_dev(parent + x, gp);
        x = gp;
    }
}

__device__ void uf_unite(int *parent, int a, int b) {
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a < b) { const int t = a; a = b; b = t; }
        if (atomicCAS(parent + a, a, b) == a) return;
    }
}

__global__ void k_pocket_init(GridGeom g, const float *__restrict__ liquid, int *parent, unsigned *flags) {
    const size_t n = (size_t)g.I * g.J * g.K;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        parent[c] = liquid[c] < 0.0f ? (int)c : -1;
        flags[c] = 0u;
    }
}

__global__ void k_pocket_union(GridGeom g, const float *__restrict__ liquid, const float *__restrict__ wU,
                               const float *__restrict__ wV, const float *__restrict__ wW, int *parent) {
    const size_t n = (size_t)g.I * g.J * g.K;
    Dims DU{g.I + 1, g.J, g.K}, DV{g.I, g.J + 1, g.K}, DW{g.I, g.J, g.K + 1};
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        if (!(liquid[c] < 0.0f)) continue;
        int i, j, k;
        unflatten3(c, g.I, g.J, i, j, k);
        // Each face once, from the cell on its low side.
        if (i + 1 < g.I && liquid[c + 1] < 0.0f && wU[DU.at(i + 1, j, k)] >= kOpenFace)
            uf_unite(parent, (int)c, (int)(c + 1));
        if (j + 1 < g.J && liquid[c + g.I] < 0.0f && wV[DV.at(i, j + 1, k)] >= kOpenFace)
            uf_unite(parent, (int)c, (int)(c + g.I));
        const size_t sk = (size_t)g.I * g.J;
        if (k + 1 < g.K && liquid[c + sk] < 0.0f && wW[DW.at(i, j, k + 1)] >= kOpenFace)
            uf_unite(parent, (int)c, (int)(c + sk));
    }
}

__global__ void k_pocket_flags(GridGeom g, const float *__restrict__ liquid, const float *__restrict__ wU,
                               const float *__restrict__ wV, const float *__restrict__ wW, int *parent,
                               unsigned *flags) {
    const size_t n = (size_t)g.I * g.J * g.K;
    Dims DU{g.I + 1, g.J, g.K}, DV{g.I, g.J + 1, g.K}, DW{g.I, g.J, g.K + 1};
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        if (ld_dev(parent + c) < 0) continue;
        const int r = uf_find(parent, (int)c);
        int i, j, k;
        unflatten3(c, g.I, g.J, i, j, k);
        unsigned f = (r != (int)c) ? kPocketMulti : 0u;
        if (interior(g, i, j, k)) {
            f |= kPocketInterior;
            // _computeBordersAirGridThread (pressuresolver.cpp:246-268).
            const size_t sk = (size_t)g.I * g.J;
            const bool air = (wU[DU.at(i, j, k)] >= kOpenFace && liquid[c - 1] >= 0.0f) ||
                             (wU[DU.at(i + 1, j, k)] >= kOpenFace && liquid[c + 1] >= 0.0f) ||
                             (wV[DV.at(i, j, k)] >= kOpenFace && liquid[c - g.I] >= 0.0f) ||
                             (wV[DV.at(i, j + 1, k)] >= kOpenFace && liquid[c + g.I] >= 0.0f) ||
                             (wW[DW.at(i, j, k)] >= kOpenFace && liquid[c - sk] >= 0.0f) ||
                             (wW[DW.at(i, j, k + 1)] >= kOpenFace && liquid[c + sk] >= 0.0f);
            if (air) f |= kPocketAir;
        }
        if (f) atomicOr(flags + r, f);
    }
}

__global__ void k_pocket_apply(GridGeom g, int *parent, const unsigned *__restrict__ flags, float *sU, float *sV,
                               float *sW) {
    const size_t n = (size_t)g.I * g.J * g.K;
    Dims DU{g.I + 1, g.J, g.K}, DV{g.I, g.J + 1, g.K}, DW{g.I, g.J, g.K + 1};
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        if (ld_dev(parent + c) < 0) continue;
        const unsigned f = flags[uf_find(parent, (int)c)];
        if (!(f & kPocketInterior) || !(f & kPocketMulti) || (f & kPocketAir)) continue;
        int i, j, k;
        unflatten3(c, g.I, g.J, i, j, k);
        sU[DU.at(i, j, k)] = 0.0f;
        sU[DU.at(i + 1, j, k)] = 0.0f;
        sV[DV.at(i, j, k)] = 0.0f;
        sV[DV.at(i, j + 1, k)] = 0.0f;
        sW[DW.at(i, j, k)] = 0.0f;
        sW[DW.at(i, j, k + 1)] = 0.0f;
    }
}

hipError_t launch_condition_solid(const GridGeom &g, const float *liquid, const float *wU, const float *wV,
                                  const float *wW, float *sU, float *sV, float *sW, int *parent,
                                  unsigned *flags, hipStream_t s) {
    const size_t n = (size_t)g.I * g.J * g.K;
    if (n == 0) return hipSuccess;
    if (n > (size_t)0x7fffffff) return hipErrorInvalidValue;  // cell indices are int
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    k_pocket_init<<<blocks, 256, 0, s>>>(g, liquid, parent, flags);
    k_pocket_union<<<blocks, 256, 0, s>>>(g, liquid, wU, wV, wW, parent);
    k_pocket_flags<<<blocks, 256, 0, s>>>(g, liquid, wU, wV, wW, parent, flags);
    k_pocket_apply<<<blocks, 256, 0, s>>>(g, parent, flags, sU, sV, sW);
    return hipGetLastError();
}

hipError_t launch_pressure_reset(PressureBuffers &pb, hipStream_t s) {
    k_reset_scalars<<<1, 1, 0, s>>>(pb);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 inner iteration (FP32 vectors r, z, s and factor;
// the SpMV reads the FP64 matrix), FP64 dots, FP64 x, and a true
// residual r = b - A x in FP64 that replaces the FP32 recursive residual
// whenever the latter has fallen by `delta` since the last replacement or
// reached the tolerance ("reliable update").  Convergence is declared only
// on the FP64 true residual, |b - A x|_inf <= tol |b|_inf.
// ---------------------------------------------------------------------------

// FP32 copy of the (FP64-computed) factor.
// The FP32 factor of the slots the sweeps read.
__global__ __launch_bounds__(kRedThreads) void k_to_f32(PressureBuffers pb) {
    for_swept_slots(pb, [&](size_t p) {
        pb.invdiag32[p] = (float)pb.invdiag[p];
        pb.Lx32[p] = (float)pb.Lx[p];
        pb.Ly32[p] = (float)pb.Ly[p];
        pb.Lz32[p] = (float)pb.Lz[p];
        pb.Lxm32[p] = (float)pb.Lxm[p];
        pb.Lym32[p] = (float)pb.Lym[p];
        pb.Lzm32[p] = (float)pb.Lzm[p];
    });
}

__global__ void k_pcg_init_vectors32(PressureBuffers pb) {
    size_t n = pb.sm.count();
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x) {
        pb.x[p] = 0.0;
        pb.r32[p] = (float)pb.rhs[p];
        pb.z32[p] = 0.0f;
        pb.s32[p] = 0.0f;
    }
}

// q = A s for the FP32 iteration: the FP64 matrix entries (an FP32 copy's
// rounding costs about 9 % more iterations on the workloads here: 126
// against 115), rebuilt from the FP32 face weights, FP32 s, FP64 sums,
// rounded once.
// The new direction s = z + beta s of the previous iteration
// (pcgsolver.h:300-302), formed where the SpMV reads it.
struct NextDirection {
    const float *__restrict__ z, *__restrict__ s;
    float beta;
    __device__ __forceinline__ float operator[](size_t q) const { return z[q] + beta * s[q]; }
};
struct SameDirection {
    const float *__restrict__ s;
    __device__ __forceinline__ float operator[](size_t q) const { return s[q]; }
};

template <class Dir>
__device__ __forceinline__ double spmv_slot_dir(const PressureBuffers &pb, double mf, const Dir &x, size_t p,
                                                int i, int j, int k) {
    const SlotMap &sm = pb.sm;
    const unsigned pzm = sm.slot32(i, j, k - 1), pym = sm.slot32(i, j - 1, k), pxm = sm.slot32(i - 1, j, k);
    const unsigned pxp = sm.slot32(i + 1, j, k), pyp = sm.slot32(i, j + 1, k), pzp = sm.slot32(i, j, k + 1);
    double s = 0.0;
    s += coupling(pb.wz[pzm], mf) * (double)x[pzm];
    s += coupling(pb.wy[pym], mf) * (double)x[pym];
    s += coupling(pb.wx[pxm], mf) * (double)x[pxm];
    s += pb.diag[p] * (double)x[p];
    s += coupling(pb.wx[p], mf) * (double)x[pxp];
    s += coupling(pb.wy[p], mf) * (double)x[pyp];
    s += coupling(pb.wz[p], mf) * (double)x[pzp];
    return s;
}

// s32b = z32 + beta s32 (or s32 as it is), q32 = A s32b and alpha = rho /
// (s . q) with an FP64 dot.  z32 and s32 are only read, so the stencil
// reads race with no write.
template <class Dir>
__device__ __forceinline__ double spmv_dot32_body(const GridGeom &g, const PressureBuffers &pb,
                                                  const Dir &dir) {
    double acc = 0.0;
    const ListRange lr = xcd_list(*pb.active_count);
    const double mf = pb.sc->mf;  // coupling scale of the assembled system
    constexpr int U = 2;
    for (int e0 = lr.begin; e0 < lr.end; e0 += U * lr.step) {
        size_t p[U];
        bool ok[U], in[U];
        float zc[U], sp[U];
        list_slots<U>(pb, lr, e0, p, ok);
#pragma unroll
        for (int u = 0; u < U; u++) {
            int i, j, k;
            in[u] = stencil_cell(g, pb.sm, p[u], i, j, k);
            zc[u] = (float)spmv_slot_dir(pb, mf, dir, p[u], i, j, k);
            sp[u] = dir[p[u]];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (!ok[u]) continue;  // uniform
            const float q = in[u] ? zc[u] : 0.0f;
            pb.q32[p[u]] = q;
            pb.s32b[p[u]] = sp[u];
            acc += (double)sp[u] * (double)q;
        }
    }
    return acc;
}

__global__ __launch_bounds__(kRedThreads) void k_spmv_dot32(GridGeom g, PressureBuffers pb) {
    __shared__ double sh[kRedThreads / 64];
    if (pb.sc->done) return;
    double acc = pb.s_plain ? spmv_dot32_body(g, pb, SameDirection{pb.s32})
                            : spmv_dot32_body(g, pb, NextDirection{pb.z32, pb.s32, pb.sc->beta_f});
    double bs = block_reduce<0>(acc, sh);
    if (threadIdx.x == 0) st_dev(&pb.partials[blockIdx.x], bs);
    if (!last_block(&pb.sc->counter[1], pb.sc->group_counter[0])) return;
    double tot = 0.0;
    for (int q = threadIdx.x; q < (int)gridDim.x; q += kRedThreads) tot += ld_dev(&pb.partials[q]);
    tot = block_reduce<0>(tot, sh);
    if (threadIdx.x == 0) {
        PcgScalars *sc = pb.sc;
        sc->sz = tot;
        double alpha = sc->rho / tot;
        sc->alpha = alpha;
        sc->alpha_f = (float)alpha;
        sc->counter[1] = 0;
    }
}

// s32 = z32: a solve's first direction (pcgsolver.h:279).  Later
// directions are formed by the SpMV that reads them.
__global__ __launch_bounds__(kRedThreads) void k_first_direction32(PressureBuffers pb) {
    if (pb.sc->done) return;
    const ListRange lr = xcd_list(*pb.active_count);
    constexpr int U = 4;
    for (int e0 = lr.begin; e0 < lr.end; e0 += U * lr.step) {
        size_t p[U];
        bool ok[U];
        float zs[U];
        list_slots<U>(pb, lr, e0, p, ok);
#pragma unroll
        for (int u = 0; u < U; u++) zs[u] = pb.z32[p[u]];
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (!ok[u]) continue;  // uniform
            pb.s32[p[u]] = zs[u];
        }
    }
}

__global__ void k_mixed_init_scalars(PressureBuffers pb) {
    PcgScalars *sc = pb.sc;
    sc->rmax = sc->bnorm;
    sc->delta = 1e-2;
    sc->refresh = 0;
    sc->refinements = 0;
}

hipError_t launch_pcg_init(const GridGeom &g, double tolf, int maxiter, bool mixed,
                           PressureBuffers &pb, hipStream_t s) {
    if (mixed) {
        k_to_f32<<<kRedBlocks, kRedThreads, 0, s>>>(pb);
        k_pcg_init_vectors32<<<kRedBlocks, kRedThreads, 0, s>>>(pb);
        k_pcg_init_scalars<<<1, 1, 0, s>>>(pb, tolf, maxiter);
        k_mixed_init_scalars<<<1, 1, 0, s>>>(pb);
        hipError_t e = launch_sweep(kForward, g, pb, fwd_ops32(pb), kDotNone, s);
        if (e != hipSuccess) return e;
        e = launch_sweep(kBackward, g, pb, bwd_ops32(pb), kDotInit, s);
        if (e != hipSuccess) return e;
        k_first_direction32<<<kRedBlocks, kRedThreads, 0, s>>>(pb);
        pb.s_plain = 1;  // the first SpMV takes s = z as it is
        return hipGetLastError();
    }
    k_pcg_init_vectors<<<kRedBlocks, kRedThreads, 0, s>>>(g, pb);
    k_pcg_init_scalars<<<1, 1, 0, s>>>(pb, tolf, maxiter);
    // z = M^-1 r, rho = z . r, s = z (pcgsolver.h:270-279).
    hipError_t e = launch_sweep(kForward, g, pb, fwd_ops64(pb, pb.r, pb.z), kDotNone, s);
    if (e != hipSuccess) return e;
    e = launch_sweep(kBackward, g, pb, bwd_ops64(pb, pb.z, pb.r), kDotInit, s);
    if (e != hipSuccess) return e;
    k_update_s<<<kRedBlocks, kRedThreads, 0, s>>>(g, pb, 1);
    return hipGetLastError();
}

// Mixed precision folds the axpy into the forward sweep and the direction
// update into the next SpMV; the FP64 path has no true-residual step.
bool pcg_piece_used(int piece, bool mixed) {
    if (mixed) return piece != kPcgUpdate && piece != kPcgAxpy;
    return piece != kPcgResidual;
}

hipError_t launch_pcg_piece(int piece, const GridGeom &g, bool mixed, PressureBuffers &pb,
                            hipStream_t s, const KernelTiming &kt) {
    const dim3 rg(kRedBlocks), rb(kRedThreads);
    if (mixed) {
        switch (piece) {
            case kPcgSpmv: {
                hipError_t e = launch_on(k_spmv_dot32, rg, rb, s, kt, g, pb);
                // The launch took s32 -> s32b by value; the new direction is s32 from here on.
                std::swap(pb.s32, pb.s32b);
                pb.s_plain = 0;
                return e;
            }
            // The forward sweep does the axpy; when its residual norm calls
            // for a true residual, the sweep is redone over that residual.
            case kPcgForward: return launch_sweep(kForward, g, pb, fwd_ops32(pb), kAxpyNorm, s, kt);
            case kPcgResidual: return launch_sweep(kForward, g, pb, fwd_ops32(pb), kTrueResidual, s, kt);
            case kPcgBackward: return launch_sweep(kBackward, g, pb, bwd_ops32(pb), kDotIter, s, kt);
            // kPcgAxpy and kPcgUpdate are folded into the forward sweep and
            // the SpMV.
            default: return hipErrorInvalidValue;
        }
    }
    switch (piece) {
        case kPcgSpmv: return launch_on(k_spmv_dot, rg, rb, s, kt, g, pb);
        case kPcgAxpy: return launch_on(k_axpy_norm, dim3(kAxpyBlocks), rb, s, kt, g, pb);
        case kPcgForward: return launch_sweep(kForward, g, pb, fwd_ops64(pb, pb.r, pb.z), kDotNone, s, kt);
        case kPcgBackward: return launch_sweep(kBackward, g, pb, bwd_ops64(pb, pb.z, pb.r), kDotIter, s, kt);
        case kPcgUpdate: return launch_on(k_update_s, rg, rb, s, kt, g, pb, 0);
        case kPcgResidual: return hipSuccess;
        default: return hipErrorInvalidValue;
    }
}

// Execution order of the pieces within one iteration.
static const int kPieceOrder[kPcgPieces] = {kPcgSpmv,     kPcgAxpy,     kPcgForward,
                                            kPcgResidual, kPcgBackward, kPcgUpdate};

hipError_t launch_pcg_iteration(const GridGeom &g, bool mixed, PressureBuffers &pb, hipStream_t s) {
    for (int q = 0; q < kPcgPieces; q++) {
        int piece = kPieceOrder[q];
        if (!pcg_piece_used(piece, mixed)) continue;
        hipError_t e = launch_pcg_piece(piece, g, mixed, pb, s, KernelTiming{});
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int pcg_piece_at(int q) { return kPieceOrder[q]; }

__global__ void k_spmv_plain(GridGeom g, PressureBuffers pb, const double *__restrict__ x,
                             double *y) {
    size_t n = pb.sm.count();
    const double mf = pb.sc->mf;  // coupling scale of the assembled system
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        decode_slot(pb.sm, p, i, j, k);
        y[p] = interior(g, i, j, k) ? spmv_slot(pb, mf, x, p, i, j, k) : 0.0;
    }
}

hipError_t launch_spmv(const GridGeom &g, PressureBuffers &pb, const double *x, double *y,
                       hipStream_t s) {
    k_spmv_plain<<<kRedBlocks, kRedThreads, 0, s>>>(g, pb, x, y);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pressure gradient update (pressuresolver.cpp:842-1047, no surface tension)
// ---------------------------------------------------------------------------
template <int COMP>
__global__ void k_apply(GridGeom g, SlotMap sm, double dt, const float *__restrict__ liquid,
                        const float *__restrict__ weight, const double *__restrict__ p,
                        float *vel, uint8_t *valid) {
    const int ni = g.I + (COMP == 0), nj = g.J + (COMP == 1), nk = g.K + (COMP == 2);
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= (size_t)ni * nj * nk) return;
    int i, j, k;
    unflatten3(f, ni, nj, i, j, k);
    int idx = COMP == 0 ? i : (COMP == 1 ? j : k);
    int lim = COMP == 0 ? g.I : (COMP == 1 ? g.J : g.K);
    uint8_t vflag = 0;
    if (idx == 0 || idx == lim - 1) {
        valid[f] = 0;  // valid grid is reset before the update (pressuresolver.cpp:862)
        return;
    }
    if (idx == lim) {
        // The outermost face layer lies inside the domain walls (weight 0),
        // so the reference zeroes it; it has no cell on its + side.
        vel[f] = 0.0f;
        valid[f] = 0;
        return;
    }
    int ai = i - (COMP == 0), aj = j - (COMP == 1), ak = k - (COMP == 2);  // cell on the - side
    auto cell = [&](int a, int b, int c) { return (size_t)a + (size_t)g.I * ((size_t)b + (size_t)g.J * c); };
    bool fa = liquid[cell(ai, aj, ak)] < 0.0;
    bool fb = liquid[cell(i, j, k)] < 0.0;
    bool border = fb || fa;
    if (weight[f] > 0 && border) {
        const float eps = 1e-6f;
        float factor = (float)(dt / g.dx);
        float p1, p2;
        if (fa && fb) {
            p1 = (float)p[sm.slot32(ai, aj, ak)];
            p2 = (float)p[sm.slot32(i, j, k)];
        } else {
            float phi1 = liquid[cell(ai, aj, ak)], phi2 = liquid[cell(i, j, k)];
            const double tension = 0.0;
            if (fa) {
                float tp = phi2 / (phi1 + eps);
                double tpd = (double)tp;
                tp = (float)(tpd < -kMaxTheta ? -kMaxTheta : (tpd > kMaxTheta ? kMaxTheta : tpd));
                float tt = (phi1 - phi2) / (phi1 + eps);
                double ttd = (double)tt;
                tt = (float)(ttd < -kMaxTheta ? -kMaxTheta : (ttd > kMaxTheta ? kMaxTheta : ttd));
                p1 = (float)p[sm.slot32(ai, aj, ak)];
                p2 = (float)(tt * tension + tp * p1);
            } else {
                float tp = phi1 / (phi2 + eps);
                double tpd = (double)tp;
                tp = (float)(tpd < -kMaxTheta ? -kMaxTheta : (tpd > kMaxTheta ? kMaxTheta : tpd));
                float tt = (phi2 - phi1) / (phi2 + eps);
                double ttd = (double)tt;
                tt = (float)(ttd < -kMaxTheta ? -kMaxTheta : (ttd > kMaxTheta ? kMaxTheta : ttd));
                p2 = (float)p[sm.slot32(i, j, k)];
                p1 = (float)(tt * tension + tp * p2);
            }
        }
        vel[f] += (float)(-factor * (p2 - p1));
        vflag = 1;
    } else {
        vel[f] = 0.0f;
    }
    valid[f] = vflag;
}

hipError_t launch_pressure_apply(const GridGeom &g, const SlotMap &sm, double dt, const float *liquid,
                                 const float *wU, const float *wV, const float *wW,
                                 const double *p, float *u, float *v, float *w, uint8_t *vu,
                                 uint8_t *vv, uint8_t *vw, hipStream_t s) {
    size_t nU = (size_t)(g.I + 1) * g.J * g.K, nV = (size_t)g.I * (g.J + 1) * g.K;
    size_t nW = (size_t)g.I * g.J * (g.K + 1);
    k_apply<0><<<(unsigned)((nU + 255) / 256), 256, 0, s>>>(g, sm, dt, liquid, wU, p, u, vu);
    k_apply<1><<<(unsigned)((nV + 255) / 256), 256, 0, s>>>(g, sm, dt, liquid, wV, p, v, vv);
    k_apply<2><<<(unsigned)((nW + 255) / 256), 256, 0, s>>>(g, sm, dt, liquid, wW, p, w, vw);
    return hipGetLastError();
}

}  // namespace flipmi
