// Pressure projection: assembly of the variational 7-point Poisson system,
// MIC(0)-preconditioned conjugate gradients and the pressure-gradient update.
//
// References: PressureSolver (pressuresolver.cpp:44-1047) and
// PCGSolver / factorModifiedIncompleteColesky0 / solveLower /
// solveLowerTransposeInPlace (pcgsolver/pcgsolver.h:69-302).
//
// Layout: every vector and matrix array is a dense cell grid (i fastest);
// a cell is a row of the system when it is interior and its liquid phi is
// negative.  Non-row entries hold zeros, so stencils need no row lookup.
//
// The MIC(0) factorisation and both triangular solves run as a pipelined
// wavefront: one 64-lane wave owns an 8 x 8 block of (j, k) columns and
// marches along i with lane (a, b) skewed by a + b steps, so every cell's
// -x, -y and -z dependencies were produced one step earlier by itself or a
// neighbouring lane (exchanged with lane shuffles).  Across blocks a wave
// waits on its upstream neighbours' published progress.  Each cell performs
// exactly the reference's operations in the reference's order, so the
// preconditioner is the reference's bit for bit.
#include "device_math.h"
#include "flipmi_kernels.h"

namespace flipmi {

constexpr int kTile = 8;
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;
constexpr double kMaxTheta = 25.0;  // pressuresolver.h:163
constexpr unsigned kSpinLimit = 1u << 20;  // ~1 s of polling: a legitimate wait is microseconds

// ---------------------------------------------------------------------------
// Assembly
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool is_row_cell(const GridGeom &g, const float *liquid, int i, int j,
                                            int k) {
    if (i < 1 || j < 1 || k < 1 || i > g.I - 2 || j > g.J - 2 || k > g.K - 2) return false;
    return liquid[(size_t)i + (size_t)g.I * ((size_t)j + (size_t)g.J * k)] < 0;
}

// Right-hand side (pressuresolver.cpp:580-669) and matrix coefficients
// (pressuresolver.cpp:723-808) of one row; couplings to the +x/+y/+z rows
// are stored on this row, couplings to -x/-y/-z on the neighbour.
__global__ void k_assemble(GridGeom g, double dt, const float *__restrict__ liquid,
                           const float *__restrict__ wU, const float *__restrict__ wV,
                           const float *__restrict__ wW, const float *__restrict__ wC,
                           const float *__restrict__ u, const float *__restrict__ v,
                           const float *__restrict__ w, const float *__restrict__ sU,
                           const float *__restrict__ sV, const float *__restrict__ sW,
                           PressureBuffers pb) {
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nc = (size_t)g.I * g.J * g.K;
    if (c >= nc) return;
    int i = (int)(c % g.I), j = (int)((c / g.I) % g.J), k = (int)(c / ((size_t)g.I * g.J));
    double rhs = 0, diag = 0, ox = 0, oy = 0, oz = 0;
    if (is_row_cell(g, liquid, i, j, k)) {
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

        double mf = dt / (g.dx * g.dx);
        double eps = 1e-9;
        double pc = liquid[c];
        diag = (volRight + volLeft + volTop + volBottom + volFront + volBack) * mf;
        auto side = [&](double vol, int a, int b, int d, double *off) {
            double pn = liquid[(size_t)a + (size_t)g.I * ((size_t)b + (size_t)g.J * d)];
            if (pn < 0.0) {
                if (off && is_row_cell(g, liquid, a, b, d)) *off = -vol * mf;
            } else {
                double theta = pn / (pc + eps);
                theta = theta < -kMaxTheta ? -kMaxTheta : (theta > kMaxTheta ? kMaxTheta : theta);
                diag -= vol * mf * theta;
            }
        };
        side(volRight, i + 1, j, k, &ox);
        side(volLeft, i - 1, j, k, nullptr);
        side(volTop, i, j + 1, k, &oy);
        side(volBottom, i, j - 1, k, nullptr);
        side(volFront, i, j, k + 1, &oz);
        side(volBack, i, j, k - 1, nullptr);
        diag = diag > 0.0 ? diag : 0.0;
        // max|rhs| as an order-independent integer maximum over IEEE bits.
        unsigned long long bits = (unsigned long long)__double_as_longlong(fabs(rhs));
        atomicMax(&pb.sc->absmax_bits, bits);
        atomicAdd(&pb.sc->rows, 1u);
    }
    pb.rhs[c] = rhs;
    pb.diag[c] = diag;
    pb.offx[c] = ox;
    pb.offy[c] = oy;
    pb.offz[c] = oz;
}

hipError_t launch_pressure_assemble(const GridGeom &g, double dt, const float *liquid,
                                    const float *wU, const float *wV, const float *wW,
                                    const float *wC, const float *u, const float *v,
                                    const float *w, const float *sU, const float *sV,
                                    const float *sW, PressureBuffers &pb, hipStream_t s) {
    size_t nc = (size_t)g.I * g.J * g.K;
    k_assemble<<<(unsigned)((nc + 255) / 256), 256, 0, s>>>(g, dt, liquid, wU, wV, wW, wC, u, v,
                                                            w, sU, sV, sW, pb);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Wavefront sweeps (factorisation, forward and backward substitution)
// ---------------------------------------------------------------------------

enum { kFactor = 0, kForward = 1, kBackward = 2 };
enum { kDotNone = 0, kDotInit = 1, kDotIter = 2 };

// Cross-workgroup exchange without cache maintenance.  Values that another
// workgroup reads within the same kernel are stored and loaded as relaxed
// device-scope atomics, which are coherent across XCDs by themselves; a
// producer orders its data stores before its progress flag by draining its
// outstanding memory operations (s_waitcnt), and a consumer issues its data
// loads only after the flag load has returned.  No acquire/release fence is
// needed, so the sweeps never write back or invalidate the L2 caches.
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
// Waits until every memory operation this wave issued has completed.
__device__ __forceinline__ void drain_memory() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }

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

// Extent of each 8 x 8 column tile's rows along i, so sweeps can skip the
// leading and trailing steps where a tile has nothing to do.
__global__ __launch_bounds__(64) void k_tile_irange(GridGeom g, const float *__restrict__ liquid,
                                                    PressureBuffers pb) {
    const int tile = blockIdx.x, lane = threadIdx.x;
    const int jt = tile % pb.ntiles_j, kt = tile / pb.ntiles_j;
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

// MODE kFactor : computes invdiag, Lx, Ly, Lz from diag/offx/offy/offz
//                (the factor arrays must be zero on entry).
// MODE kForward: out = L^-1 in        (solveLower, pcgsolver.h:191-204)
// MODE kBackward: out = L^-T out (in place), optional dot(out, rvec)
//                 (solveLowerTransposeInPlace, pcgsolver.h:207-221).
// Cells without rows are never written: `out` and the factor hold zeros there.
//
// The march is software-pipelined: the per-cell operands of the next chunk
// of kChunk steps are loaded while the current chunk computes, and the
// values a tile reads from its upstream neighbours are fetched once per
// chunk, after the neighbours have published enough progress.
constexpr int kChunk = 8;  // 5-6 loads per step: keeps a chunk's prefetch under the 63 outstanding loads vmcnt tracks

template <int MODE>
struct SweepVals {
    static constexpr int NV = MODE == kFactor ? 4 : (MODE == kForward ? 5 : 6);
    static constexpr int NE = MODE == kFactor ? 3 : 1;
};

template <int MODE>
__global__ __launch_bounds__(64) void k_wavefront(GridGeom g, PressureBuffers pb,
                                                  const double *__restrict__ in, double *out,
                                                  const double *__restrict__ rvec, int dotmode) {
    constexpr int C = kChunk;
    constexpr int NV = SweepVals<MODE>::NV;
    constexpr int NE = SweepVals<MODE>::NE;
    if (ld_dev(&pb.sc->done) != 0) return;
    const int lane = threadIdx.x;
    const int ntj = pb.ntiles_j, ntk = pb.ntiles_k, ntiles = ntj * ntk;
    int ticket = 0;
    if (lane == 0) ticket = (int)atomicAdd(pb.ticket, 1u);
    ticket = __shfl(ticket, 0, 64);
    int jt, kt;
    if (MODE == kBackward) {
        jt = ntj - 1 - ticket % ntj;
        kt = ntk - 1 - ticket / ntj;
    } else {
        jt = ticket % ntj;
        kt = ticket / ntj;
    }
    const int tile = jt + ntj * kt;
    const int a = lane & 7, b = lane >> 3;
    const int j = jt * kTile + a, k = kt * kTile + b;
    const bool col_ok = j >= 1 && k >= 1 && j <= g.J - 2 && k <= g.K - 2;
    const int up_j = MODE == kBackward ? (jt + 1 < ntj ? tile + 1 : -1) : (jt > 0 ? tile - 1 : -1);
    const int up_k = MODE == kBackward ? (kt + 1 < ntk ? tile + ntj : -1) : (kt > 0 ? tile - ntj : -1);
    const int skew = MODE == kBackward ? (7 - a) + (7 - b) : a + b;
    const int src_j = (MODE == kBackward ? lane + 1 : lane - 1) & 63;  // neighbour lane along j
    const int src_k = (MODE == kBackward ? lane + 8 : lane - 8) & 63;  // neighbour lane along k
    const bool edge_j = MODE == kBackward ? (a == 7) : (a == 0);       // reads the upstream tile
    const bool edge_k = MODE == kBackward ? (b == 7) : (b == 0);
    // Lanes whose results a downstream tile reads publish them coherently.
    const bool pub = MODE == kBackward ? (a == 0 || b == 0) : (a == 7 || b == 7);
    const long sj = (long)g.I, sk = (long)g.I * g.J;
    const long base = sj * j + sk * k;
    const long dy = MODE == kBackward ? sj : -sj, dz = MODE == kBackward ? sk : -sk;
    const unsigned total = (unsigned)(g.I + 14);
    const int imin = pb.tile_irange[2 * tile], imax = pb.tile_irange[2 * tile + 1];

    double dot_acc = 0.0;
    bool ok = true;
    if (imax >= 0) {
        // Steps before tb and from te on touch no row of this tile.
        const unsigned tb = MODE == kBackward ? (unsigned)(g.I - 1 - imax) : (unsigned)imin;
        unsigned te = (MODE == kBackward ? (unsigned)(g.I - 1 - imin) : (unsigned)imax) + 15;
        if (te > total) te = total;
        if (lane == 0) st_dev(&pb.progress[tile], tb);

        // Cell index of this lane at step t, or -1 if it is not interior.
        auto cell_at = [&](unsigned t) -> long {
            int step = (int)t - skew;
            int i = MODE == kBackward ? g.I - 1 - step : step;
            if (!col_ok || step < 0 || step >= g.I || i < 1 || i > g.I - 2) return -1;
            return base + i;
        };
        auto prefetch = [&](unsigned t0, double (&v)[NV][C]) {
#pragma unroll
            for (int u = 0; u < C; u++) {
                long c = cell_at(t0 + u);
                if (c < 0) {
#pragma unroll
                    for (int q = 0; q < NV; q++) v[q][u] = 0.0;
                    continue;
                }
                if (MODE == kFactor) {
                    v[0][u] = pb.diag[c]; v[1][u] = pb.offx[c];
                    v[2][u] = pb.offy[c]; v[3][u] = pb.offz[c];
                } else if (MODE == kForward) {
                    v[0][u] = in[c]; v[1][u] = pb.Lz[c + dz]; v[2][u] = pb.Ly[c + dy];
                    v[3][u] = pb.Lx[c - 1]; v[4][u] = pb.invdiag[c];
                } else {
                    v[0][u] = out[c]; v[1][u] = pb.Lx[c]; v[2][u] = pb.Ly[c];
                    v[3][u] = pb.Lz[c]; v[4][u] = pb.invdiag[c];
                    v[5][u] = dotmode != kDotNone ? rvec[c] : 0.0;
                }
            }
        };

        double cur = 0.0, curLx = 0.0, curLy = 0.0, curLz = 0.0;
        double nxt[NV][C];
        prefetch(tb, nxt);
        for (unsigned t0 = tb; t0 < te; t0 += C) {
            const unsigned t1 = t0 + C < te ? t0 + C : te;
            unsigned need = t1 + 7;
            if (need > total) need = total;
            double v[NV][C];
#pragma unroll
            for (int q = 0; q < NV; q++)
#pragma unroll
                for (int u = 0; u < C; u++) v[q][u] = nxt[q][u];
            if (up_j >= 0 || up_k >= 0) ok = wait_progress2(pb, up_j, up_k, need);
            if (!ok) break;
            // Upstream tiles' values for this chunk (one batch of loads).
            double ey[NE][C], ez[NE][C];
#pragma unroll
            for (int u = 0; u < C; u++) {
                long c = cell_at(t0 + u);
                bool live = c >= 0 && t0 + u < t1;
#pragma unroll
                for (int q = 0; q < NE; q++) { ey[q][u] = 0.0; ez[q][u] = 0.0; }
                if (live && edge_j) {
                    if (MODE == kFactor) {
                        ey[0][u] = ld_dev(&pb.Lx[c + dy]); ey[1][u] = ld_dev(&pb.Ly[c + dy]);
                        ey[2][u] = ld_dev(&pb.Lz[c + dy]);
                    } else {
                        ey[0][u] = ld_dev(&out[c + dy]);
                    }
                }
                if (live && edge_k) {
                    if (MODE == kFactor) {
                        ez[0][u] = ld_dev(&pb.Lx[c + dz]); ez[1][u] = ld_dev(&pb.Ly[c + dz]);
                        ez[2][u] = ld_dev(&pb.Lz[c + dz]);
                    } else {
                        ez[0][u] = ld_dev(&out[c + dz]);
                    }
                }
            }
            // Next chunk's operands stream in while this chunk computes; loads
            // complete in issue order, so the edge values above are not held
            // up by them.
            if (t0 + C < te) prefetch(t0 + C, nxt);
#pragma unroll
            for (int u = 0; u < C; u++) {
                const unsigned t = t0 + u;
                // Values produced one step earlier by the neighbouring lanes.
                double nj_cur = shfl_d(cur, src_j), nk_cur = shfl_d(cur, src_k);
                double nj_Lx = 0, nj_Ly = 0, nj_Lz = 0, nk_Lx = 0, nk_Ly = 0, nk_Lz = 0;
                if (MODE == kFactor) {
                    nj_Lx = shfl_d(curLx, src_j); nj_Ly = shfl_d(curLy, src_j);
                    nj_Lz = shfl_d(curLz, src_j); nk_Lx = shfl_d(curLx, src_k);
                    nk_Ly = shfl_d(curLy, src_k); nk_Lz = shfl_d(curLz, src_k);
                }
                if (t >= t1) continue;  // uniform across the wave
                long c = cell_at(t);
                if (c < 0) {
                    cur = 0.0; curLx = curLy = curLz = 0.0;
                    continue;
                }
                if (MODE == kFactor) {
                    double yLx = edge_j ? ey[0][u] : nj_Lx, yLy = edge_j ? ey[1][u] : nj_Ly;
                    double yLz = edge_j ? ey[2][u] : nj_Lz;
                    double zLx = edge_k ? ez[0][u] : nk_Lx, zLy = edge_k ? ez[1][u] : nk_Ly;
                    double zLz = edge_k ? ez[2][u] : nk_Lz;
                    const double tau = 0.97, sigma = 0.25;  // pcgsolver.h:72-73
                    double A = v[0][u];
                    double d = A;
                    double L, missing;
                    L = zLz; missing = 0; missing += zLx; missing += zLy;   // -z neighbour
                    d -= L * L; d -= tau * L * missing;
                    L = yLy; missing = 0; missing += yLx; missing += yLz;   // -y neighbour
                    d -= L * L; d -= tau * L * missing;
                    L = curLx; missing = 0; missing += curLy; missing += curLz;  // -x neighbour
                    d -= L * L; d -= tau * L * missing;
                    double inv;
                    double fx = v[1][u], fy = v[2][u], fz = v[3][u];
                    if (A == 0) {
                        inv = d;
                    } else {
                        inv = 1 / sqrt(d < sigma * A ? A : d);
                        fx = fx * inv; fy = fy * inv; fz = fz * inv;
                    }
                    if (pub) {
                        st_dev(&pb.invdiag[c], inv); st_dev(&pb.Lx[c], fx);
                        st_dev(&pb.Ly[c], fy); st_dev(&pb.Lz[c], fz);
                    } else {
                        pb.invdiag[c] = inv; pb.Lx[c] = fx; pb.Ly[c] = fy; pb.Lz[c] = fz;
                    }
                    curLx = fx; curLy = fy; curLz = fz;
                } else if (MODE == kForward) {
                    double qy = edge_j ? ey[0][u] : nj_cur;
                    double qz = edge_k ? ez[0][u] : nk_cur;
                    double q = v[0][u];
                    q -= v[1][u] * qz;
                    q -= v[2][u] * qy;
                    q -= v[3][u] * cur;
                    q *= v[4][u];
                    if (pub) st_dev(&out[c], q); else out[c] = q;
                    cur = q;
                } else {
                    double yy = edge_j ? ey[0][u] : nj_cur;
                    double yz = edge_k ? ez[0][u] : nk_cur;
                    double q = v[0][u];
                    q -= v[1][u] * cur;
                    q -= v[2][u] * yy;
                    q -= v[3][u] * yz;
                    q *= v[4][u];
                    if (pub) st_dev(&out[c], q); else out[c] = q;
                    cur = q;
                    dot_acc += q * v[5][u];
                }
            }
            // Publish: this chunk's stores have completed before the flag moves.
            drain_memory();
            if (lane == 0) st_dev(&pb.progress[tile], t1);
        }
    }
    drain_memory();
    if (lane == 0) st_dev(&pb.progress[tile], total);

    if (MODE == kBackward && dotmode != kDotNone) {
        double part = wave_sum_fixed(dot_acc);
        if (lane == 0) st_dev(&pb.partials[tile], part);
    }
    // The last wave to finish reduces the dot product and resets the
    // progress counters and the ticket for the next launch.
    drain_memory();
    unsigned prev = 0;
    if (lane == 0) prev = atomicAdd(&pb.sc->counter[0], 1u);
    prev = __shfl(prev, 0, 64);
    if (prev != (unsigned)ntiles - 1) return;
    for (int q = lane; q < ntiles; q += 64) pb.progress[q] = 0;
    if (lane == 0) {
        *pb.ticket = 0;
        pb.sc->counter[0] = 0;
    }
    if (MODE == kBackward && dotmode != kDotNone) {
        // Fixed-order sum of the per-tile partials.
        double acc = 0.0;
        for (int q = lane; q < ntiles; q += 64) acc += ld_dev(&pb.partials[q]);
        double rho = wave_sum_fixed(acc);
        if (lane == 0) {
            PcgScalars *sc = pb.sc;
            if (dotmode == kDotInit) {
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

static hipError_t launch_wave(int mode, const GridGeom &g, PressureBuffers &pb, const double *in,
                              double *out, const double *rvec, int dotmode, hipStream_t s) {
    int ntiles = pb.ntiles_j * pb.ntiles_k;
    if (mode == kFactor) k_wavefront<kFactor><<<ntiles, 64, 0, s>>>(g, pb, in, out, rvec, dotmode);
    else if (mode == kForward) k_wavefront<kForward><<<ntiles, 64, 0, s>>>(g, pb, in, out, rvec, dotmode);
    else k_wavefront<kBackward><<<ntiles, 64, 0, s>>>(g, pb, in, out, rvec, dotmode);
    return hipGetLastError();
}

hipError_t launch_mic0_factor(const GridGeom &g, const float *liquid, PressureBuffers &pb,
                              hipStream_t s) {
    size_t nc = (size_t)g.I * g.J * g.K;
    hipError_t e;
    if ((e = launch_fill_f64(pb.invdiag, nc, 0.0, s)) != hipSuccess) return e;
    if ((e = launch_fill_f64(pb.Lx, nc, 0.0, s)) != hipSuccess) return e;
    if ((e = launch_fill_f64(pb.Ly, nc, 0.0, s)) != hipSuccess) return e;
    if ((e = launch_fill_f64(pb.Lz, nc, 0.0, s)) != hipSuccess) return e;
    k_tile_irange<<<pb.ntiles_j * pb.ntiles_k, 64, 0, s>>>(g, liquid, pb);
    return launch_wave(kFactor, g, pb, nullptr, nullptr, nullptr, kDotNone, s);
}

hipError_t launch_precond_apply(const GridGeom &g, PressureBuffers &pb, const double *r,
                                double *z, hipStream_t s) {
    hipError_t e = launch_wave(kForward, g, pb, r, z, nullptr, kDotNone, s);
    if (e != hipSuccess) return e;
    return launch_wave(kBackward, g, pb, nullptr, z, nullptr, kDotNone, s);
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
// last block reads every partial with coherent loads.
__device__ bool last_block(unsigned *counter) {
    __shared__ bool is_last;
    drain_memory();
    __syncthreads();
    if (threadIdx.x == 0) is_last = atomicAdd(counter, 1u) == gridDim.x - 1;
    __syncthreads();
    return is_last;
}

__device__ __forceinline__ bool interior_cell(const GridGeom &g, size_t c, int &i, int &j, int &k) {
    i = (int)(c % g.I); j = (int)((c / g.I) % g.J); k = (int)(c / ((size_t)g.I * g.J));
    return i >= 1 && j >= 1 && k >= 1 && i <= g.I - 2 && j <= g.J - 2 && k <= g.K - 2;
}

// y = A x (sparsematrix.h:187-216: row entries in increasing column order).
__device__ __forceinline__ double spmv_row(const GridGeom &g, const PressureBuffers &pb,
                                           const double *__restrict__ x, size_t c) {
    size_t sj = (size_t)g.I, sk = (size_t)g.I * g.J;
    double s = 0;
    s += pb.offz[c - sk] * x[c - sk];
    s += pb.offy[c - sj] * x[c - sj];
    s += pb.offx[c - 1] * x[c - 1];
    s += pb.diag[c] * x[c];
    s += pb.offx[c] * x[c + 1];
    s += pb.offy[c] * x[c + sj];
    s += pb.offz[c] * x[c + sk];
    return s;
}

// z = A s and alpha = rho / (s . z), alpha rounded to float as the
// reference's BLAS::addScaled takes it (blaswrapper.h:135-158).
__global__ __launch_bounds__(kRedThreads) void k_spmv_dot(GridGeom g, PressureBuffers pb) {
    __shared__ double sh[kRedThreads / 64];
    if (pb.sc->done) return;
    size_t nc = (size_t)g.I * g.J * g.K;
    double acc = 0.0;
    for (size_t c = (size_t)blockIdx.x * kRedThreads + threadIdx.x; c < nc;
         c += (size_t)gridDim.x * kRedThreads) {
        int i, j, k;
        double zc = 0.0;
        if (interior_cell(g, c, i, j, k)) zc = spmv_row(g, pb, pb.s, c);
        pb.z[c] = zc;
        acc += pb.s[c] * zc;
    }
    double bs = block_reduce<0>(acc, sh);
    if (threadIdx.x == 0) st_dev(&pb.partials[blockIdx.x], bs);
    if (!last_block(&pb.sc->counter[1])) return;
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
    size_t nc = (size_t)g.I * g.J * g.K;
    float af = pb.sc->alpha_f;
    float naf = (float)(-pb.sc->alpha);
    double mx = 0.0;
    for (size_t c = (size_t)blockIdx.x * kRedThreads + threadIdx.x; c < nc;
         c += (size_t)gridDim.x * kRedThreads) {
        pb.x[c] += af * pb.s[c];
        double r = pb.r[c] + naf * pb.z[c];
        pb.r[c] = r;
        double a = fabs(r);
        mx = a > mx ? a : mx;
    }
    double bm = block_reduce<1>(mx, sh);
    if (threadIdx.x == 0) st_dev(&pb.partials[kRedBlocks + blockIdx.x], bm);
    if (!last_block(&pb.sc->counter[2])) return;
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
__global__ void k_update_s(GridGeom g, PressureBuffers pb, int first) {
    if (pb.sc->done) return;
    size_t nc = (size_t)g.I * g.J * g.K;
    float bf = first ? 0.0f : pb.sc->beta_f;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (size_t)gridDim.x * blockDim.x) {
        if (first) pb.s[c] = pb.z[c];
        else pb.s[c] = pb.z[c] + bf * pb.s[c];
    }
}

// x = 0, r = b; scalars for a new solve.
__global__ void k_pcg_init_vectors(GridGeom g, PressureBuffers pb) {
    size_t nc = (size_t)g.I * g.J * g.K;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (size_t)gridDim.x * blockDim.x) {
        pb.x[c] = 0.0;
        pb.r[c] = pb.rhs[c];
        pb.z[c] = 0.0;
        pb.s[c] = 0.0;
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
    sc->abort = 0;
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

hipError_t launch_pressure_reset(PressureBuffers &pb, hipStream_t s) {
    k_reset_scalars<<<1, 1, 0, s>>>(pb);
    return hipGetLastError();
}

hipError_t launch_pcg_init(const GridGeom &g, double tolf, int maxiter, PressureBuffers &pb,
                           hipStream_t s) {
    k_pcg_init_vectors<<<kRedBlocks, kRedThreads, 0, s>>>(g, pb);
    k_pcg_init_scalars<<<1, 1, 0, s>>>(pb, tolf, maxiter);
    // z = M^-1 r, rho = z . r, s = z (pcgsolver.h:270-279).
    hipError_t e = launch_wave(kForward, g, pb, pb.r, pb.z, nullptr, kDotNone, s);
    if (e != hipSuccess) return e;
    e = launch_wave(kBackward, g, pb, nullptr, pb.z, pb.r, kDotInit, s);
    if (e != hipSuccess) return e;
    k_update_s<<<kRedBlocks, kRedThreads, 0, s>>>(g, pb, 1);
    return hipGetLastError();
}

hipError_t launch_pcg_piece(int piece, const GridGeom &g, PressureBuffers &pb, hipStream_t s) {
    switch (piece) {
        case kPcgSpmv: k_spmv_dot<<<kRedBlocks, kRedThreads, 0, s>>>(g, pb); break;
        case kPcgAxpy: k_axpy_norm<<<kRedBlocks, kRedThreads, 0, s>>>(g, pb); break;
        case kPcgForward: return launch_wave(kForward, g, pb, pb.r, pb.z, nullptr, kDotNone, s);
        case kPcgBackward: return launch_wave(kBackward, g, pb, nullptr, pb.z, pb.r, kDotIter, s);
        case kPcgUpdate: k_update_s<<<kRedBlocks, kRedThreads, 0, s>>>(g, pb, 0); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_pcg_iteration(const GridGeom &g, PressureBuffers &pb, hipStream_t s) {
    for (int piece = 0; piece < kPcgPieces; piece++) {
        hipError_t e = launch_pcg_piece(piece, g, pb, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

__global__ void k_spmv_plain(GridGeom g, PressureBuffers pb, const double *__restrict__ x,
                             double *y) {
    size_t nc = (size_t)g.I * g.J * g.K;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        y[c] = interior_cell(g, c, i, j, k) ? spmv_row(g, pb, x, c) : 0.0;
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
__global__ void k_apply(GridGeom g, double dt, const float *__restrict__ liquid,
                        const float *__restrict__ weight, const double *__restrict__ p,
                        float *vel, uint8_t *valid) {
    const int ni = g.I + (COMP == 0), nj = g.J + (COMP == 1), nk = g.K + (COMP == 2);
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= (size_t)ni * nj * nk) return;
    int i = (int)(f % ni), j = (int)((f / ni) % nj), k = (int)(f / ((size_t)ni * nj));
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
            p1 = (float)p[cell(ai, aj, ak)];
            p2 = (float)p[cell(i, j, k)];
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
                p1 = (float)p[cell(ai, aj, ak)];
                p2 = (float)(tt * tension + tp * p1);
            } else {
                float tp = phi1 / (phi2 + eps);
                double tpd = (double)tp;
                tp = (float)(tpd < -kMaxTheta ? -kMaxTheta : (tpd > kMaxTheta ? kMaxTheta : tpd));
                float tt = (phi2 - phi1) / (phi2 + eps);
                double ttd = (double)tt;
                tt = (float)(ttd < -kMaxTheta ? -kMaxTheta : (ttd > kMaxTheta ? kMaxTheta : ttd));
                p2 = (float)p[cell(i, j, k)];
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

hipError_t launch_pressure_apply(const GridGeom &g, double dt, const float *liquid,
                                 const float *wU, const float *wV, const float *wW,
                                 const double *p, float *u, float *v, float *w, uint8_t *vu,
                                 uint8_t *vv, uint8_t *vw, hipStream_t s) {
    size_t nU = (size_t)(g.I + 1) * g.J * g.K, nV = (size_t)g.I * (g.J + 1) * g.K;
    size_t nW = (size_t)g.I * g.J * (g.K + 1);
    k_apply<0><<<(unsigned)((nU + 255) / 256), 256, 0, s>>>(g, dt, liquid, wU, p, u, vu);
    k_apply<1><<<(unsigned)((nV + 255) / 256), 256, 0, s>>>(g, dt, liquid, wV, p, v, vv);
    k_apply<2><<<(unsigned)((nW + 255) / 256), 256, 0, s>>>(g, dt, liquid, wW, p, w, vw);
    return hipGetLastError();
}

}  // namespace flipmi
