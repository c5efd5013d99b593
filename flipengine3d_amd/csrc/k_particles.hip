// Grid-to-particle transfer and particle advection.
//
// One thread per marker particle, in the engine's canonical particle order:
//   * PIC/FLIP velocity update (fluidsimulation.cpp:4082-4092),
//   * Ralston RK3 through the projected MAC grid (fluidsimulation.cpp:4191-4198),
//   * collision resolution against the solid SDF (fluidsimulation.cpp:4221-4296),
//   * the inputs of particle removal: solid test and speed histogram
//     (fluidsimulation.cpp:4298-4353),
// followed by order-preserving compaction.  The MAC sampler is
// MACVelocityField::evaluateVelocityAtPositionLinear (macvelocityfield.cpp:
// 519-645) in double precision; every per-particle operation matches the
// reference bit for bit given the same grids.
#include "device_math.h"
#include "flipmi_kernels.h"
#include "inflow_device.h"

namespace flipmi {

struct MacView {
    const float *u, *v, *w;
};

// MACVelocityField::_interpolateLinearU/V/W (macvelocityfield.cpp:519-613).
template <int COMP>
__device__ __forceinline__ double interp_comp(const float *__restrict__ grid, const GridGeom &g,
                                              double x, double y, double z) {
    if (!(x >= 0 && y >= 0 && z >= 0 && x < g.dx * g.I && y < g.dx * g.J && z < g.dx * g.K))
        return 0.0;
    if (COMP != 0) x -= 0.5 * g.dx;
    if (COMP != 1) y -= 0.5 * g.dx;
    if (COMP != 2) z -= 0.5 * g.dx;
    const int ni = g.I + (COMP == 0), nj = g.J + (COMP == 1), nk = g.K + (COMP == 2);
    int i = (int)floor(x * g.invdx), j = (int)floor(y * g.invdx), k = (int)floor(z * g.invdx);
    double gx = (double)i * g.dx, gy = (double)j * g.dx, gz = (double)k * g.dx;
    double ix = (x - gx) * g.invdx, iy = (y - gy) * g.invdx, iz = (z - gz) * g.invdx;
    Dims D{ni, nj, nk};
    double p[8];
    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                           {1, 0, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 1}};
#pragma unroll
    for (int q = 0; q < 8; q++) {
        int a = i + off[q][0], b = j + off[q][1], c = k + off[q][2];
        p[q] = D.in(a, b, c) ? (double)grid[D.at(a, b, c)] : 0.0;
    }
    return trilerp8(p, ix, iy, iz);
}

__device__ __forceinline__ F3 mac_velocity(const MacView &m, const GridGeom &g, F3 pos) {
    double x = pos.x, y = pos.y, z = pos.z;
    if (!(x >= 0 && y >= 0 && z >= 0 && x < g.dx * g.I && y < g.dx * g.J && z < g.dx * g.K))
        return f3(0.0f, 0.0f, 0.0f);
    double u = interp_comp<0>(m.u, g, x, y, z);
    double v = interp_comp<1>(m.v, g, x, y, z);
    double w = interp_comp<2>(m.w, g, x, y, z);
    return f3((float)u, (float)v, (float)w);
}

// Interpolation::trilinearInterpolate on the nodal solid SDF (interpolation.cpp:73-112).
__device__ float solid_trilinear(const float *__restrict__ phi, const GridGeom &g, F3 p) {
    I3 c = pos_to_index(p, g.invdx);
    F3 gp = idx_to_pos(c.i, c.j, c.k, g.dx);
    double ix = (double)(p.x - gp.x) * g.invdx, iy = (double)(p.y - gp.y) * g.invdx,
           iz = (double)(p.z - gp.z) * g.invdx;
    Dims N{g.I + 1, g.J + 1, g.K + 1};
    double pts[8];
    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                           {1, 0, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 1}};
#pragma unroll
    for (int q = 0; q < 8; q++) {
        int a = c.i + off[q][0], b = c.j + off[q][1], d = c.k + off[q][2];
        pts[q] = N.in(a, b, d) ? (double)phi[N.at(a, b, d)] : 0.0;
    }
    return (float)trilerp8(pts, ix, iy, iz);
}

// Interpolation::trilinearInterpolateGradient (interpolation.cpp:193-258).
__device__ __forceinline__ double bilerp(double v00, double v10, double v01, double v11, double ix,
                                         double iy) {
    double l1 = (1 - ix) * v00 + ix * v10;
    double l2 = (1 - ix) * v01 + ix * v11;
    return (1 - iy) * l1 + iy * l2;
}
__device__ F3 solid_gradient(const float *__restrict__ phi, const GridGeom &g, F3 p) {
    I3 c = pos_to_index(p, g.invdx);
    F3 gp = idx_to_pos(c.i, c.j, c.k, g.dx);
    double ix = (double)(p.x - gp.x) * g.invdx, iy = (double)(p.y - gp.y) * g.invdx,
           iz = (double)(p.z - gp.z) * g.invdx;
    Dims N{g.I + 1, g.J + 1, g.K + 1};
    auto at = [&](int a, int b, int d) -> float { return N.in(a, b, d) ? phi[N.at(a, b, d)] : 0.0f; };
    float v000 = at(c.i, c.j, c.k), v100 = at(c.i + 1, c.j, c.k), v010 = at(c.i, c.j + 1, c.k),
          v001 = at(c.i, c.j, c.k + 1), v101 = at(c.i + 1, c.j, c.k + 1),
          v011 = at(c.i, c.j + 1, c.k + 1), v110 = at(c.i + 1, c.j + 1, c.k),
          v111 = at(c.i + 1, c.j + 1, c.k + 1);
    float ddx00 = v100 - v000, ddx10 = v110 - v010, ddx01 = v101 - v001, ddx11 = v111 - v011;
    float dvdx = (float)bilerp(ddx00, ddx10, ddx01, ddx11, iy, iz);
    float ddy00 = v010 - v000, ddy10 = v110 - v100, ddy01 = v011 - v001, ddy11 = v111 - v101;
    float dvdy = (float)bilerp(ddy00, ddy10, ddy01, ddy11, ix, iz);
    float ddz00 = v001 - v000, ddz10 = v101 - v100, ddz01 = v011 - v010, ddz11 = v111 - v110;
    float dvdz = (float)bilerp(ddz00, ddz10, ddz01, ddz11, ix, iy);
    return f3(dvdx, dvdy, dvdz);
}

struct Box {
    F3 pos;
    double w, h, d;
    __device__ bool inside(F3 p) const {  // aabb.cpp:130-133
        return p.x >= pos.x && p.y >= pos.y && p.z >= pos.z && (double)p.x < (double)pos.x + w &&
               (double)p.y < (double)pos.y + h && (double)p.z < (double)pos.z + d;
    }
    __device__ F3 nearest_inside(F3 p) const {  // aabb.cpp:501-518
        const double eps = 1e-6;
        if (inside(p)) return p;
        F3 mx = pos + f3((float)w, (float)h, (float)d);
        p.x = (float)fmax((double)p.x, (double)pos.x);
        p.y = (float)fmax((double)p.y, (double)pos.y);
        p.z = (float)fmax((double)p.z, (double)pos.z);
        p.x = (float)fmin((double)p.x, (double)mx.x - eps);
        p.y = (float)fmin((double)p.y, (double)mx.y - eps);
        p.z = (float)fmin((double)p.z, (double)mx.z - eps);
        return p;
    }
};

// k1 is the field's velocity at p0, passed in when the caller has it.
__device__ F3 rk3_from(const MacView &m, const GridGeom &g, F3 p0, F3 k1, double dt) {
    F3 k2 = mac_velocity(m, g, p0 + (float)(0.5 * dt) * k1);
    F3 k3 = mac_velocity(m, g, p0 + (float)(0.75 * dt) * k2);
    return p0 + (float)(dt / 9.0f) * ((2.0f * k1 + 3.0f * k2) + 4.0f * k3);
}
__device__ F3 rk3(const MacView &m, const GridGeom &g, F3 p0, double dt) {
    F3 k1 = mac_velocity(m, g, p0);
    F3 k2 = mac_velocity(m, g, p0 + (float)(0.5 * dt) * k1);
    F3 k3 = mac_velocity(m, g, p0 + (float)(0.75 * dt) * k2);
    return p0 + (float)(dt / 9.0f) * ((2.0f * k1 + 3.0f * k2) + 4.0f * k3);
}

// FluidSimulation::_resolveCollision (fluidsimulation.cpp:4221-4296).
__device__ F3 resolve_collision(const ParticleStepArgs &a, const GridGeom &g, const Box &bnd,
                                F3 oldp, F3 newp) {
    I3 gc = pos_to_index(newp, g.invdx);
    if (!(gc.i >= 0 && gc.j >= 0 && gc.k >= 0 && gc.i < g.I && gc.j < g.J && gc.k < g.K))
        newp = bnd.nearest_inside(newp);
    double inv_near = 1.0 / a.near_cell;
    I3 og = pos_to_index(oldp, inv_near), ng = pos_to_index(newp, inv_near);
    Dims C{a.nci, a.ncj, a.nck};
    bool near_o = C.in(og.i, og.j, og.k) && a.near_solid[C.at(og.i, og.j, og.k)];
    bool near_n = C.in(ng.i, ng.j, ng.k) && a.near_solid[C.at(ng.i, ng.j, ng.k)];
    if (!near_o && !near_n) return newp;
    const float eps = 1e-6f;
    float stepDistance = 0.5f * (float)g.dx;
    float travel = len3(newp - oldp);
    if (travel < eps) return newp;
    int numSteps = (int)ceilf(travel / stepDistance);
    F3 stepdir = normalize3(newp - oldp);
    F3 last = oldp, cur = oldp;
    bool found = false;
    float collisionPhi = 0.0f;
    for (int s = 0; s < numSteps; s++) {
        if (s == numSteps - 1) cur = newp;
        else cur = oldp + ((float)(s + 1) * stepDistance) * stepdir;
        float phi = solid_trilinear(a.solid, g, cur);
        if (phi < 0.0f || !bnd.inside(cur)) {
            collisionPhi = phi;
            found = true;
            break;
        }
        last = cur;
    }
    if (!found) return newp;
    F3 resolved;
    float maxResolved = (float)(a.cfl * g.dx);
    F3 grad = solid_gradient(a.solid, g, cur);
    if (len3(grad) > eps) {
        grad = normalize3(grad);
        resolved = cur - (float)((double)collisionPhi - (double)0.1f * g.dx) * grad;
        float rphi = solid_trilinear(a.solid, g, resolved);
        float rdist = len3(resolved - cur);
        if (rphi < 0 || rdist > maxResolved) resolved = last;
    } else {
        resolved = last;
    }
    if (!bnd.inside(resolved)) {
        F3 orig = resolved;
        resolved = bnd.nearest_inside(resolved);
        float rphi = solid_trilinear(a.solid, g, resolved);
        float rdist = len3(resolved - orig);
        if (rphi < 0.0f || rdist > maxResolved) resolved = last;
    }
    return resolved;
}

__device__ __forceinline__ void particle_step_one(const GridGeom &g, const ParticleStepArgs &a, int i,
                                                  int *hist);

__global__ __launch_bounds__(256) void k_particle_step(GridGeom g, ParticleStepArgs a) {
    __shared__ int hist[64];
    for (int q = threadIdx.x; q < a.maxsteps; q += blockDim.x) hist[q] = 0;
    __syncthreads();
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n) particle_step_one(g, a, i, hist);
    __syncthreads();
    for (int q = threadIdx.x; q < a.maxsteps; q += blockDim.x)
        if (hist[q]) atomicAdd(&a.speed_hist[q], hist[q]);
}

__device__ __forceinline__ void particle_step_one(const GridGeom &g, const ParticleStepArgs &a, int i,
                                                  int *hist) {
    MacView mac{a.u, a.v, a.w}, saved{a.su, a.sv, a.sw};
    F3 p = f3(a.px[i], a.py[i], a.pz[i]);
    F3 vel = f3(a.vx[i], a.vy[i], a.vz[i]);
    // PIC/FLIP blend (fluidsimulation.cpp:4082-4092).
    F3 vpic = mac_velocity(mac, g, p);
    F3 vflip = (vel + vpic) - mac_velocity(saved, g, p);
    F3 vnew = a.pic_ratio_f * vpic + a.flip_ratio_f * vflip;
    // Inflow sources hold their particles' velocity (fluidsimulation.cpp:4172-4185).
    for (int s = 0; s < a.inflow.n; s++) {
        const InflowView &src = a.inflow.v[s];
        if (src.constrain && inflow_holds_particle(src, g, p)) vnew = f3(src.vel[0], src.vel[1], src.vel[2]);
    }
    // Advection + collision (fluidsimulation.cpp:4200-4219).
    Box bnd{f3(a.bpos[0], a.bpos[1], a.bpos[2]), a.bext[0], a.bext[1], a.bext[2]};
    // RK3's first stage samples the projected field at p: that is vpic.
    F3 p1 = rk3_from(mac, g, p, vpic, a.dt);
    p1 = resolve_collision(a, g, bnd, p, p1);
    a.px[i] = p1.x; a.py[i] = p1.y; a.pz[i] = p1.z;
    a.vx[i] = vnew.x; a.vy[i] = vnew.y; a.vz[i] = vnew.z;
    // Removal inputs (fluidsimulation.cpp:4298-4336).
    a.solid_removed[i] = solid_trilinear(a.solid, g, p1) < 0.0f ? 1 : 0;
    double speed = (double)len3(vnew);
    int idx = (int)fmin(floor(speed / a.speed_step), (double)(a.maxsteps - 1));
    atomicAdd(&hist[idx], 1);
}

hipError_t launch_particle_step(const GridGeom &g, const ParticleStepArgs &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    k_particle_step<<<(a.n + 255) / 256, 256, 0, s>>>(g, a);
    return hipGetLastError();
}

// Maximum squared speed as float bits (non-negative floats order as
// unsigned integers), FluidSimulation::_getMaximumMarkerParticleSpeed
// (fluidsimulation.cpp:5553-5565).  Each thread keeps four loads in flight;
// the block reduces its waves' maxima in LDS and takes one atomic, so few
// atomics meet on the one address.  (The maximum is order-independent; NaN
// speeds are skipped.)
constexpr int kMaxSpeedBlocks = 512;
__global__ __launch_bounds__(256) void k_max_speed(const float *__restrict__ vx, const float *__restrict__ vy,
                                                   const float *__restrict__ vz, int n, unsigned int *out) {
    __shared__ unsigned wave_best[4];
    unsigned best = 0;
    auto take = [&](size_t i) {
        float d = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
        unsigned b = __float_as_uint(d);
        if (d == d && b > best) best = b;
    };
    const size_t stride = (size_t)gridDim.x * blockDim.x, end = (size_t)n;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < end; i += 4 * stride) {
        float x[4], y[4], z[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            x[u] = vx[i + u * stride];
            y[u] = vy[i + u * stride];
            z[u] = vz[i + u * stride];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            float d = x[u] * x[u] + y[u] * y[u] + z[u] * z[u];
            unsigned b = __float_as_uint(d);
            if (d == d && b > best) best = b;
        }
    }
    for (; i < end; i += stride) take(i);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        unsigned o = __shfl_xor(best, off, 64);
        best = o > best ? o : best;
    }
    if ((threadIdx.x & 63) == 0) wave_best[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned m = wave_best[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = wave_best[w] > m ? wave_best[w] : m;
        atomicMax(out, m);
    }
}

hipError_t launch_max_speed(const float *vx, const float *vy, const float *vz, int n,
                            unsigned int *out_bits, hipStream_t s) {
    hipError_t e = launch_fill_i32((int *)out_bits, 1, 0, s);
    if (e != hipSuccess) return e;
    if (n <= 0) return hipSuccess;
    int blocks = (n + 255) / 256;
    if (blocks > kMaxSpeedBlocks) blocks = kMaxSpeedBlocks;
    k_max_speed<<<blocks, 256, 0, s>>>(vx, vy, vz, n, out_bits);
    return hipGetLastError();
}

// Per-cell rank among particles that passed the solid test, in particle
// order (the reference's running count grid, fluidsimulation.cpp:4330-4343).
// Only cells holding more than the 250-particle cap need ranks; a thread
// walks its cell's ascending id list.  Nothing runs unless the sort raised
// `over` (then the order exists).
__global__ void k_cell_rank(const int *__restrict__ offset, int ncell, const int *__restrict__ order,
                            const uint8_t *__restrict__ solid_removed, const int *__restrict__ over,
                            int *rank) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell || (over && *over == 0)) return;
    int b = offset[c], e = offset[c + 1];
    if (e - b <= 250) return;
    int r = 0;
    for (int q = b; q < e; q++) {
        int id = order[q];
        if (solid_removed[id]) continue;
        rank[id] = r++;
    }
}

hipError_t launch_cell_rank_cells(const int *offset, int ncell, const int *order,
                                  const uint8_t *solid_removed, const int *over, int *rank,
                                  hipStream_t s) {
    k_cell_rank<<<(ncell + 255) / 256, 256, 0, s>>>(offset, ncell, order, solid_removed, over, rank);
    return hipGetLastError();
}

// Keep flag per particle (FluidSimulation::_removeMarkerParticles,
// fluidsimulation.cpp:4324-4353): solid, over the per-cell cap, or faster
// than the speed limit.  rank < 0 means the cell is under the cap.
__global__ void k_keep(const float *__restrict__ vx, const float *__restrict__ vy,
                       const float *__restrict__ vz, int n, const uint8_t *__restrict__ solid_removed,
                       const int *__restrict__ rank, double maxspeedsq, int *keep) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = 1;
    if (solid_removed[i]) k = 0;
    else if (rank[i] >= 250) k = 0;
    else {
        float d = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
        if ((double)d > maxspeedsq) k = 0;
    }
    keep[i] = k;
}

hipError_t launch_keep_flags(const float *vx, const float *vy, const float *vz, int n,
                             const uint8_t *solid_removed, const int *rank, double maxspeedsq,
                             int *keep, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_keep<<<(n + 255) / 256, 256, 0, s>>>(vx, vy, vz, n, solid_removed, rank, maxspeedsq, keep);
    return hipGetLastError();
}

__global__ void k_compact(const int *__restrict__ keep, const int *__restrict__ pos, int n,
                          const float *__restrict__ a0, const float *__restrict__ a1,
                          const float *__restrict__ a2, const float *__restrict__ a3,
                          const float *__restrict__ a4, const float *__restrict__ a5, float *b0,
                          float *b1, float *b2, float *b3, float *b4, float *b5, int *new_count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == n - 1) *new_count = pos[i] + keep[i];
    if (!keep[i]) return;
    int o = pos[i];
    b0[o] = a0[i]; b1[o] = a1[i]; b2[o] = a2[i];
    b3[o] = a3[i]; b4[o] = a4[i]; b5[o] = a5[i];
}

hipError_t launch_compact_keep(const int *keep, int n, int *pos, int *scan_tmp,
                               const float *const src[6], float *const dst[6], int *new_count,
                               hipStream_t s) {
    if (n <= 0) return launch_fill_i32(new_count, 1, 0, s);
    hipError_t e = launch_exclusive_scan(keep, pos, n, scan_tmp, s);
    if (e != hipSuccess) return e;
    k_compact<<<(n + 255) / 256, 256, 0, s>>>(keep, pos, n, src[0], src[1], src[2], src[3], src[4],
                                              src[5], dst[0], dst[1], dst[2], dst[3], dst[4],
                                              dst[5], new_count);
    return hipGetLastError();
}

// ParticleAdvector::_RK1.._RK4 and trilinearInterpolate
// (particleadvector.cpp:338-445, 498-516) on an AoS float3 array.
__global__ void k_advect_rk(GridGeom g, MacView m, const float *__restrict__ pin, int n, double dt,
                            int order, float *pout) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    F3 p0 = f3(pin[3 * q], pin[3 * q + 1], pin[3 * q + 2]), p1;
    if (order == 4) {
        F3 k1 = mac_velocity(m, g, p0);
        F3 k2 = mac_velocity(m, g, p0 + (float)(0.5 * dt) * k1);
        F3 k3 = mac_velocity(m, g, p0 + (float)(0.5 * dt) * k2);
        F3 k4 = mac_velocity(m, g, p0 + (float)dt * k3);
        p1 = p0 + (float)(dt / 6.0f) * (((k1 + 2.0f * k2) + 2.0f * k3) + k4);
    } else if (order == 3) {
        p1 = rk3(m, g, p0, dt);
    } else if (order == 2) {
        F3 k1 = mac_velocity(m, g, p0);
        F3 k2 = mac_velocity(m, g, p0 + (float)(0.5 * dt) * k1);
        p1 = p0 + (float)dt * k2;
    } else {
        F3 k1 = mac_velocity(m, g, p0);
        p1 = p0 + (float)dt * k1;
    }
    pout[3 * q] = p1.x; pout[3 * q + 1] = p1.y; pout[3 * q + 2] = p1.z;
}

__global__ void k_trilinear(GridGeom g, MacView m, const float *__restrict__ pin, int n, float *out) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    F3 r = mac_velocity(m, g, f3(pin[3 * q], pin[3 * q + 1], pin[3 * q + 2]));
    if (isinf(r.x) || isnan(r.x) || isinf(r.y) || isnan(r.y) || isinf(r.z) || isnan(r.z))
        r = f3(0.0f, 0.0f, 0.0f);
    out[3 * q] = r.x; out[3 * q + 1] = r.y; out[3 * q + 2] = r.z;
}

hipError_t launch_advect_rk(const GridGeom &g, const float *u, const float *v, const float *w,
                            const float *pin, int n, double dt, int order, float *pout,
                            hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_advect_rk<<<(n + 255) / 256, 256, 0, s>>>(g, MacView{u, v, w}, pin, n, dt, order, pout);
    return hipGetLastError();
}

hipError_t launch_trilinear_velocity(const GridGeom &g, const float *u, const float *v,
                                     const float *w, const float *pin, int n, float *out,
                                     hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_trilinear<<<(n + 255) / 256, 256, 0, s>>>(g, MacView{u, v, w}, pin, n, out);
    return hipGetLastError();
}

}  // namespace flipmi
