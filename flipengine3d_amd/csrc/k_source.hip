// Mesh inflow sources (MeshFluidSource with a static mesh, inflow).
//
// Set-up, once per source:
//   * the source SDF on its sub-grid: exact-band distances to the triangles
//     (MeshLevelSet::_computeExactBandDistanceFieldMultiThreaded,
//     meshlevelset.cpp:781-876, 1115-1163) negated at the nodes inside the
//     mesh (_computeDistanceFieldSigns, meshlevelset.cpp:1298-1317);
//   * the source cells: the cells around every domain node inside the mesh
//     (MeshObject::getCells, meshobject.cpp:101-140).
//   Both inside tests are MeshUtils::getCellsInsideTriangleMesh
//   (meshutils.cpp:99-214): one thread per ray column.
// Per substep:
//   * the particle mask of the source cells (ParticleMaskGrid,
//     particlemaskgrid.cpp:54-105),
//   * new particles at the empty subcells (FluidSimulation::_addNewFluidCells,
//     fluidsimulation.cpp:4397-4431, 4523-4562), appended in the reference's
//     order (source cell list order, then subcell offset order) by a scan.
// Outflow sources remove particles through keep flags and the stable
// compaction (k_outflow_keep).  The velocity constraints of particles and
// faces live in the particle step and body-force kernels (k_particles.hip,
// k_grid.hip) through inflow_holds_particle() and inflow_holds_face()
// (inflow_device.h).  Obstacles reuse the set-up kernels for their level
// sets and k_union_min / k_negate to join them to the solid SDF.
#include "device_math.h"
#include "flipmi_kernels.h"

namespace flipmi {

__device__ __forceinline__ F3 cross3(F3 a, F3 b) {  // vmath.h:80-84
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Collision::lineIntersectsTriangle (collision.cpp:79-118) for dir = +z;
// *z receives the crossing's z, p.z + (float)t.
__device__ __forceinline__ bool ray_hit_z(F3 p, F3 v0, F3 v1, F3 v2, float *z) {
    const F3 dir = f3(0.0f, 0.0f, 1.0f);
    F3 e1 = v1 - v0, e2 = v2 - v0;
    F3 h = cross3(dir, e2);
    float a = dot3(e1, h);
    if ((double)fabsf(a) < 10e-9) return false;
    float f = 1.0f / a;
    F3 s = p - v0;
    float u = f * dot3(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    F3 q = cross3(s, e1);
    float v = f * dot3(dir, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    double t = f * dot3(e2, q);
    *z = p.z + (float)t * dir.z;
    return true;
}

// Exact band: one thread per node; each node keeps the minimum distance to
// the triangles whose band box covers it, taken in the frame of the 10^3
// node block that owns it.
__global__ void k_exact_band(const BoxTri *__restrict__ tris, int ntri, Dims N, double dx,
                             float upper, float *phi) {
    const int bw = 10;  // meshlevelset.h:370 (_blockwidth)
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= N.count()) return;
    int i = (int)(tid % N.ni), j = (int)((tid / N.ni) % N.nj), k = (int)(tid / ((size_t)N.ni * N.nj));
    int bi = i / bw, bj = j / bw, bk = k / bw;
    F3 boff = idx_to_pos(bi, bj, bk, bw * dx);
    F3 gpos = idx_to_pos(i - bi * bw, j - bj * bw, k - bk * bw, dx);
    float best = upper;
    for (int t = 0; t < ntri; t++) {
        const BoxTri &T = tris[t];
        if (i < T.gmin[0] || j < T.gmin[1] || k < T.gmin[2] || i > T.gmax[0] || j > T.gmax[1] ||
            k > T.gmax[2])
            continue;
        F3 v0 = f3(T.v[0][0], T.v[0][1], T.v[0][2]) - boff;
        F3 v1 = f3(T.v[1][0], T.v[1][1], T.v[1][2]) - boff;
        F3 v2 = f3(T.v[2][0], T.v[2][1], T.v[2][2]) - boff;
        F3 cp = closest_point_on_triangle(gpos, v0, v1, v2);
        float d = len3(cp - gpos);
        if (d < best) best = d;
    }
    phi[tid] = best;
}

constexpr int kRayThreads = 64;

// Inside test, one thread per ray column (i, j) of the local ci x cj x ck
// cell grid.  The column's crossings are collected in LDS; a column with an
// odd count is outside everywhere, otherwise cell k is inside when an odd
// number of crossings lie below its centre.  Inside cell (i, j, k) is node
// (i, j, k) + go of the target node grid N: mode 0 writes 1 into `nodes`,
// mode 1 negates `phi`.
__global__ __launch_bounds__(kRayThreads) void k_ray_inside(const RayTri *__restrict__ tris,
                                                            int ntri, int ci, int cj, int ck,
                                                            double dx, F3 jitter, int goi, int goj,
                                                            int gok, Dims N, int mode,
                                                            uint8_t *nodes, float *phi,
                                                            int *overflow) {
    __shared__ float zs[kMaxRayCrossings][kRayThreads];  // lane-minor: no bank conflicts
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ci * cj) return;
    int i = col % ci, j = col / ci;
    F3 gp = f3(idx_to_center1(i, dx), idx_to_center1(j, dx), idx_to_center1(-1, dx)) + jitter;
    float *z = &zs[0][threadIdx.x];  // z[q * kRayThreads]
    int nz = 0;
    bool lost = false;
    for (int t = 0; t < ntri; t++) {
        const RayTri &T = tris[t];
        if (i < T.c0[0] || i > T.c1[0] || j < T.c0[1] || j > T.c1[1]) continue;
        float zh;
        if (ray_hit_z(gp, f3(T.v[0][0], T.v[0][1], T.v[0][2]), f3(T.v[1][0], T.v[1][1], T.v[1][2]),
                      f3(T.v[2][0], T.v[2][1], T.v[2][2]), &zh)) {
            if (nz < kMaxRayCrossings) z[nz * kRayThreads] = zh;
            else lost = true;
            nz++;
        }
    }
    if (lost) {
        atomicAdd(overflow, 1);
        return;
    }
    if (nz & 1) return;
    for (int k = 0; k < ck; k++) {
        float zc = idx_to_center1(k, dx);
        int below = 0;
        for (int q = 0; q < nz; q++) below += z[q * kRayThreads] < zc ? 1 : 0;
        if (!(below & 1)) continue;
        int a = i + goi, b = j + goj, c = k + gok;
        if (!N.in(a, b, c)) continue;
        size_t id = N.at(a, b, c);
        if (mode == 0) nodes[id] = 1;
        else phi[id] = -phi[id];
    }
}

// Source cells: every cell with an inside node at one of its 8 corners
// (MeshObject::getCells with Grid3d::getVertexGridIndexNeighbours,
// grid3d.h:326-335); the flag doubles as the input of the list scan.
__global__ void k_cells_from_nodes(const uint8_t *__restrict__ nodes, Dims C, uint8_t *cells,
                                   int *flags) {
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= C.count()) return;
    int i = (int)(tid % C.ni), j = (int)((tid / C.ni) % C.nj), k = (int)(tid / ((size_t)C.ni * C.nj));
    Dims N{C.ni + 1, C.nj + 1, C.nk + 1};
    uint8_t in = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) in |= nodes[N.at(i + (q & 1), j + ((q >> 1) & 1), k + (q >> 2))];
    cells[tid] = in;
    flags[tid] = in;
}

__global__ void k_scan_total(const int *__restrict__ flags, const int *__restrict__ pos, size_t n,
                             int *count) {
    *count = pos[n - 1] + flags[n - 1];
}

__global__ void k_cell_list(const int *__restrict__ flags, const int *__restrict__ pos, size_t n,
                            int *list) {
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n) return;
    if (flags[tid]) list[pos[tid]] = (int)tid;
}

hipError_t launch_mesh_exact_band(const BoxTri *tris, int ntri, int ni, int nj, int nk, double dx,
                                  float upper, float *phi, hipStream_t s) {
    Dims N{ni, nj, nk};
    size_t n = N.count();
    k_exact_band<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(tris, ntri, N, dx, upper, phi);
    return hipGetLastError();
}

hipError_t launch_ray_inside(const RayTri *tris, int ntri, int ci, int cj, int ck, double dx,
                             const float jitter[3], const int go[3], int ni, int nj, int nk,
                             int mode, uint8_t *nodes, float *phi, int *overflow, hipStream_t s) {
    if (ci <= 0 || cj <= 0 || ck <= 0) return hipSuccess;
    if ((mode == 0 && !nodes) || (mode == 1 && !phi)) return hipErrorInvalidValue;
    int cols = ci * cj;
    k_ray_inside<<<(cols + kRayThreads - 1) / kRayThreads, kRayThreads, 0, s>>>(
        tris, ntri, ci, cj, ck, dx, F3{jitter[0], jitter[1], jitter[2]}, go[0], go[1], go[2],
        Dims{ni, nj, nk}, mode, nodes, phi, overflow);
    return hipGetLastError();
}

hipError_t launch_source_cell_scan(const uint8_t *nodes, int I, int J, int K, uint8_t *cells,
                                  int *flags, int *pos, int *scan_tmp, int *count, hipStream_t s) {
    Dims C{I, J, K};
    size_t n = C.count();
    k_cells_from_nodes<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(nodes, C, cells, flags);
    hipError_t e = launch_exclusive_scan(flags, pos, (int)n, scan_tmp, s);
    if (e != hipSuccess) return e;
    k_scan_total<<<1, 1, 0, s>>>(flags, pos, n, count);
    return hipGetLastError();
}

hipError_t launch_source_cell_list(const int *flags, const int *pos, size_t n, int *list,
                                   hipStream_t s) {
    k_cell_list<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(flags, pos, n, list);
    return hipGetLastError();
}

// MeshLevelSet::calculateUnion (meshlevelset.cpp:603-638, 1758-1790) of a
// sub-grid level set at node offset go into a nodal field: the smaller value
// wins.
__global__ void k_union_min(const float *__restrict__ phi, Dims S, int goi, int goj, int gok, Dims N,
                            float *dst) {
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= S.count()) return;
    int i = (int)(tid % S.ni), j = (int)((tid / S.ni) % S.nj), k = (int)(tid / ((size_t)S.ni * S.nj));
    int a = i + goi, b = j + goj, c = k + gok;
    if (!N.in(a, b, c)) return;
    size_t id = N.at(a, b, c);
    float v = phi[tid];
    if (v < dst[id]) dst[id] = v;
}

__global__ void k_negate(float *phi, size_t n) {  // MeshLevelSet::negate (meshlevelset.cpp:702-708)
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) phi[tid] = -phi[tid];
}

hipError_t launch_negate(float *phi, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_negate<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(phi, n);
    return hipGetLastError();
}

hipError_t launch_union_min(const float *phi, int ni, int nj, int nk, const int go[3], int Ni, int Nj,
                            int Nk, float *dst, hipStream_t s) {
    Dims S{ni, nj, nk};
    size_t n = S.count();
    if (n == 0) return hipSuccess;
    k_union_min<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(phi, S, go[0], go[1], go[2], Dims{Ni, Nj, Nk}, dst);
    return hipGetLastError();
}

// ---- per substep ----------------------------------------------------------

// ParticleMaskGrid::addParticle: the bit of the particle's subcell parity in
// the byte of its cell (floor(p / dx)).  The mask is one byte per domain
// cell packed four to a word; only the bytes of source cells are ever read,
// so only particles in a source cell's box set bits.
__device__ __forceinline__ uint32_t subcell_bit(I3 s) {
    int c = ((s.i % 2 == 1) ? 1 : 0) | ((s.j % 2 == 1) ? 2 : 0) | ((s.k % 2 == 1) ? 4 : 0);
    return 1u << c;
}
__device__ __forceinline__ void mask_add(uint32_t *mask, const GridGeom &g, F3 p, double invsub) {
    I3 s = pos_to_index(p, invsub);
    I3 c = pos_to_index(p, g.invdx);
    size_t cell = (size_t)c.i + (size_t)g.I * ((size_t)c.j + (size_t)g.J * c.k);
    atomicOr(&mask[cell >> 2], subcell_bit(s) << (8 * (cell & 3)));
}

__global__ void k_mask_clear(const int *__restrict__ list, int count, uint8_t *mask_bytes) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    mask_bytes[list[q]] = 0;
}

__global__ __launch_bounds__(256) void k_mask_particles(GridGeom g, const float *__restrict__ px,
                                                        const float *__restrict__ py,
                                                        const float *__restrict__ pz, int n,
                                                        InflowViews views, uint32_t *mask) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    F3 p = f3(px[q], py[q], pz[q]);
    I3 c = pos_to_index(p, g.invdx);
    bool hit = false;
    for (int s = 0; s < views.n; s++) {
        const InflowView &v = views.v[s];
        if (!v.emit) continue;
        if (c.i < v.cmin[0] || c.j < v.cmin[1] || c.k < v.cmin[2] || c.i > v.cmax[0] ||
            c.j > v.cmax[1] || c.k > v.cmax[2])
            continue;
        if (v.cells[(size_t)c.i + (size_t)g.I * ((size_t)c.j + (size_t)g.J * c.k)]) hit = true;
    }
    if (hit) mask_add(mask, g, p, 1.0 / (0.5 * g.dx));
}

// One emission pass of one source, candidate = 8 * (list index) + subcell.
__global__ __launch_bounds__(256) void k_emit_flags(GridGeom g, InflowView v,
                                                    const float *__restrict__ solid,
                                                    const uint32_t *__restrict__ mask,
                                                    EmitParams ep, float *cand, int *flags) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 8 * v.ncells) return;
    int o = q & 7;
    int cell = v.list[q >> 3];
    int i = cell % g.I, j = (cell / g.I) % g.J, k = cell / (g.I * g.J);
    const float qf = (float)(0.25 * g.dx);
    F3 off = f3((o & 1) ? qf : -qf, (o & 2) ? qf : -qf, (o & 4) ? qf : -qf);
    F3 p = f3(idx_to_center1(i, g.dx), idx_to_center1(j, g.dx), idx_to_center1(k, g.dx)) + off;
    int keep = 0;
    double d = nodal_trilinear(v.sdf, Dims{v.ni, v.nj, v.nk}, g.dx, g.invdx,
                               p - f3(v.off[0], v.off[1], v.off[2]));
    if (!(d > 0)) {
        if (ep.jitter_surface || d < -g.dx) {
            // _jitterMarkerParticlePosition (fluidsimulation.cpp:2628-2635):
            // p += _randomDouble(-j, j) per axis, drawn from the keyed stream.
            uint64_t item = 3ull * (uint64_t)q;
            const double jt = ep.jitter;
            p.x = (float)((double)p.x + (-jt + draw_unit(ep.seed, ep.stream, item) * (jt - -jt)));
            p.y = (float)((double)p.y + (-jt + draw_unit(ep.seed, ep.stream, item + 1) * (jt - -jt)));
            p.z = (float)((double)p.z + (-jt + draw_unit(ep.seed, ep.stream, item + 2) * (jt - -jt)));
        }
        if (nodal_trilinear(solid, Dims{g.I + 1, g.J + 1, g.K + 1}, g.dx, g.invdx, p) > 0) {
            // ParticleMaskGrid::isSubCellSet (particlemaskgrid.cpp:75-105).
            I3 s = pos_to_index(p, 1.0 / (0.5 * g.dx));
            size_t mc = (size_t)(s.i / 2) + (size_t)g.I * ((size_t)(s.j / 2) + (size_t)g.J * (s.k / 2));
            uint32_t byte = (mask[mc >> 2] >> (8 * (mc & 3))) & 0xffu;
            // _addMarkerParticle keeps particles inside the grid
            // (fluidsimulation.cpp:2637-2642).
            I3 c = pos_to_index(p, g.invdx);
            bool in = c.i >= 0 && c.j >= 0 && c.k >= 0 && c.i < g.I && c.j < g.J && c.k < g.K;
            keep = (!(byte & subcell_bit(s)) && in) ? 1 : 0;
        }
    }
    cand[3 * (size_t)q] = p.x;
    cand[3 * (size_t)q + 1] = p.y;
    cand[3 * (size_t)q + 2] = p.z;
    flags[q] = keep;
}

__global__ __launch_bounds__(256) void k_emit_scatter(GridGeom g, InflowView v,
                                                      const float *__restrict__ cand,
                                                      const int *__restrict__ flags,
                                                      const int *__restrict__ pos,
                                                      const int *__restrict__ nbase, EmitOut out,
                                                      uint32_t *mask) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 8 * v.ncells || !flags[q]) return;
    size_t dst = (size_t)*nbase + pos[q];
    F3 p = f3(cand[3 * (size_t)q], cand[3 * (size_t)q + 1], cand[3 * (size_t)q + 2]);
    out.px[dst] = p.x; out.py[dst] = p.y; out.pz[dst] = p.z;
    out.vx[dst] = v.vel[0]; out.vy[dst] = v.vel[1]; out.vz[dst] = v.vel[2];
    mask_add(mask, g, p, 1.0 / (0.5 * g.dx));
}

__global__ void k_emit_advance(int m, const int *__restrict__ flags, const int *__restrict__ pos,
                               int *nbase) {
    *nbase += pos[m - 1] + flags[m - 1];
}

// FluidSimulation::_updateOutflowMeshFluidSource (fluidsimulation.cpp:
// 4608-4673): a particle in an outflow cell (a source cell, or any other
// cell when inversed) goes where the source SDF, sampled with the sub-grid
// offset, is negative (not negative when inversed).
__global__ __launch_bounds__(256) void k_outflow_keep(GridGeom g, const float *__restrict__ px,
                                                      const float *__restrict__ py,
                                                      const float *__restrict__ pz, int n,
                                                      InflowViews views, int *keep) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    F3 p = f3(px[q], py[q], pz[q]);
    I3 c = pos_to_index(p, g.invdx);
    const size_t cell = (size_t)c.i + (size_t)g.I * ((size_t)c.j + (size_t)g.J * c.k);
    int stay = 1;
    for (int s = 0; s < views.n; s++) {
        const InflowView &v = views.v[s];
        if (!v.outflow) continue;
        const bool in = v.cells[cell] != 0;
        if (v.inversed ? in : !in) continue;
        float d = (float)nodal_trilinear(v.sdf, Dims{v.ni, v.nj, v.nk}, g.dx, g.invdx,
                                         p - f3(v.off[0], v.off[1], v.off[2]));
        if (v.inversed ? d >= 0.0f : d < 0.0f) stay = 0;
    }
    keep[q] = stay;
}

hipError_t launch_outflow_keep(const GridGeom &g, const float *px, const float *py, const float *pz, int n,
                               const InflowViews &views, int *keep, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_outflow_keep<<<(n + 255) / 256, 256, 0, s>>>(g, px, py, pz, n, views, keep);
    return hipGetLastError();
}

hipError_t launch_inflow_mask(const GridGeom &g, const float *px, const float *py, const float *pz,
                              int n, const InflowViews &views, uint32_t *mask, hipStream_t s) {
    for (int q = 0; q < views.n; q++) {
        if (!views.v[q].emit) continue;
        if (views.v[q].ncells == 0) continue;
        k_mask_clear<<<(views.v[q].ncells + 255) / 256, 256, 0, s>>>(views.v[q].list,
                                                                     views.v[q].ncells,
                                                                     (uint8_t *)mask);
    }
    if (n > 0) k_mask_particles<<<(n + 255) / 256, 256, 0, s>>>(g, px, py, pz, n, views, mask);
    return hipGetLastError();
}

hipError_t launch_inflow_emit(const GridGeom &g, const InflowView &v, const float *solid,
                              uint32_t *mask, const EmitParams &ep, float *cand, int *flags,
                              int *pos, int *scan_tmp, int *nbase, const EmitOut &out,
                              hipStream_t s) {
    int m = 8 * v.ncells;
    if (m <= 0) return hipSuccess;
    unsigned nb = (unsigned)((m + 255) / 256);
    k_emit_flags<<<nb, 256, 0, s>>>(g, v, solid, mask, ep, cand, flags);
    hipError_t e = launch_exclusive_scan(flags, pos, m, scan_tmp, s);
    if (e != hipSuccess) return e;
    k_emit_scatter<<<nb, 256, 0, s>>>(g, v, cand, flags, pos, nbase, out, mask);
    k_emit_advance<<<1, 1, 0, s>>>(m, flags, pos, nbase);
    return hipGetLastError();
}

}  // namespace flipmi
