// Kernel launch interface between the host orchestration (flipmi_scene.cpp)
// and the HIP kernels (k_*.hip).  Every launcher enqueues on the given
// stream and returns the hipError_t of the launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flipmi {

// One triangle of the domain box with its exact-band node range
// (MeshLevelSet::_initializeTriangleData, meshlevelset.cpp:874-915).
struct BoxTri {
    float v[3][3];
    int gmin[3], gmax[3];
};

struct BoxMesh {
    BoxTri tri[12];
    int ntri;
    float bpos[3];   // boundary AABB min corner (float)
    double bext[3];  // boundary AABB extents (double)
};

// A mesh triangle for the ray-column inside test: vertices in the test's
// local frame and the range of ray columns its bounding box covers.
struct RayTri {
    float v[3][3];
    int c0[2], c1[2];
};
constexpr int kMaxRayCrossings = 64;  // crossings one ray column may collect

// Grid geometry shared by all kernels.
struct GridGeom {
    int I, J, K;
    double dx, invdx;
};

// ---- setup (k_setup.hip) ----
hipError_t launch_domain_solid(const GridGeom &g, const BoxMesh &mesh, float *phi_nodal,
                               hipStream_t s);
hipError_t launch_weights(const GridGeom &g, const float *phi_nodal, float *wU, float *wV,
                          float *wW, float *wC, hipStream_t s);
hipError_t launch_near_solid(const GridGeom &g, const float *phi_nodal, uint8_t *near_a,
                             uint8_t *near_b, int ci, int cj, int ck, int feather_layers,
                             uint8_t **result, hipStream_t s);

// ---- sort (k_sort.hip) ----
struct SortBuffers {
    int *cell;       // [n] cell id per particle (-1: outside the grid)
    int *slot;       // [n] arrival slot within the cell
    int *count;      // [ncell]
    int *offset;     // [ncell + 1]
    int *order;      // [n] particle ids grouped by cell, ascending within a cell
    int *scan_tmp;   // scratch for the scan
    size_t scan_tmp_elems;
    // When set: a flag raised if some cell holds more than 250 particles;
    // without that, only the counts and offsets are formed (order is left
    // as it was).
    int *over = nullptr;
};
size_t scan_scratch_elems(int n);
hipError_t launch_exclusive_scan(const int *in, int *out, int n, int *tmp, hipStream_t s);
hipError_t launch_sort_by_cell(const GridGeom &g, const float *px, const float *py,
                               const float *pz, int n, SortBuffers &sb, hipStream_t s);
// Half-cell bins: 2I x 2J x 2K, sx fastest; a bin's cell is (sx>>1, sy>>1, sz>>1).
inline size_t subcell_bins(const GridGeom &g) { return (size_t)8 * g.I * g.J * g.K; }
// Counting sort into half-cell bins (count/offset of subcell_bins(g) (+1)
// entries; sb.cell holds the bin, sb.over is not used).  Order within a bin
// is arbitrary.
hipError_t launch_sort_by_subcell(const GridGeom &g, const float *px, const float *py, const float *pz, int n,
                                  SortBuffers &sb, hipStream_t s);
// Count grid of the removal pass: removed_in_cell[c] = solid-removed
// particles in cell c (cleared first), then out = min(count - removed, cap).
hipError_t launch_removed_in_cell(const GridGeom &g, const float *px, const float *py, const float *pz,
                                  const uint8_t *removed, int n, int *removed_in_cell, hipStream_t s);
hipError_t launch_removal_counts(const GridGeom &g, const int *count, const int *removed_in_cell, int cap,
                                 int *out, hipStream_t s);
// Gathers the first *nsorted (= the sort's last offset) entries of the
// order from the SoA particles src (x, y, z, u, v, w): spos = (x, y, z,
// particle id as int bits), svel = the velocity components.
hipError_t launch_gather_sorted(const int *order, const int *nsorted, int n, const float *const src[6],
                                float4 *spos, float *const svel[3], hipStream_t s);

// ---- classification (k_classify.hip) ----
// From particles sorted into half-cell bins (spos from launch_gather_sorted,
// sub_offset: subcell_bins(g) + 1).
hipError_t launch_liquid_sdf(const GridGeom &g, const float4 *spos, const int *sub_offset, double radius,
                             const float *solid_phi_nodal, float *phi, hipStream_t s);
// ParticleLevelSet::postProcessSignedDistanceField on an existing field.
hipError_t launch_liquid_postprocess(const GridGeom &g, const float *solid_phi_nodal, float *phi,
                                     hipStream_t s);

// ---- P2G (k_p2g.hip) ----
// P2G of the three MAC components from particles sorted into half-cell bins
// (spos / svel from launch_gather_sorted), each face summed in ascending
// particle id.  ovf_count (1 int) and ovf_list (faces of U, V and W
// together) are scratch for faces with more than 32 contributions.
// n = particle count (ids below n).
hipError_t launch_p2g(const GridGeom &g, const float4 *spos, const float *const svel[3], const int *sub_offset,
                      int n, double radius, float *const out[3], uint8_t *const valid[3], int *ovf_count,
                      int *ovf_list, hipStream_t s);
// Faces of U, V and W together (the size of launch_p2g's overflow list).
inline size_t p2g_face_count(const GridGeom &g) {
    return (size_t)(g.I + 1) * g.J * g.K + (size_t)g.I * (g.J + 1) * g.K + (size_t)g.I * g.J * (g.K + 1);
}

// ---- grid passes (k_grid.hip) ----
hipError_t launch_extrapolate(int ni, int nj, int nk, float *grid, const uint8_t *valid,
                              uint8_t *status, int layers, hipStream_t s);
// Looks up the device address of the no-allocation scratch array on
// `device` (current) once; noalloc.hip.
hipError_t noalloc_scratch_init(int device);
hipError_t launch_save_and_body_force(size_t n, const float *mac, float *saved, float add,
                                      int do_add, hipStream_t s);
hipError_t launch_constrain(size_t n, float *mac, const float *weight, const float *solid_vel,
                            hipStream_t s);
hipError_t launch_fill_u8(uint8_t *p, size_t n, uint8_t v, hipStream_t s);
hipError_t launch_fill_f32(float *p, size_t n, float v, hipStream_t s);
hipError_t launch_fill_f64(double *p, size_t n, double v, hipStream_t s);
hipError_t launch_fill_i32(int *p, size_t n, int v, hipStream_t s);

// ---- pressure (k_pressure.hip) ----
struct PcgScalars {
    double rho, alpha, beta, resid, tol, bnorm, sz;
    float alpha_f, beta_f;
    int iters, done, converged, failed, abort, maxiter, precision;
    unsigned int counter[4];
    // Per-XCD arrival counts for the last-block tickets of counter[1] and
    // counter[2], one 128-byte line each (atomics on one line serialise).
    alignas(128) unsigned int group_counter[2][8][32];
    unsigned int rows;                  // rows of the system (fluid interior cells)
    unsigned long long absmax_bits;
    double mf;                          // dt / dx^2 of the assembled system (coupling scale)
    // mixed precision: true-residual replacement ("reliable update")
    double rmax;                        // |r| at the last true residual
    double delta;                       // replace when |r32| < delta * rmax
    int refresh;                        // a true residual is due
    int refinements;                    // true residuals computed
    // Epoch of the next sweep launch: read by every workgroup of a sweep,
    // advanced by its last tile, so launches replayed from a captured graph
    // still tag their records afresh.  Starts at 1 (records start at 0).
    unsigned int sweep_epoch;
};

// Sweep-ordered ("slot") storage of the pressure system.  Cells are grouped
// into 8 x 8 column tiles over (j, k); a cell of tile (jt, kt) has lane
// a + 8 b (a = j % 8, b = k % 8) and wavefront step t = i + a + b,
// 0 <= t < S (S >= I + 14, a multiple of 32).  Steps are stored in groups of
// four: the slot of (tile, t, lane) is
//     ((tile * S/4 + t/4) * 64 + lane) * 4 + t % 4,
// so a wave that owns a tile loads four steps of all its 64 cells with one
// contiguous 16-byte-per-lane access (two for FP64), and each cell's -x, -y
// and -z neighbours of the same tile sit one step earlier in lanes l, l - 1
// and l - 8.  Slots that are not grid cells (skew corners, step padding,
// padding of J and K up to a multiple of 8) hold zeros in every array.
struct SlotMap {
    int I, J, K, ntj, ntk, S;
    __host__ __device__ size_t count() const { return (size_t)ntj * ntk * S * 64; }
    // Elements allocated per slot array: count() plus two trailing groups,
    // a zero group (read in place of a group without rows, never written)
    // and a sink group (written in place of such a group, never read).
    __host__ __device__ size_t alloc_count() const { return count() + 512; }
    __host__ __device__ size_t zero_group() const { return count(); }
    __host__ __device__ size_t sink_group() const { return count() + 256; }
    __host__ __device__ size_t slot(int i, int j, int k) const {
        int a = j & 7, b = k & 7, t = i + a + b;
        size_t tile = (size_t)(j >> 3) + (size_t)ntj * (size_t)(k >> 3);
        size_t grp = tile * (size_t)(S >> 2) + (size_t)(t >> 2);
        return ((grp * 64 + (size_t)(a + 8 * b)) << 2) + (size_t)(t & 3);
    }
    // slot() in 32-bit arithmetic, for kernels (the store holds at most 2^32
    // slots: PressureStore::alloc checks).
    __host__ __device__ unsigned slot32(int i, int j, int k) const {
        const unsigned a = (unsigned)(j & 7), b = (unsigned)(k & 7), t = (unsigned)i + a + b;
        const unsigned tile = (unsigned)(j >> 3) + (unsigned)ntj * (unsigned)(k >> 3);
        const unsigned grp = tile * ((unsigned)S >> 2) + (t >> 2);
        return ((grp * 64u + a + 8u * b) << 2) + (t & 3u);
    }
};
inline SlotMap make_slot_map(int I, int J, int K) {
    SlotMap m;
    m.I = I; m.J = J; m.K = K;
    m.ntj = (J + 7) / 8; m.ntk = (K + 7) / 8;
    m.S = (I + 14 + 31) & ~31;
    return m;
}

struct PressureBuffers {
    SlotMap sm;
    double *diag;                       // matrix: diagonal
    // Couplings to the +x/+y/+z rows, stored as the FP32 face weight w each
    // is formed from: entry = -(double)w * mf (PcgScalars::mf), bit for bit
    // the FP64 entry; -0.0f where a slot has no such coupling (entry +0.0).
    float *wx, *wy, *wz;
    double *rhs;
    double *x, *r, *z, *s;              // PCG vectors (FP64)
    double *invdiag, *Lx, *Ly, *Lz;     // MIC(0) factor: column entries (rows +x/+y/+z)
    double *Lxm, *Lym, *Lzm;            // the same entries read from the row: columns -x/-y/-z
    // Mixed precision (null unless allocated): FP32 copies of the factor and
    // the FP32 PCG vectors; x and the matrix stay FP64.
    float *invdiag32, *Lx32, *Ly32, *Lz32, *Lxm32, *Lym32, *Lzm32;
    float *r32, *z32, *s32;
    // Mixed precision: the SpMV forms the new direction s = z + beta s from
    // the old one, so s lives in two buffers: s32 is read, s32b written, and
    // the host swaps them after each SpMV launch.  s_plain: the next SpMV
    // reads s32 as it is (right after initialisation).
    float *s32b;
    int s_plain;
    float *q32;  // mixed precision: q = A s, read by the axpy
    double *partials;                   // per-block partial sums
    int npartials;
    unsigned long long *edge;           // tagged values a tile hands to its j / k neighbours
    // Groups (4 steps x 64 lanes) holding at least one row: bit g % 64 of
    // word g / 64 of the tile's mask_words words (null: treat all as active).
    unsigned long long *group_mask;
    int mask_words;
    // The active groups in ascending order (rebuilt with the mask for every
    // solve) and their number; the per-group PCG kernels walk this list.
    int *active_list;
    int *active_count;
    unsigned int epoch;                 // sweep launches so far (host count, an upper bound of sweep_epoch)
    unsigned int *progress;             // factorisation: tile progress counters
    int ntiles_j, ntiles_k;
    unsigned int *ticket;
    int *tile_irange;                   // per (j,k) tile: min/max i of its rows (-1: none)
    PcgScalars *sc;                     // device scalars
};

hipError_t launch_pressure_assemble(const GridGeom &g, double dt, const float *liquid,
                                    const float *wU, const float *wV, const float *wW,
                                    const float *wC, const float *u, const float *v,
                                    const float *w, const float *sU, const float *sV,
                                    const float *sW, PressureBuffers &pb, hipStream_t s);
hipError_t launch_pressure_reset(PressureBuffers &pb, hipStream_t s);
// PressureSolver::_conditionSolidVelocityField: zeroes the solid face
// velocities (sU, sV, sW, in place) around isolated fluid pockets.  Scratch:
// parent and flags, one element per cell each.
hipError_t launch_condition_solid(const GridGeom &g, const float *liquid, const float *wU, const float *wV,
                                  const float *wW, float *sU, float *sV, float *sW, int *parent,
                                  unsigned *flags, hipStream_t s);
// Marks the slot groups that hold a row (after assembly).
hipError_t launch_group_mask(PressureBuffers &pb, hipStream_t s);
// Zeroes the factor, finds each column tile's row extent and factorises.
hipError_t launch_mic0_factor(const GridGeom &g, const float *liquid, PressureBuffers &pb,
                              hipStream_t s);
// mixed: FP32 inner iteration with FP64 dots, FP64 x and true-residual
// replacement (requires the FP32 arrays of pb).
hipError_t launch_pcg_init(const GridGeom &g, double tolf, int maxiter, bool mixed,
                           PressureBuffers &pb, hipStream_t s);
hipError_t launch_pcg_iteration(const GridGeom &g, bool mixed, PressureBuffers &pb, hipStream_t s);
// The kernels of one PCG iteration.  The FP64 path runs spmv, axpy,
// forward, backward, update.  The mixed path folds the axpy into the forward
// sweep and the direction update into the SpMV, and runs kPcgResidual (a
// true residual and a forward sweep over it) between the forward and the
// backward sweep; it exits at once unless the forward sweep's residual norm
// called for a true residual.
enum {
    kPcgSpmv = 0, kPcgAxpy = 1, kPcgForward = 2, kPcgBackward = 3, kPcgUpdate = 4,
    kPcgResidual = 5, kPcgPieces = 6
};
bool pcg_piece_used(int piece, bool mixed);
// Sweep launches per iteration (each takes a record epoch).
inline unsigned pcg_sweeps_per_iteration(bool mixed) { return mixed ? 3u : 2u; }
int pcg_piece_at(int q);  // piece executed q-th within an iteration
// With kt.start set, the kernel's dispatch records its start and end in the
// two events.
struct KernelTiming {
    hipEvent_t start = nullptr, stop = nullptr;
};
hipError_t launch_pcg_piece(int piece, const GridGeom &g, bool mixed, PressureBuffers &pb,
                            hipStream_t s, const KernelTiming &kt);
// p is in slot order (pb.sm).
hipError_t launch_pressure_apply(const GridGeom &g, const SlotMap &sm, double dt, const float *liquid,
                                 const float *wU, const float *wV, const float *wW,
                                 const double *p, float *u, float *v, float *w, uint8_t *vu,
                                 uint8_t *vv, uint8_t *vw, hipStream_t s);
hipError_t launch_precond_apply(const GridGeom &g, PressureBuffers &pb, const double *r,
                                double *z, hipStream_t s);
hipError_t launch_spmv(const GridGeom &g, PressureBuffers &pb, const double *x, double *y,
                       hipStream_t s);
// Number of 64-bit words of pb.edge.
size_t edge_record_count(const SlotMap &sm);

// ---- particles (k_particles.hip) ----
// ---- mesh inflow sources (k_source.hip) ----
// Device view of one inflow source.  sdf is nodal on the source's sub-grid
// (ni x nj x nk nodes) whose origin is `off` in grid coordinates; cells is
// the I*J*K byte grid of source cells, cmin/cmax their index box, list the
// ncells source cells in raster order.
struct InflowView {
    const float *sdf;
    int ni, nj, nk;
    float off[3];
    const uint8_t *cells;
    int cmin[3], cmax[3];
    const int *list;
    int ncells;
    float vel[3];
    int emit;       // enabled: emits this substep
    int constrain;  // enabled with constrained fluid velocity
    int outflow;    // enabled outflow source: removes particles
    int inversed;   // the outflow acts outside the mesh
};
constexpr int kMaxInflow = 8;
struct InflowViews {
    int n;
    InflowView v[kMaxInflow];
};
struct EmitParams {
    double jitter;       // 0.25 * (jitter factor - 1e-3) * dx
    int jitter_surface;  // jitter every new particle, not only the deep ones
    uint64_t seed, stream;
};
struct EmitOut {
    float *px, *py, *pz, *vx, *vy, *vz;
};
hipError_t launch_mesh_exact_band(const BoxTri *tris, int ntri, int ni, int nj, int nk, double dx,
                                  float upper, float *phi, hipStream_t s);
hipError_t launch_ray_inside(const RayTri *tris, int ntri, int ci, int cj, int ck, double dx,
                             const float jitter[3], const int go[3], int ni, int nj, int nk,
                             int mode, uint8_t *nodes, float *phi, int *overflow, hipStream_t s);
// Cells around inside nodes: byte grid, scan flags/positions and count.
hipError_t launch_source_cell_scan(const uint8_t *nodes, int I, int J, int K, uint8_t *cells,
                                  int *flags, int *pos, int *scan_tmp, int *count, hipStream_t s);
hipError_t launch_source_cell_list(const int *flags, const int *pos, size_t n, int *list,
                                   hipStream_t s);
hipError_t launch_negate(float *phi, size_t n, hipStream_t s);
hipError_t launch_union_min(const float *phi, int ni, int nj, int nk, const int go[3], int Ni, int Nj,
                            int Nk, float *dst, hipStream_t s);
hipError_t launch_inflow_mask(const GridGeom &g, const float *px, const float *py, const float *pz,
                              int n, const InflowViews &views, uint32_t *mask, hipStream_t s);
hipError_t launch_inflow_emit(const GridGeom &g, const InflowView &v, const float *solid,
                              uint32_t *mask, const EmitParams &ep, float *cand, int *flags,
                              int *pos, int *scan_tmp, int *nbase, const EmitOut &out,
                              hipStream_t s);
// Keep flags of the outflow sources (1 = the particle stays).
hipError_t launch_outflow_keep(const GridGeom &g, const float *px, const float *py, const float *pz, int n,
                               const InflowViews &views, int *keep, hipStream_t s);
// Body force on one MAC component that skips the valid faces an inflow
// source constrains (fluidsimulation.cpp:3372-3438, 3450-3489).
hipError_t launch_save_and_body_force_inflow(const GridGeom &g, int comp, const float *mac,
                                             float *saved, float add, int do_add,
                                             const uint8_t *valid, const InflowViews &views,
                                             hipStream_t s);

struct ParticleStepArgs {
    float *px, *py, *pz, *vx, *vy, *vz;  // in/out, canonical order
    int n;
    const float *u, *v, *w;              // projected MAC
    const float *su, *sv, *sw;           // saved (pre-force) MAC
    const float *solid;                  // nodal solid phi
    const uint8_t *near_solid;
    int nci, ncj, nck;
    double near_cell;                    // near-solid coarse cell size
    float pic_ratio_f, flip_ratio_f;
    double dt;
    double cfl;
    float bpos[3];                       // collision boundary (already shrunk)
    double bext[3];
    uint8_t *solid_removed;              // [n] out
    int *speed_hist;                     // [maxsteps] out (accumulated)
    double speed_step;
    int maxsteps;
    InflowViews inflow;                  // sources whose particles take their velocity
};
hipError_t launch_particle_step(const GridGeom &g, const ParticleStepArgs &a, hipStream_t s);
hipError_t launch_max_speed(const float *vx, const float *vy, const float *vz, int n,
                            unsigned int *out_bits, hipStream_t s);
hipError_t launch_cell_rank_cells(const int *offset, int ncell, const int *order,
                                  const uint8_t *solid_removed, const int *over, int *rank,
                                  hipStream_t s);
hipError_t launch_keep_flags(const float *vx, const float *vy, const float *vz, int n,
                             const uint8_t *solid_removed, const int *rank, double maxspeedsq,
                             int *keep, hipStream_t s);
hipError_t launch_compact_keep(const int *keep, int n, int *pos, int *scan_tmp,
                               const float *const src[6], float *const dst[6], int *new_count,
                               hipStream_t s);
hipError_t launch_advect_rk(const GridGeom &g, const float *u, const float *v, const float *w,
                            const float *pin, int n, double dt, int order, float *pout,
                            hipStream_t s);
hipError_t launch_trilinear_velocity(const GridGeom &g, const float *u, const float *v,
                                     const float *w, const float *pin, int n, float *out,
                                     hipStream_t s);

}  // namespace flipmi
