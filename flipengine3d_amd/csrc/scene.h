// Host-side scene: device-resident state of one FLIP simulation and the
// FluidSimulation::update(dt) substep loop that drives the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/flipmi.h"
#include "flipmi_kernels.h"

namespace flipmi {

struct Ctx {
    int device = 0;
    std::string last_error;
    hipStream_t stream = nullptr;  // stream for stage drop-ins
};

// Device buffer that frees itself.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() {}
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    hipError_t alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void **)&p, sizeof(T) * count);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return sizeof(T) * n; }
};

struct StatusError {
    int code;
    std::string msg;
};

// Builds the domain wall box mesh and the per-triangle exact-band node
// ranges of FluidSimulation::_getBoundaryAABB / _getTriangleMeshFromAABB /
// MeshLevelSet::_initializeTriangleData (fluidsimulation.cpp:2805-2845,
// meshlevelset.cpp:874-915).
BoxMesh make_domain_box(int isize, int jsize, int ksize, double dx);

// Collision boundary: the wall box shrunk by the 0.1 dx solid buffer
// (fluidsimulation.cpp:4213-4215).
void collision_box(int isize, int jsize, int ksize, double dx, float pos[3], double ext[3]);

// One mesh inflow source: MeshFluidSource with a static mesh, inflow
// (meshfluidsource.cpp:32-44, 74-76, 114-116), in grid coordinates.
struct InflowSource {
    std::vector<float> verts;  // xyz
    std::vector<int> tris;     // three vertex indices per triangle
    float vel[3] = {0, 0, 0};
    bool enabled = true;       // enable() / disable()
    bool constrain = true;     // _isConstrainedFluidVelocity (meshfluidsource.h:124)
    int emissions = 1;         // _substepEmissions (meshfluidsource.h:119)
    bool outflow = false;      // setOutflow() (meshfluidsource.cpp:82-84)
    bool inversed = false;     // outflowInverse() (meshfluidsource.cpp:158-160)
    // Placement of the source SDF (MeshFluidSource::update, :178-222).
    int go[3] = {0, 0, 0};     // sub-grid offset in cells
    float off[3] = {0, 0, 0};  // its position
    int ni = 0, nj = 0, nk = 0;  // sub-grid node dims
    // Device state, built at the first substep that needs it.
    bool built = false;
    DBuf<float> sdf;
    DBuf<uint8_t> cells;
    DBuf<int> list;
    int ncells = 0;
    int cmin[3] = {0, 0, 0}, cmax[3] = {-1, -1, -1};
};

// One static mesh obstacle: MeshObject with updateMeshStatic, added with
// FluidSimulation::addMeshObstacle (meshobject.cpp:50-59,
// fluidsimulation.cpp:1994-2008), in grid coordinates.  Only the vertices
// some triangle uses are kept (TriangleMesh::removeExtraneousVertices).
struct ObstacleMesh {
    std::vector<float> verts;
    std::vector<int> tris;
    bool enabled = true;
    bool inversed = false;  // MeshObject::inverse() (meshobject.cpp:293-295): a container
};

struct SurfaceStore;

struct Scene {
    Ctx *ctx = nullptr;
    flipmi_scene_desc desc{};
    GridGeom g{};
    hipStream_t stream = nullptr;
    bool own_stream = false;

    // particles (canonical order) and work buffers
    int n = 0, cap = 0;
    DBuf<float> part[6], part_alt[6];
    DBuf<float4> spos;     // half-cell order: (x, y, z, particle id bits)
    DBuf<float> svel[3];   // half-cell order: velocity
    DBuf<int> cell, slot, count, offset, order, scan_tmp;
    // Half-cell bins of the step's particle sort (8 I J K, +1), read by the
    // liquid SDF and P2G; P2G's list of faces with more than 32 contributions.
    DBuf<int> sub_count, sub_offset, p2g_ovf, p2g_ovf_count;
    DBuf<int> keep, keep_pos, rank, new_count, over_cap;
    DBuf<uint8_t> solid_removed;
    // Solid-removed particles per cell at the last removal pass; with `count`
    // it gives the removal pass's count grid (FLIPMI_GRID_CELL_COUNT).
    DBuf<int> removed_in_cell;
    bool removal_counts_ready = false;
    DBuf<int> speed_hist;
    DBuf<unsigned> max_speed_bits;

    // grids
    DBuf<float> solid, wU, wV, wW, wC, sU, sV, sW;
    DBuf<uint8_t> near_a, near_b;
    uint8_t *near = nullptr;
    int nci = 0, ncj = 0, nck = 0;
    DBuf<float> liquid, u, v, w, su, sv, sw;
    DBuf<uint8_t> vu, vv, vw, status;

    // pressure
    struct PressureStore *press = nullptr;

    // pinned host staging
    void *host_stage = nullptr;

    // host-side state mirroring FluidSimulation's update bookkeeping
    bool initialized = false, solid_ready = false, weights_ready = false;
    bool output_initialized = false;
    int current_frame = 0, step_number = 0;
    double frame_dt = 0, remaining = 0, cur_step = 0;
    double radius = 0;
    std::vector<float> pending_pos, pending_vel;  // loaded, not yet in the scene
    // setDomainOffset / setDomainScale (fluidsimulation.h:1612-1613):
    // particle positions cross the boundary in domain coordinates.
    float domain_offset[3] = {0.0f, 0.0f, 0.0f};
    double domain_scale = 1.0;

    // mesh inflow sources and the marker-particle jitter
    // (setMarkerParticleJitterFactor, enable/disableJitterSurfaceMarkerParticles,
    // fluidsimulation.cpp:185-210)
    std::vector<std::unique_ptr<InflowSource>> sources;
    std::vector<std::unique_ptr<ObstacleMesh>> obstacles;
    double jitter_factor = 1.0;
    bool jitter_surface = false;
    uint64_t jitter_seed = 0;
    uint64_t emission_count = 0;  // emission passes so far (keys the jitter)
    DBuf<uint32_t> inflow_mask;   // ParticleMaskGrid bytes, four cells per word
    DBuf<float> emit_cand;
    DBuf<int> emit_flags, emit_pos, emit_scan, emit_nbase;

    // Surface reconstruction (fluidsimulation.h:1586-1611) and the last mesh.
    flipmi_surface_desc surface{};
    SurfaceStore *mesh = nullptr;  // owned

    // PCG iteration batches launched as captured HIP graphs.
    bool launch_graphs = false;

    // last pressure solve
    int last_iters = 0, last_ok = 1;
    double last_resid = 0;
    int64_t last_rows = 0;
    int last_refinements = 0;

    // profiling
    int profiling = 0;  // 0 off, 1 stage events, 2 also each selected PCG kernel
    unsigned kernel_mask = ~0u;  // PCG kernels timed at level 2 (bit = kPcg* piece)
    int kernel_every = 1;         // level 2: time them on every n-th PCG iteration
    int64_t pcg_iteration_count = 0;  // PCG iterations launched, for kernel_every
    double stage_ms[FLIPMI_NUM_STAGES] = {0};
    int64_t stage_calls[FLIPMI_NUM_STAGES] = {0};
    struct PendingEvent {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<PendingEvent> pending_events;
    std::vector<hipEvent_t> event_pool;

    ~Scene();
    void check(hipError_t e, const char *what);
    void alloc_all();
    void load(const float *pos, const float *vel, int64_t count);
    void initialize();
    // Domain to grid coordinates and the grid-box filter of the pending
    // particles (FluidSimulation::_loadMarkerParticles).
    void admit_pending();
    void upload_pending();
    // Grid to domain coordinates of n positions (xyz), in place.
    void to_domain(float *pos_xyz, int64_t n) const;
    // Reuses an initialized scene for a new particle set on the same grid and
    // settings (the state right after initialize()); false if the set does not
    // fit the particle capacity.
    bool rearm(const float *pos, const float *vel, int64_t count);
    size_t device_bytes() const;
    flipmi_frame_stats update(double dt);
    void step_fluid(double dt);
    double next_time_step(double dt);
    void ensure_solid();
    void pressure_solve(double dt);
    void advance_particles(double dt);
    void sort_particles(int *over = nullptr);
    void sort_subcells();
    // The isomesh of the current particles (_outputSurfaceMeshThread).
    void build_mesh();
    void extrapolate_all(int layers);
    // Mesh inflow sources: add (checks the mesh), build the device state of
    // every source, their device views, and the emission at the end of a
    // substep (FluidSimulation::_updateFluidObjects).
    int add_inflow(const float *verts, int nv, const int *tris, int nt, const double vel[3]);
    void ensure_inflow();
    void build_inflow(InflowSource &src, int id);
    InflowViews inflow_views(bool emitting) const;
    bool any_inflow_constraint() const;
    void emit_inflow();
    // FluidSimulation::_updateOutflowMeshFluidSources at the end of a substep.
    void remove_outflow();
    // Static mesh obstacles: add (checks the mesh) and their union into the
    // solid SDF (FluidSimulation::_addStaticObjectsToSDF).
    int add_obstacle(const float *verts, int nv, const int *tris, int nt);
    void union_obstacles();
    // Grows every particle-sized buffer to hold at least `need` particles.
    void grow(int64_t need);
    void begin_stage(int stage, hipEvent_t *a);
    void end_stage(int stage, hipEvent_t a);
    // Books a pair of events already recorded (by a timed dispatch).
    void add_timed(int stage, hipEvent_t a, hipEvent_t b);
    void collect_events();
    hipEvent_t take_event();
    void sync();
};

// Pressure solve over caller grids already on the device (used by the
// PressureSolver drop-in and by Scene::pressure_solve).
// PCG iteration batches captured as HIP graphs, one per (batch size, s32
// buffer the batch starts from); replayed instead of launching each kernel.
struct PcgGraphs {
    std::map<std::pair<int, const void *>, hipGraphExec_t> exec;
    PcgGraphs() {}
    PcgGraphs(const PcgGraphs &) = delete;
    PcgGraphs &operator=(const PcgGraphs &) = delete;
    ~PcgGraphs() { clear(); }
    void clear() {
        for (auto &e : exec) (void)hipGraphExecDestroy(e.second);
        exec.clear();
    }
};

struct PressureWork {
    GridGeom g;
    double dt, tol, acceptable;
    int maxiter, precision;
    const float *liquid, *wU, *wV, *wW, *wC, *sU, *sV, *sW;
    float *u, *v, *w;
    uint8_t *vu, *vv, *vw;
    PcgGraphs *graphs = nullptr;  // non-null: iteration batches run as captured graphs
};
struct PressureOutcome {
    int iterations = 0;
    double residual = 0;
    bool success = true;
    bool solved = false;  // the pressure update was applied
    int refinements = 0;
    int64_t rows = 0;
};
PressureOutcome run_pressure(const PressureWork &w, PressureBuffers &pb, PcgScalars *host_sc,
                             hipStream_t s, Scene *prof);

// Device storage of one pressure system in slot order: arr[0..12] are the
// FP64 slot arrays (diag, rhs, x, r, z, s, invdiag, Lx, Ly, Lz, Lxm, Lym,
// Lzm), arr[13] the reduction partials, wgt[0..2] the coupling weights
// (wx, wy, wz).
struct PressureStore {
    static constexpr int kSlotArrays = 13;
    static constexpr int kWeightArrays = 3;
    static constexpr int kSlotArrays32 = 12;  // FP32 factor and vectors, s twice, q (mixed precision)
    DBuf<double> arr[kSlotArrays + 1];
    DBuf<float> wgt[kWeightArrays];
    DBuf<float> arr32[kSlotArrays32];
    DBuf<unsigned long long> edge, mask;
    DBuf<unsigned> progress, ticket;
    DBuf<int> irange;
    DBuf<int> alist;  // active group list (groups) and its length (last element)
    DBuf<PcgScalars> scal;
    PressureBuffers pb{};
    PcgGraphs graphs;             // captured iteration batches (pointers into these buffers)
    void alloc(const GridGeom &g, hipStream_t s, bool mixed);
    size_t device_bytes() const;
};

// Device bytes a scene with these settings and particle capacity allocates.
size_t estimate_scene_bytes(const flipmi_scene_desc &d, int64_t cap);

// Argument checks of flipmi_scene_create; throws StatusError.
void validate_scene_desc(const flipmi_scene_desc &d);
GridGeom make_geom(int I, int J, int K, double dx);

}  // namespace flipmi

// The opaque handle of the C ABI.
struct flipmi_ctx {
    flipmi::Ctx c;
};
