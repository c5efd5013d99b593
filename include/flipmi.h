/*
 * flipmi -- MI355X-native FLIP fluid step behind a plain C ABI.
 *
 * Every entry point replaces one call of the reference engine
 * (jklae/FLIPEngine3D, src/engine) and keeps its argument meaning and error
 * behaviour; the citation above each declaration names the reference
 * interface it stands in for.  No C++ or framework types cross the ABI:
 * plain pointers (host memory unless stated), sizes and POD structs.
 *
 * Grids are flat float/byte arrays, i fastest (Array3d layout,
 * src/engine/array3d.h:425-428):
 *   cell grid         I x J x K
 *   MAC U / V / W     (I+1) x J x K,  I x (J+1) x K,  I x J x (K+1)
 *   nodal solid SDF   (I+1) x (J+1) x (K+1)
 * Particle data are xyz float triples, byte-compatible with the reference's
 * vmath::vec3 blobs (fluidsimulation.cpp:2408-2420, 2488-2509).
 *
 * All functions return a flipmi_status (0 = success).  On failure
 * flipmi_last_error() holds a message.  A scene is used by one host thread
 * at a time; distinct scenes may be driven concurrently.
 */
#ifndef FLIPMI_H
#define FLIPMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLIPMI_ABI_VERSION 1

typedef enum {
    FLIPMI_OK = 0,
    FLIPMI_ERR_INVALID_ARG = -1,    /* std::domain_error in the reference */
    FLIPMI_ERR_NOT_INITIALIZED = -2,/* std::runtime_error "must be initialized" */
    FLIPMI_ERR_OOM = -3,
    FLIPMI_ERR_NOT_CONVERGED = -4,  /* PressureSolver::solve returned false */
    FLIPMI_ERR_HIP = -5,
    FLIPMI_ERR_BUDGET = -6,         /* serving: admission budget exceeded */
    FLIPMI_ERR_TIMEOUT = -7         /* a device-side bounded wait ran out */
} flipmi_status;

typedef enum {
    FLIPMI_PCG_FP64 = 0,  /* reference arithmetic: FP64 vectors, alpha/beta as float */
    FLIPMI_PCG_MIXED = 1  /* FP32 PCG, FP64 dots + iterative refinement */
} flipmi_pcg_precision;

typedef struct flipmi_ctx flipmi_ctx;
typedef struct flipmi_scene flipmi_scene;

/* Scene description: FluidSimulation(isize, jsize, ksize, dx)
 * (fluidsimulation.h:127) plus the setters the workloads use:
 * addBodyForce (fluidsimulation.cpp:1582), setCFLConditionNumber (:1765),
 * setMinTimeStepsPerFrame / setMaxTimeStepsPerFrame (:1801, :1819),
 * setPICFLIPRatio (:1891) and the pressure-solver tolerance / iteration cap
 * (fluidsimulation.h:1657-1659). */
typedef struct {
    int isize, jsize, ksize;
    double dx;
    double body_force[3];
    double cfl;                /* default 5 */
    int min_substeps;          /* default 1 */
    int max_substeps;          /* default 6 */
    double pic_ratio;          /* default 0.05 */
    double pcg_tolerance;      /* default 1e-9 (relative to |b|_inf) */
    int pcg_max_iterations;    /* default 1000 */
    int pcg_precision;         /* flipmi_pcg_precision, default FP64 */
    int64_t max_particles;     /* particle capacity; 0 = size of the loaded set */
} flipmi_scene_desc;

/* FluidSimulationFrameStats (fluidsimulation.h:67-100) plus device timing. */
typedef struct {
    int substeps;
    int64_t num_particles;
    int pcg_iterations;        /* last substep */
    double pcg_residual;       /* last substep, |r|_inf ("Estimated Error") */
    int pressure_ok;           /* 0 if the last pressure solve failed */
    double gpu_ms;             /* device time of the frame */
    int64_t pressure_rows;     /* last substep: rows of the Poisson system */
    int pcg_refinements;       /* last substep, mixed precision: FP64 true residuals */
} flipmi_frame_stats;

/* Per-stage device time accumulated since the last reset (hipEvents). */
#define FLIPMI_NUM_STAGES 24
typedef struct {
    double ms[FLIPMI_NUM_STAGES];
    int64_t calls[FLIPMI_NUM_STAGES];
} flipmi_stage_times;

enum {
    FLIPMI_STAGE_SORT = 0,
    FLIPMI_STAGE_CLASSIFY = 1,
    FLIPMI_STAGE_P2G = 2,
    FLIPMI_STAGE_EXTRAPOLATE = 3,
    FLIPMI_STAGE_ASSEMBLE = 4,
    FLIPMI_STAGE_MIC0 = 5,
    FLIPMI_STAGE_PCG = 6,           /* all PCG iterations of a solve */
    FLIPMI_STAGE_APPLY = 7,
    FLIPMI_STAGE_G2P_ADVECT = 8,
    FLIPMI_STAGE_REMOVE = 9,
    FLIPMI_STAGE_GLUE = 10,         /* save/body force/constrain */
    FLIPMI_STAGE_PCG_ITERS = 11,    /* calls = iterations, ms unused */
    FLIPMI_STAGE_K_SPMV = 12,       /* single kernels of a PCG iteration */
    FLIPMI_STAGE_K_AXPY = 13,
    FLIPMI_STAGE_K_FORWARD = 14,
    FLIPMI_STAGE_K_BACKWARD = 15,
    FLIPMI_STAGE_K_UPDATE = 16,
    FLIPMI_STAGE_K_RESIDUAL = 17,   /* mixed precision: FP64 true residual + forward sweep over it */
    FLIPMI_STAGE_INFLOW = 18,       /* mesh inflow: particle mask and emission */
    FLIPMI_STAGE_SURFACE = 19       /* surface reconstruction (sort, field, marching cubes, smoothing) */
};

/* Grids readable with flipmi_scene_get_grid (sizes as in the header). */
enum {
    FLIPMI_GRID_LIQUID_PHI = 0,
    FLIPMI_GRID_SOLID_PHI = 1,
    FLIPMI_GRID_U = 2,
    FLIPMI_GRID_V = 3,
    FLIPMI_GRID_W = 4,
    FLIPMI_GRID_WEIGHT_U = 5,
    FLIPMI_GRID_WEIGHT_V = 6,
    FLIPMI_GRID_WEIGHT_W = 7,
    FLIPMI_GRID_WEIGHT_C = 8,
    FLIPMI_GRID_VALID_U = 9,   /* uint8 */
    FLIPMI_GRID_VALID_V = 10,  /* uint8 */
    FLIPMI_GRID_VALID_W = 11,  /* uint8 */
    /* int32, I x J x K: the count grid of the last substep's removal pass
     * (FluidSimulation::_removeMarkerParticles, fluidsimulation.cpp:4325-4343):
     * per cell, the advected particles not removed by the solid test, capped
     * at 250.  All zero before the first substep. */
    FLIPMI_GRID_CELL_COUNT = 12
};

/* ---- context ----------------------------------------------------------- */
int flipmi_ctx_create(int device, flipmi_ctx **out);
int flipmi_ctx_destroy(flipmi_ctx *ctx);
const char *flipmi_last_error(const flipmi_ctx *ctx);
void flipmi_scene_desc_default(flipmi_scene_desc *desc);

/* ---- FluidSimulation (fluidsimulation.h:127-153) ------------------------ */
/* FluidSimulation::FluidSimulation(int, int, int, double) */
int flipmi_scene_create(flipmi_ctx *ctx, const flipmi_scene_desc *desc, flipmi_scene **out);
int flipmi_scene_destroy(flipmi_scene *scene);
/* FluidSimulation::loadMarkerParticleData (fluidsimulation.cpp:2488-2509);
 * call before flipmi_scene_initialize, which admits the particles
 * (_loadMarkerParticles, fluidsimulation.cpp:2773-2785: domain to grid
 * coordinates, then the grid-box filter).  vel may be NULL (zero velocity). */
int flipmi_scene_load_particles(flipmi_scene *scene, const float *pos_xyz, const float *vel_xyz,
                                int64_t n);
/* FluidSimulation::initialize (fluidsimulation.cpp:82-88) */
int flipmi_scene_initialize(flipmi_scene *scene);
/* FluidSimulation::update(double dt) (fluidsimulation.cpp:5755-5846).
 * Returns FLIPMI_ERR_NOT_INITIALIZED before initialize and
 * FLIPMI_ERR_INVALID_ARG for dt < 0, as the reference throws.  A failed
 * pressure solve leaves the velocity unprojected (as the reference does) and
 * is reported in stats->pressure_ok, not as an error. stats may be NULL. */
int flipmi_scene_update(flipmi_scene *scene, double dt, flipmi_frame_stats *stats);
/* FluidSimulation::getNumMarkerParticles (0 before initialize) */
int flipmi_scene_num_particles(flipmi_scene *scene, int64_t *n);
/* FluidSimulation::getMarkerParticlePositionData / VelocityData
 * (fluidsimulation.cpp:2408-2420): vec3 float blobs, 12 bytes per particle,
 * positions in domain coordinates (position * scale + offset); either
 * pointer may be NULL. */
int flipmi_scene_get_particles(flipmi_scene *scene, float *pos_xyz, float *vel_xyz);
/* FluidSimulation::getMarkerParticlePositionDataRange / VelocityDataRange
 * (fluidsimulation.cpp:2314-2338): particles [start, end).  An invalid
 * range (start < 0, end > count, start > end) is FLIPMI_ERR_INVALID_ARG, as
 * the reference throws domain_error.  Either pointer may be NULL. */
int flipmi_scene_get_particles_range(flipmi_scene *scene, int64_t start, int64_t end, float *pos_xyz,
                                     float *vel_xyz);
/* FluidSimulation::setDomainOffset(double, double, double) and
 * setDomainScale(double) (fluidsimulation.cpp:307-330): particles are
 * loaded and read back in domain coordinates, p_domain = p * scale +
 * offset.  Takes effect for particles admitted by a later initialize and
 * for every read-back. */
int flipmi_scene_set_domain(flipmi_scene *scene, double offset_x, double offset_y, double offset_z,
                            double scale);
/* Static mesh obstacle: MeshObject(isize, jsize, ksize, dx) +
 * updateMeshStatic(mesh) added with FluidSimulation::addMeshObstacle
 * (meshobject.cpp:50-59, fluidsimulation.cpp:1994-2008).  Its level set
 * (exact band 3, meshobject.cpp:239-262, 519-553) is unioned into the solid
 * SDF (fluidsimulation.cpp:2932-2985), which is rebuilt at the next substep.
 * Grid coordinates.  The mesh must lie at least 4 cells inside the domain
 * grid (the reference splits meshes that leave its level set's grid,
 * meshutils.cpp:392-428), have fewer than 25 islands (the reference's
 * per-island path, meshobject.cpp:254-262) and no ray column may cross it
 * more than 64 times; otherwise FLIPMI_ERR_INVALID_ARG. */
int flipmi_scene_add_obstacle(flipmi_scene *scene, const float *vertices_xyz, int num_vertices,
                              const int *triangles, int num_triangles, int *obstacle_id);
/* MeshObject::enable()/disable() of an obstacle: the solid SDF is rebuilt at
 * the next substep (fluidsimulation.cpp:3028-3036). */
int flipmi_scene_set_obstacle(flipmi_scene *scene, int obstacle_id, int enabled);
/* MeshObject::inverse() state of an obstacle (meshobject.cpp:293-295): an
 * inversed obstacle is a container, solid outside its mesh
 * (fluidsimulation.cpp:2968-2984).  Like the reference, changing it does
 * not rebuild a solid SDF that already exists. */
int flipmi_scene_set_obstacle_inversed(flipmi_scene *scene, int obstacle_id, int inversed);
/* Mesh inflow source: MeshFluidSource(isize, jsize, ksize, dx) +
 * updateMeshStatic(mesh) + setInflow() + setVelocity(v)
 * (meshfluidsource.cpp:32-44, 74-76, 114-116), added with
 * FluidSimulation::addMeshFluidSource (fluidsimulation.cpp:1953-1965).
 * vertices_xyz: float triples in grid coordinates; triangles: int index
 * triples.  Each substep ends with the source's emission
 * (_updateInflowMeshFluidSources, fluidsimulation.cpp:4564-4697); its
 * particles and faces keep the source velocity (:3372-3438, 4109-4185).
 * The mesh must lie at least 4 cells inside the domain grid (the reference
 * splits meshes that leave the grid, meshutils.cpp:392-428) and no ray
 * column may cross it more than 64 times; otherwise FLIPMI_ERR_INVALID_ARG.
 * At most 8 sources.  *source_id receives the source's index. */
int flipmi_scene_add_inflow_source(flipmi_scene *scene, const float *vertices_xyz, int num_vertices,
                                   const int *triangles, int num_triangles, const double velocity[3],
                                   int *source_id);
/* MeshFluidSource::getMeshLevelSet / getMeshLevelSetOffset / getCells
 * (meshfluidsource.cpp:228-246) after initialize: dims receives the
 * sub-grid offset (cells) and node dims (6 ints); sdf (node dims, i
 * fastest) and cells (isize*jsize*ksize bytes, 1 = source cell) may be
 * NULL. */
int flipmi_scene_get_inflow_source(flipmi_scene *scene, int source_id, int dims[6], float *sdf,
                                   uint8_t *cells);
/* MeshFluidSource::enable()/disable(), enable/disableConstrainedFluidVelocity()
 * and setSubstepEmissions(n) (meshfluidsource.cpp:51-72, 146-156): n > 0
 * emits n times per substep, n = 0 only in the first substep of a frame;
 * negative n counts as 0. */
int flipmi_scene_set_inflow_source(flipmi_scene *scene, int source_id, int enabled,
                                   int constrain_velocity, int substep_emissions);
/* MeshFluidSource::setOutflow() / setInflow() and outflowInverse()
 * (meshfluidsource.cpp:74-84, 158-160): an outflow source adds nothing and
 * removes, at the end of every substep, the particles in its cells where
 * its SDF is negative (inversed: the particles outside its cells, or inside
 * them where the SDF is not negative), keeping the others in order
 * (fluidsimulation.cpp:4608-4673). */
int flipmi_scene_set_outflow_source(flipmi_scene *scene, int source_id, int outflow, int inversed);
/* FluidSimulation::setMarkerParticleJitterFactor (fluidsimulation.cpp:185-196,
 * factor in [0, 1]) and enable/disableJitterSurfaceMarkerParticles (:198-210)
 * for the particles inflow sources add.  The reference draws the jitter from
 * rand(); here it comes from a SplitMix64 stream keyed by (seed, emission
 * pass, candidate), so a run is reproducible. */
int flipmi_scene_set_marker_jitter(flipmi_scene *scene, double factor, int jitter_surface,
                                   uint64_t seed);
/* Copies an internal grid (FLIPMI_GRID_*) to host memory, e.g.
 * FluidSimulation::getVelocityField (fluidsimulation.cpp:2242). */
int flipmi_scene_get_grid(flipmi_scene *scene, int which, void *out);

/* Surface reconstruction settings: FluidSimulation's surface setters
 * (fluidsimulation.cpp:168-540, fields of fluidsimulation.h:1586-1611).
 * flipmi_surface_desc_default gives the reference's defaults, enabled.
 * Values the reference rejects (subdivisions < 1, slices < 1, iterations
 * < 0, particle scale < 0, polyhedron count < 0, distance < 0) are
 * FLIPMI_ERR_INVALID_ARG, as it throws
 * domain_error.  Not built here, also FLIPMI_ERR_INVALID_ARG: more than one
 * polygonizer slice, a minimum polyhedron triangle count above 0, and a
 * kernel search box wider than 40 half-cells per 10^3 node block (particle
 * scale above about 9 at subdivision level 1). */
typedef struct {
    int enabled;                      /* enable/disableSurfaceReconstruction, default 1 */
    int subdivisions;                 /* setSurfaceSubdivisionLevel, default 1 */
    int polygonizer_slices;           /* setNumPolygonizerSlices, default 1 */
    double smoothing_value;           /* setSurfaceSmoothingValue, default 0.5 */
    int smoothing_iterations;         /* setSurfaceSmoothingIterations, default 2 */
    int min_polyhedron_triangles;     /* setMinPolyhedronTriangleCount, default 0 */
    double particle_scale;            /* setMarkerParticleScale, default 3.0 */
    double obstacle_meshing_offset;   /* setObstacleMeshingOffset (cells), default 0 */
    int remove_near_domain;           /* enable/disableRemoveSurfaceNearDomain, default 0 */
    int remove_near_domain_distance;  /* setRemoveSurfaceNearDomainDistance, default 0 */
    int invert_contact_normals;       /* enable/disableInvertedContactNormals, default 0 */
} flipmi_surface_desc;
void flipmi_surface_desc_default(flipmi_surface_desc *desc);
/* A scene starts with surface reconstruction off.  When on, the mesh of the
 * particles is built at the end of the first substep of every frame, as the
 * reference's _outputSimulationData does (fluidsimulation.cpp:5445-5463). */
int flipmi_scene_set_surface(flipmi_scene *scene, const flipmi_surface_desc *desc);
/* FluidSimulation::getIsomesh (fluidsimulation.cpp:2543-2546): the last
 * mesh, vertices in domain coordinates (xyz floats) and triangles (three
 * int32 vertex indices each).  Either array pointer may be NULL to query
 * the counts only; an empty mesh before the first one is built. */
int flipmi_scene_get_isomesh(flipmi_scene *scene, int64_t *num_vertices, int64_t *num_triangles,
                             float *vertices_xyz, int32_t *triangles);

/* The surface of a particle set over a nodal solid SDF, as the simulation
 * builds it with `desc` (enabled is ignored): ParticleMesher::meshParticles
 * (particlemesher.cpp:36-55) followed by the post-processing of
 * _outputSurfaceMeshThread.  With smoothing_iterations = 0, no removal and
 * no inversion, and an identity domain transform it is meshParticles'
 * mesh.  kernel_radius: the particle kernel radius (ParticleMesherParameters
 * ::radius); 0 = the simulation's marker particle radius times
 * desc->particle_scale.  Positions in grid coordinates; domain_offset may be
 * NULL.  The result stays on the device until copied and destroyed. */
typedef struct flipmi_mesh flipmi_mesh;
int flipmi_mesh_particles(flipmi_ctx *ctx, int isize, int jsize, int ksize, double dx, const float *pos_xyz,
                          int64_t n, const float *solid_phi_nodal, const flipmi_surface_desc *desc,
                          double kernel_radius, const double *domain_offset, double domain_scale,
                          flipmi_mesh **out);
int flipmi_mesh_counts(flipmi_mesh *mesh, int64_t *num_vertices, int64_t *num_triangles);
int flipmi_mesh_copy(flipmi_mesh *mesh, float *vertices_xyz, int32_t *triangles);
int flipmi_mesh_destroy(flipmi_mesh *mesh);
int flipmi_scene_stage_times(flipmi_scene *scene, flipmi_stage_times *out, int reset);
/* hipEvent timing on the scene's stream: level 0 off, 1 one event pair per
 * stage, 2 additionally one pair around every PCG kernel launch. */
int flipmi_scene_set_profiling(flipmi_scene *scene, int level);
/* PCG kernels timed at profiling level 2, one bit per FLIPMI_STAGE_K_*
 * stage minus FLIPMI_STAGE_K_SPMV (bit 5: the true residual and, on the
 * mixed path, the forward sweep over it; the mixed path's forward sweep also
 * does its axpy, so its FLIPMI_STAGE_K_AXPY stays empty); default all.
 * Each timed launch adds an event pair, which delays the launches around it. */
int flipmi_scene_set_kernel_timing(flipmi_scene *scene, unsigned mask);
/* Times the selected kernels on every `every`-th PCG iteration only (default
 * 1), so a long run is timed at a fraction of the event pairs' delay.  No
 * counterpart in the reference: measurement only. */
int flipmi_scene_set_kernel_sampling(flipmi_scene *scene, int every);
/* Not in the reference: run the PCG's iteration batches as HIP graphs,
 * each (batch size, starting buffer) captured on first use and replayed
 * afterwards, so a batch of k iterations is one launch instead of 5 k.
 * Results are bit-identical either way.  Off by default; graphs are not
 * used while PCG kernels are timed (profiling level 2). */
int flipmi_scene_set_launch_graphs(flipmi_scene *scene, int enabled);
/* Blocks until all work queued by the context's scenes has finished. */
int flipmi_ctx_synchronize(flipmi_ctx *ctx);

/* ---- stage drop-ins on host buffers ------------------------------------ */
/* Domain-wall solid SDF (nodal) and the weight grid: FluidSimulation::
 * _addStaticObjectsToSDF + _updateWeightGrid (fluidsimulation.cpp:2918-2930,
 * 3646-3730).  Weight pointers may be NULL. */
int flipmi_domain_solid(flipmi_ctx *ctx, int isize, int jsize, int ksize, double dx,
                        float *phi_nodal, float *wU, float *wV, float *wW, float *wC);

/* ParticleLevelSet::calculateSignedDistanceField + postProcessSignedDistanceField
 * (particlelevelset.cpp:161-196).  solid_phi_nodal may be NULL (no post-process). */
int flipmi_liquid_sdf(flipmi_ctx *ctx, int isize, int jsize, int ksize, double dx,
                      const float *pos_xyz, int64_t n, double radius,
                      const float *solid_phi_nodal, float *phi_out);

/* The counting loop of FluidSimulation::_removeMarkerParticles
 * (fluidsimulation.cpp:4325-4343) on given particle positions: particles in
 * order, those flagged in `removed` (the solid test, may be NULL) skipped,
 * each cell counting up to 250 (fluidsimulation.h:1684).  counts_out is an
 * int32 I x J x K grid; particles outside the grid count nowhere. */
int flipmi_removal_counts(flipmi_ctx *ctx, int isize, int jsize, int ksize, double dx,
                          const float *pos_xyz, int64_t n, const uint8_t *removed, int32_t *counts_out);

/* ParticleLevelSet::postProcessSignedDistanceField(MeshLevelSet&)
 * (particlelevelset.cpp:172-196) on a field from flipmi_liquid_sdf computed
 * without a solid, in place. */
int flipmi_liquid_sdf_postprocess(flipmi_ctx *ctx, int isize, int jsize, int ksize, double dx,
                                  const float *solid_phi_nodal, float *phi_inout);

/* VelocityAdvector::advect(VelocityAdvectorParameters) (velocityadvector.h:75)
 * onto a cleared MAC grid: velocities and valid flags of U, V and W. */
int flipmi_p2g(flipmi_ctx *ctx, int isize, int jsize, int ksize, double dx, const float *pos_xyz,
               const float *vel_xyz, int64_t n, double radius, float *u, float *v, float *w,
               uint8_t *valid_u, uint8_t *valid_v, uint8_t *valid_w);

/* MACVelocityField::extrapolateVelocityField / GridUtils::extrapolateGrid
 * (gridutils.cpp:32-228) on one component grid, in place. */
int flipmi_extrapolate(flipmi_ctx *ctx, int ni, int nj, int nk, float *grid,
                       const uint8_t *valid, int layers);

/* Size of the per-device scratch array behind flipmi_extrapolate_noalloc
 * (one status byte per face): 64 MiB, held by every device the library
 * runs on until the process ends. */
#define FLIPMI_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)

/* No counterpart in the reference.  The same result as flipmi_extrapolate
 * (GridUtils::extrapolateGrid, gridutils.cpp:32-228), for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops.
 * It calls no allocator.  grid_dev and valid_dev are DEVICE pointers owned by
 * the caller (ni*nj*nk floats / bytes); the BFS status grid lives in a fixed
 * per-device array of FLIPMI_NOALLOC_SCRATCH_BYTES, so ni*nj*nk above that
 * is rejected with FLIPMI_ERR_INVALID_ARG.  The work is enqueued on `stream`
 * (a hipStream_t; NULL = the context's stream) and not waited for.  All calls
 * on one device share that device's array: calls that overlap on a device
 * race. */
int flipmi_extrapolate_noalloc(flipmi_ctx *ctx, int ni, int nj, int nk, float *grid_dev,
                               const uint8_t *valid_dev, int layers, void *stream);

/* PressureSolverParameters (pressuresolver.h:63-79). */
typedef struct {
    int isize, jsize, ksize;
    double dx, dt;
    double tolerance;          /* relative to |rhs|_inf */
    double acceptable_tolerance; /* accept a capped solve below this |r|_inf (1.0) */
    int max_iterations;
    int precision;             /* flipmi_pcg_precision */
    const float *liquid_phi;   /* I x J x K */
    const float *weight_u, *weight_v, *weight_w, *weight_c;
    /* Solid face velocities (MeshLevelSet::getFaceVelocityU/V/W), NULL = 0.
     * In/out: zeroed around isolated fluid pockets, as the reference's
     * _conditionSolidVelocityField does (pressuresolver.cpp:124-268). */
    float *solid_u, *solid_v, *solid_w;
    float *u, *v, *w;          /* MAC velocity, updated in place */
    uint8_t *valid_u, *valid_v, *valid_w;      /* updated in place */
    double *pressure_rows;     /* optional out: solution in reference row order */
} flipmi_pressure_args;

typedef struct {
    int iterations;            /* "Pressure Solver Iterations" */
    double residual;           /* "Estimated Error" (|r|_inf) */
    int rows;                  /* fluid interior cells */
    int success;               /* the bool PressureSolver::solve returns */
    int refinements;           /* mixed precision: outer refinement passes */
} flipmi_pressure_report;

/* bool PressureSolver::solve(PressureSolverParameters) + getSolverStatus()
 * (pressuresolver.h:91-92).  Returns FLIPMI_OK when the solve succeeded (or
 * was accepted), FLIPMI_ERR_NOT_CONVERGED otherwise (velocity unchanged). */
int flipmi_pressure_solve(flipmi_ctx *ctx, const flipmi_pressure_args *args,
                          flipmi_pressure_report *report);

/* One application of the MIC(0) preconditioner z = M^-1 r on the system
 * assembled from (liquid_phi, weights), vectors in reference row order
 * (pcgsolver.h:69-221).  invdiag_out may be NULL. */
int flipmi_mic0_apply(flipmi_ctx *ctx, int isize, int jsize, int ksize, double dx, double dt,
                      const float *liquid_phi, const float *wU, const float *wV,
                      const float *wW, const float *wC, const double *r_rows, double *z_rows,
                      double *invdiag_rows);

/* ParticleAdvector::advectParticlesRK4/RK3/RK2/RK1 (particleadvector.h:53-71):
 * rk_order in {1, 2, 3, 4}. */
int flipmi_advect_particles(flipmi_ctx *ctx, int isize, int jsize, int ksize, double dx,
                            const float *u, const float *v, const float *w,
                            const float *pos_in, int64_t n, double dt, int rk_order,
                            float *pos_out);

/* ParticleAdvector::trilinearInterpolate (particleadvector.h:73-79). */
int flipmi_trilinear_velocity(flipmi_ctx *ctx, int isize, int jsize, int ksize, double dx,
                              const float *u, const float *v, const float *w,
                              const float *pos, int64_t n, float *vel_out);

/* ---- serving ------------------------------------------------------------
 * Not in the reference: a device-resident request loop over independent
 * scenes.  Requests are admitted in arrival order whenever one of max_active
 * worker slots (a host thread and a HIP stream each) is free and the scene's
 * device memory fits memory_budget, so scenes of mixed sizes run
 * concurrently and new arrivals join while others are on the device.
 * Finished scenes stay resident and are re-armed for the next request of the
 * same grid and settings; idle ones are evicted (least recently used) when a
 * new kind needs the memory.  A request's result does not depend on what else
 * runs beside it. */
typedef struct flipmi_server flipmi_server;

typedef struct {
    int max_active;                    /* scenes stepped at once, 1..64 (default 8) */
    int64_t memory_budget;             /* device bytes for resident scenes (default 200 GiB) */
    double particle_capacity_per_cell; /* particle capacity of a new scene per cell, so later
                                          requests of the same kind fit (default 3) */
    int launch_graphs;                 /* the scenes' PCG iteration batches run as captured HIP
                                          graphs (flipmi_scene_set_launch_graphs; default 1) */
} flipmi_server_desc;

typedef struct {
    flipmi_scene_desc scene;           /* grid and settings, as for flipmi_scene_create */
    const float *pos_xyz;              /* n particles, copied at submit */
    const float *vel_xyz;              /* may be NULL (zero velocity) */
    int64_t n;
    double dt;                         /* FluidSimulation::update(dt) per step */
    int steps;                         /* number of update calls */
} flipmi_request;

typedef struct {
    int status;                        /* FLIPMI_OK or the error that ended the request */
    int steps_done;
    int64_t num_particles;             /* at the end */
    int64_t particle_substeps;         /* sum over substeps of the particles stepped */
    int pressure_failures;             /* substeps whose pressure solve failed */
    int reused_scene;                  /* 1 if a resident scene was re-armed */
    double queue_ms, run_ms;           /* submit -> start, start -> done (host clock) */
    double step_ms_p50, step_ms_p99;   /* latency of its update calls */
} flipmi_request_report;

typedef struct {
    int64_t submitted, completed, failed, queued, active;
    int64_t particle_substeps;         /* completed requests */
    int64_t steps;                     /* update calls of completed requests */
    double step_ms_p50, step_ms_p99;   /* over those update calls */
    int64_t resident_bytes, peak_resident_bytes;
    int64_t scenes_created, scenes_reused, scenes_evicted;
} flipmi_server_stats;

void flipmi_server_desc_default(flipmi_server_desc *desc);
int flipmi_server_create(flipmi_ctx *ctx, const flipmi_server_desc *desc, flipmi_server **out);
/* Queues a request; *request_id identifies it in wait/take. */
int flipmi_server_submit(flipmi_server *server, const flipmi_request *request, int64_t *request_id);
/* Blocks until the request is done (timeout_s < 0: no limit; else
 * FLIPMI_ERR_TIMEOUT when it is not done in time); returns its status. */
int flipmi_server_wait(flipmi_server *server, int64_t request_id, double timeout_s,
                       flipmi_request_report *report);
/* Copies the final particles of a finished request (report.num_particles
 * triples each; either pointer may be NULL) and releases it. */
int flipmi_server_take(flipmi_server *server, int64_t request_id, float *pos_xyz, float *vel_xyz);
int flipmi_server_get_stats(flipmi_server *server, flipmi_server_stats *stats);
/* Runs the queued and running requests to the end, then frees everything. */
int flipmi_server_destroy(flipmi_server *server);

#ifdef __cplusplus
}
#endif

#endif /* FLIPMI_H */
