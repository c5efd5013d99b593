// CPU oracle for the FLIP step -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// A single-threaded C++17 restatement of the reference FLIP engine step
// (jklae/FLIPEngine3D src/engine, FluidSimulation::update and the stages it
// calls).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it, and only as a checker or as the timed CPU baseline.  The
// shipped library (flipengine3d_amd/) never links or calls it.
//
// Every function in flip_oracle.cpp cites the reference file:line it follows.
// Where the reference's floating-point result does not depend on evaluation
// order (minimum reductions, per-element arithmetic, sequential loops) the
// oracle reproduces it bit for bit; it is pinned against the compiled
// reference (oracle/_ref) by tests/test_oracle_vs_ref.py.
#pragma once
#include <cstdint>

#ifdef __cplusplus
extern "C" {
#endif

// Pressure-solver modes for orc_pressure_solve / orc_sim_set_pcg_mode.
//   0: reference-exact -- alpha/beta rounded to float before each axpy and
//      sequential dot products (pcgsolver.h:248-302, blaswrapper.h:35-158).
//   1: pure FP64 -- alpha/beta kept in double, sequential dots, stop on the
//      recursive residual.
//   2: converged FP64 error oracle -- alpha/beta in double, compensated dot
//      products, and the solve ends only when the TRUE residual b - Ax is
//      below the tolerance (run at 1e-12 to measure other solvers' error).
enum { ORC_PCG_REF_EXACT = 0, ORC_PCG_FP64 = 1, ORC_PCG_FP64_TRUE = 2 };

void orc_domain_solid(int isize, int jsize, int ksize, double dx, float *phi_nodal, float *wU,
                      float *wV, float *wW, float *wC);

void orc_liquid_sdf(int isize, int jsize, int ksize, double dx, const float *pos, int n,
                    double radius, const float *solid_phi_nodal, float *phi_out);

void orc_p2g(int isize, int jsize, int ksize, double dx, const float *pos, const float *vel, int n,
             double radius, float *u, float *v, float *w, uint8_t *vu, uint8_t *vv, uint8_t *vw);

void orc_extrapolate(int ni, int nj, int nk, float *grid, uint8_t *valid, int layers);

int orc_pressure_solve(int isize, int jsize, int ksize, double dx, double dt, double tol,
                       int maxiter, int mode, const float *liquid_phi,
                       const float *solid_phi_nodal, const float *wU, const float *wV,
                       const float *wW, const float *wC, float *u, float *v, float *w,
                       uint8_t *vu, uint8_t *vv, uint8_t *vw, float *su, float *sv, float *sw,
                       double *rhs_out, double *soln_out, int *n_rows, int *iters, double *err);

void orc_mic0_apply(int isize, int jsize, int ksize, double dx, double dt, const float *liquid_phi,
                    const float *wU, const float *wV, const float *wW, const float *wC,
                    const double *r_in, double *invdiag_out, double *z_out);

void orc_advect_particles(int isize, int jsize, int ksize, double dx, const float *u,
                          const float *v, const float *w, const float *pos, int n, double dt,
                          int order, float *out);

void orc_trilinear_velocity(int isize, int jsize, int ksize, double dx, const float *u,
                            const float *v, const float *w, const float *pos, int n, float *out);

// Whole-step driver (FluidSimulation::update, fluidsimulation.cpp:5755-5846).
void *orc_sim_create(int isize, int jsize, int ksize, double dx);
void orc_sim_destroy(void *h);
void orc_sim_add_body_force(void *h, double fx, double fy, double fz);
void orc_sim_set_cfl(void *h, double cfl);
void orc_sim_set_substeps(void *h, int minsteps, int maxsteps);
void orc_sim_set_pic_ratio(void *h, double r);
void orc_sim_set_pcg(void *h, int mode, double tol, int maxiter);
void orc_sim_load_particles(void *h, const float *pos, const float *vel, int n);
int orc_sim_initialize(void *h);
int orc_sim_update(void *h, double dt);
int orc_sim_num_particles(void *h);
void orc_sim_get_particles(void *h, float *pos, float *vel);
void orc_sim_get_float_grid(void *h, int which, float *out);
void orc_sim_get_bool_grid(void *h, int which, uint8_t *out);
int orc_sim_frame_substeps(void *h);
void orc_sim_get_count_grid(void *h, int32_t *out, int *speed_last, int *speed_total);
void orc_removal_counts(int I, int J, int K, double dx, const float *pos, int n, const uint8_t *removed,
                        int32_t *out);
void orc_sim_pressure_info(void *h, int *iters, double *err, int *ok);
// Mesh inside tests, mesh level sets and inflow sources
// (meshutils.cpp:99-428, meshlevelset.cpp:572-601, meshfluidsource.cpp,
// fluidsimulation.cpp:3372-3438, 4109-4185, 4397-4806).
int orc_mesh_nodes_inside(const float *verts, int nv, const int *tris, int nt, double dx, int nni,
                          int nnj, int nnk, const float *jitter, uint8_t *out);
int orc_mesh_sdf(const float *verts, int nv, const int *tris, int nt, int isize, int jsize,
                 int ksize, double dx, int band, const float *jitter, float *phi);
int orc_sim_add_inflow(void *h, const float *verts, int nv, const int *tris, int nt, double vx,
                       double vy, double vz);
void orc_sim_set_inflow(void *h, int id, int enabled, int constrain, int emissions);
void orc_sim_set_outflow(void *h, int id, int outflow, int inversed);
void orc_sim_set_jitter(void *h, double factor, int surface, uint64_t seed);
int orc_sim_inflow_sdf(void *h, int id, int *dims, float *phi);
int orc_sim_inflow_cells(void *h, int id, uint8_t *cells);
void orc_ray_jitter(uint64_t seed, int id, int which, double dx, float *out);
int orc_sim_add_obstacle(void *h, const float *verts, int nv, const int *tris, int nt,
                         const float *jitter);
void orc_sim_set_obstacle(void *h, int id, int enabled);
void orc_sim_set_obstacle_inversed(void *h, int id, int inversed);
int orc_inflow_build(const float *verts, int nv, const int *tris, int nt, int I, int J, int K,
                     double dx, uint64_t seed, int id, int *dims, float *phi, uint8_t *cells);

// Surface reconstruction (FluidSimulation::_outputSurfaceMeshThread,
// fluidsimulation.cpp:5134-5243; ParticleMesher, Polygonizer3d,
// TriangleMesh::smooth).  One polygonizer slice, no meshing volume, no
// preview or motion blur, minimum polyhedron size 0.
void orc_sim_set_surface(void *h, int enabled, int subdivisions, double smooth_value, int smooth_iterations,
                         double particle_scale, double obstacle_offset, int remove_near_domain,
                         int remove_distance, int invert_contact, double contact_threshold);
void orc_sim_isomesh(void *h, int *nverts, int *ntris, float *verts, int *tris);
void *orc_surface_mesh(int I, int J, int K, double dx, const float *pos, int n, const float *solid_phi,
                       int subdivisions, double smooth_value, int smooth_iterations, double particle_scale,
                       double obstacle_offset, int remove_near_domain, int remove_distance, int invert_contact,
                       double contact_threshold, const float *domain_offset, double domain_scale);
void orc_mesh_get(void *h, int *nverts, int *ntris, float *verts, int *tris);
void orc_mesh_free(void *h);

#ifdef __cplusplus
}
#endif
