// Headless harness around the reference FLIP engine (jklae/FLIPEngine3D,
// src/engine).  It is compiled together with the reference's own engine
// sources, unmodified and in place, by oracle/Makefile into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
// calls this file; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may use oracle/_ref as a checker or as the timed CPU
// baseline.
//
// The harness exposes a small extern "C" surface so Python (ctypes) can
//   * drive FluidSimulation::update(dt) end to end
//     (src/engine/fluidsimulation.cpp:5755-5846),
//   * read the engine's internal grids after a step (liquid SDF, solid SDF,
//     MAC velocity, valid flags, weight grid) for parity checks, and
//   * call single stages on caller-supplied inputs: the liquid SDF
//     (particlelevelset.cpp:161-196), P2G (velocityadvector.cpp:38-459),
//     velocity extrapolation (gridutils.cpp:32-228), the pressure projection
//     (pressuresolver.cpp:44-1047) and ParticleAdvector::advectParticlesRK*
//     (particleadvector.cpp:137-336).
//
// Private members are reached by re-declaring `private`/`protected` as
// `public` for this translation unit only, after every standard header has
// been included.  Class layouts are unchanged by this.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <queue>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#define private public
#define protected public
#include "fluidsimulation.h"
#include "pressuresolver.h"
#include "velocityadvector.h"
#include "particleadvector.h"
#include "particlelevelset.h"
#include "meshlevelset.h"
#include "meshfluidsource.h"
#include "meshutils.h"
#include "trianglemesh.h"
#include "macvelocityfield.h"
#include "gridutils.h"
#include "threadutils.h"
#include "pcgsolver/pcgsolver.h"
#undef private
#undef protected

namespace {

struct RefSim {
    FluidSimulation *sim = nullptr;
    std::vector<MeshFluidSource *> sources;
    std::vector<MeshObject *> obstacles;
    std::string last_error;
};

TriangleMesh make_tri_mesh(const float *verts, int nv, const int *tris, int nt) {
    TriangleMesh m;
    for (int i = 0; i < nv; i++)
        m.vertices.push_back(vmath::vec3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]));
    for (int i = 0; i < nt; i++) {
        Triangle t;
        t.tri[0] = tris[3 * i]; t.tri[1] = tris[3 * i + 1]; t.tri[2] = tris[3 * i + 2];
        m.triangles.push_back(t);
    }
    return m;
}

// The ray jitter MeshUtils::_getCollisionGridZ draws right after srand(seed)
// (meshutils.cpp:107-110), written as the same expression so the compiler
// assigns the three draws to the axes in the same order.
vmath::vec3 predicted_ray_jitter(unsigned seed, double dx) {
    srand(seed);
    double jit = 0.001 * dx;
    vmath::vec3 jitter(MeshUtils::_randomDouble(jit, -jit), MeshUtils::_randomDouble(jit, -jit),
                       MeshUtils::_randomDouble(jit, -jit));
    return jitter;
}

template <class T>
void copy_array(Array3d<T> &a, T *out) {
    int n = a.width * a.height * a.depth;
    const T *raw = a.getRawArray();
    std::memcpy(out, raw, sizeof(T) * (size_t)n);
}

void copy_bool_array(Array3d<bool> &a, uint8_t *out) {
    int n = a.width * a.height * a.depth;
    bool *raw = a.getRawArray();
    for (int i = 0; i < n; i++) out[i] = raw[i] ? 1 : 0;
}

template <class T>
void fill_array(Array3d<T> &a, const T *in) {
    int n = a.width * a.height * a.depth;
    T *raw = a.getRawArray();
    std::memcpy(raw, in, sizeof(T) * (size_t)n);
}

void fill_bool_array(Array3d<bool> &a, const uint8_t *in) {
    int n = a.width * a.height * a.depth;
    bool *raw = a.getRawArray();
    for (int i = 0; i < n; i++) raw[i] = in[i] != 0;
}

// Builds the domain-boundary solid SDF exactly as FluidSimulation does for a
// scene with no obstacles (fluidsimulation.cpp:2918-2930 via
// _addStaticObjectsToSDF).  The MeshObject must outlive the level set.
void build_domain_solid(FluidSimulation &sim) {
    sim._solidSDF.reset();
    sim._addStaticObjectsToSDF(0.0, sim._solidSDF);
    sim._solidSDF.normalizeVelocityGrid();
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Whole-engine driver
// ---------------------------------------------------------------------------

void *ref_sim_create(int isize, int jsize, int ksize, double dx, int nthreads) {
    ThreadUtils::setMaxThreadCount(nthreads > 0 ? nthreads : 1);
    RefSim *h = new RefSim();
    h->sim = new FluidSimulation(isize, jsize, ksize, dx);
    h->sim->disableSurfaceReconstruction();
    h->sim->disableConsoleOutput();
    return h;
}

void ref_sim_destroy(void *hp) {
    RefSim *h = (RefSim *)hp;
    if (!h) return;
    delete h->sim;
    for (MeshFluidSource *src : h->sources) delete src;
    for (MeshObject *ob : h->obstacles) delete ob;
    delete h;
}

void ref_set_thread_count(int n) { ThreadUtils::setMaxThreadCount(n > 0 ? n : 1); }

// Surface reconstruction settings through the engine's own setters
// (fluidsimulation.cpp:168-540), meshing synchronously.
void ref_sim_set_surface(void *hp, int enabled, int subdivisions, double smooth_value, int smooth_iterations,
                         double particle_scale, double obstacle_offset, int remove_near_domain,
                         int remove_distance, int invert_contact) {
    FluidSimulation *s = ((RefSim *)hp)->sim;
    if (enabled) s->enableSurfaceReconstruction(); else s->disableSurfaceReconstruction();
    s->disableAsynchronousMeshing();
    s->setSurfaceSubdivisionLevel(subdivisions);
    s->setSurfaceSmoothingValue(smooth_value);
    s->setSurfaceSmoothingIterations(smooth_iterations);
    s->setMarkerParticleScale(particle_scale);
    s->setObstacleMeshingOffset(obstacle_offset);
    if (remove_near_domain) s->enableRemoveSurfaceNearDomain(); else s->disableRemoveSurfaceNearDomain();
    s->setRemoveSurfaceNearDomainDistance(remove_distance);
    if (invert_contact) s->enableInvertedContactNormals(); else s->disableInvertedContactNormals();
}

// FluidSimulation::getIsomesh (fluidsimulation.cpp:2543-2546).
void ref_sim_isomesh(void *hp, int *nverts, int *ntris, float *verts, int *tris) {
    TriangleMesh &m = ((RefSim *)hp)->sim->getIsomesh();
    *nverts = (int)m.vertices.size();
    *ntris = (int)m.triangles.size();
    if (verts)
        for (size_t v = 0; v < m.vertices.size(); v++) {
            verts[3 * v] = m.vertices[v].x;
            verts[3 * v + 1] = m.vertices[v].y;
            verts[3 * v + 2] = m.vertices[v].z;
        }
    if (tris)
        for (size_t t = 0; t < m.triangles.size(); t++)
            for (int c = 0; c < 3; c++) tris[3 * t + c] = m.triangles[t].tri[c];
}

void ref_sim_add_body_force(void *hp, double fx, double fy, double fz) {
    ((RefSim *)hp)->sim->addBodyForce(fx, fy, fz);
}

void ref_sim_set_cfl(void *hp, double cfl) { ((RefSim *)hp)->sim->setCFLConditionNumber(cfl); }

void ref_sim_set_substeps(void *hp, int minsteps, int maxsteps) {
    ((RefSim *)hp)->sim->setMinTimeStepsPerFrame(minsteps);
    ((RefSim *)hp)->sim->setMaxTimeStepsPerFrame(maxsteps);
}

void ref_sim_set_pic_ratio(void *hp, double r) { ((RefSim *)hp)->sim->setPICFLIPRatio(r); }

void ref_sim_load_particles(void *hp, const float *pos, const float *vel, int n) {
    FluidSimulationMarkerParticleData d;
    d.size = n;
    d.positions = (char *)pos;
    d.velocities = (char *)vel;
    ((RefSim *)hp)->sim->loadMarkerParticleData(d);
}

void ref_sim_set_domain(void *hp, double ox, double oy, double oz, double scale) {
    FluidSimulation *s = ((RefSim *)hp)->sim;
    s->setDomainOffset(ox, oy, oz);
    s->setDomainScale(scale);
}

int ref_sim_get_particle_range(void *hp, int start, int end, float *pos, float *vel) {
    RefSim *h = (RefSim *)hp;
    try {
        h->sim->getMarkerParticlePositionDataRange(start, end, (char *)pos);
        h->sim->getMarkerParticleVelocityDataRange(start, end, (char *)vel);
    } catch (std::exception &e) {
        h->last_error = e.what();
        return -1;
    }
    return 0;
}

int ref_sim_initialize(void *hp) {
    RefSim *h = (RefSim *)hp;
    try {
        h->sim->initialize();
    } catch (std::exception &e) {
        h->last_error = e.what();
        return -1;
    }
    return 0;
}

int ref_sim_update(void *hp, double dt) {
    RefSim *h = (RefSim *)hp;
    try {
        h->sim->update(dt);
    } catch (std::exception &e) {
        h->last_error = e.what();
        return -1;
    }
    return 0;
}

int ref_sim_num_particles(void *hp) { return (int)((RefSim *)hp)->sim->_markerParticles.size(); }

void ref_sim_get_particles(void *hp, float *pos, float *vel) {
    FluidSimulation *s = ((RefSim *)hp)->sim;
    for (size_t i = 0; i < s->_markerParticles.size(); i++) {
        MarkerParticle mp = s->_markerParticles[i];
        pos[3 * i + 0] = mp.position.x;
        pos[3 * i + 1] = mp.position.y;
        pos[3 * i + 2] = mp.position.z;
        vel[3 * i + 0] = mp.velocity.x;
        vel[3 * i + 1] = mp.velocity.y;
        vel[3 * i + 2] = mp.velocity.z;
    }
}

// which: 0 liquid phi (I*J*K), 1 solid phi nodal ((I+1)(J+1)(K+1)),
// 2/3/4 MAC U/V/W, 5/6/7 weight U/V/W, 8 weight center,
// 9/10/11 solid face velocity U/V/W
void ref_sim_get_float_grid(void *hp, int which, float *out) {
    FluidSimulation *s = ((RefSim *)hp)->sim;
    switch (which) {
        case 0: copy_array(s->_liquidSDF._phi, out); break;
        case 1: copy_array(s->_solidSDF._phi, out); break;
        case 2: copy_array(s->_MACVelocity._u, out); break;
        case 3: copy_array(s->_MACVelocity._v, out); break;
        case 4: copy_array(s->_MACVelocity._w, out); break;
        case 5: copy_array(s->_weightGrid.U, out); break;
        case 6: copy_array(s->_weightGrid.V, out); break;
        case 7: copy_array(s->_weightGrid.W, out); break;
        case 8: copy_array(s->_weightGrid.center, out); break;
        case 9: copy_array(s->_solidSDF._velocityData.field._u, out); break;
        case 10: copy_array(s->_solidSDF._velocityData.field._v, out); break;
        case 11: copy_array(s->_solidSDF._velocityData.field._w, out); break;
        default: break;
    }
}

// which: 0/1/2 valid U/V/W; 3 near-solid grid (coarse)
void ref_sim_get_bool_grid(void *hp, int which, uint8_t *out) {
    FluidSimulation *s = ((RefSim *)hp)->sim;
    switch (which) {
        case 0: copy_bool_array(s->_validVelocities.validU, out); break;
        case 1: copy_bool_array(s->_validVelocities.validV, out); break;
        case 2: copy_bool_array(s->_validVelocities.validW, out); break;
        case 3: copy_bool_array(s->_nearSolidGrid, out); break;
        default: break;
    }
}

void ref_sim_near_solid_dims(void *hp, int *dims) {
    FluidSimulation *s = ((RefSim *)hp)->sim;
    dims[0] = s->_nearSolidGrid.width;
    dims[1] = s->_nearSolidGrid.height;
    dims[2] = s->_nearSolidGrid.depth;
}

// Frame statistics of the last update(): substeps, particle count and the
// pressure-solver status text (pressuresolver.cpp:821-837).
int ref_sim_frame_substeps(void *hp) { return ((RefSim *)hp)->sim->getFrameStatsData().substeps; }

void ref_sim_pressure_status(void *hp, char *buf, int buflen) {
    std::string st = ((RefSim *)hp)->sim->_pressureSolverStatus;
    std::snprintf(buf, (size_t)buflen, "%s", st.c_str());
}

void ref_sim_last_error(void *hp, char *buf, int buflen) {
    std::snprintf(buf, (size_t)buflen, "%s", ((RefSim *)hp)->last_error.c_str());
}

// ---------------------------------------------------------------------------
// Stage-level entry points on caller-supplied inputs
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Mesh inside tests, mesh level sets and inflow sources
// ---------------------------------------------------------------------------

// MeshUtils::getGridNodesInsideTriangleMesh(mesh, dx, Array3d<bool>)
// (meshutils.cpp:374-428) with rand() seeded by `seed`; the ray jitter that
// seed produces is written to jitter_out.
void ref_mesh_nodes_inside(const float *verts, int nv, const int *tris, int nt, double dx,
                           int nni, int nnj, int nnk, unsigned seed, float *jitter_out,
                           uint8_t *out) {
    TriangleMesh m = make_tri_mesh(verts, nv, tris, nt);
    vmath::vec3 j = predicted_ray_jitter(seed, dx);
    jitter_out[0] = j.x; jitter_out[1] = j.y; jitter_out[2] = j.z;
    Array3d<bool> nodes(nni, nnj, nnk, false);
    srand(seed);
    MeshUtils::getGridNodesInsideTriangleMesh(m, dx, nodes);
    copy_bool_array(nodes, out);
}

// MeshLevelSet(isize, jsize, ksize, dx).fastCalculateSignedDistanceField(mesh,
// band) (meshlevelset.cpp:572-601) without velocity data, rand() seeded by
// `seed`.
void ref_mesh_sdf(const float *verts, int nv, const int *tris, int nt, int isize, int jsize,
                  int ksize, double dx, int band, unsigned seed, float *jitter_out, float *phi) {
    TriangleMesh m = make_tri_mesh(verts, nv, tris, nt);
    vmath::vec3 j = predicted_ray_jitter(seed, dx);
    jitter_out[0] = j.x; jitter_out[1] = j.y; jitter_out[2] = j.z;
    MeshLevelSet ls(isize, jsize, ksize, dx);
    ls.disableVelocityData();
    srand(seed);
    ls.fastCalculateSignedDistanceField(m, band);
    copy_array(ls._phi, phi);
}

// MeshObject(isize, jsize, ksize, dx) + updateMeshStatic(mesh), added with
// FluidSimulation::addMeshObstacle (meshobject.cpp:50-59, fluidsimulation.cpp:
// 1994-2008).
int ref_sim_add_obstacle(void *hp, const float *verts, int nv, const int *tris, int nt) {
    RefSim *h = (RefSim *)hp;
    FluidSimulation *sim = h->sim;
    MeshObject *ob = new MeshObject(sim->_isize, sim->_jsize, sim->_ksize, sim->_dx);
    ob->updateMeshStatic(make_tri_mesh(verts, nv, tris, nt));
    sim->addMeshObstacle(ob);
    h->obstacles.push_back(ob);
    return (int)h->obstacles.size() - 1;
}

void ref_sim_set_obstacle_inversed(void *hp, int id, int inversed) {
    MeshObject *ob = ((RefSim *)hp)->obstacles[id];
    if ((inversed != 0) != ob->isInversed()) ob->inverse();
}

void ref_sim_set_obstacle(void *hp, int id, int enabled) {
    MeshObject *ob = ((RefSim *)hp)->obstacles[id];
    if (enabled) ob->enable(); else ob->disable();
}

// srand(seed), then the ray jitters of the next `count` inside tests, as
// MeshUtils::_getCollisionGridZ draws them one after another.  Call
// ref_seed_rand(seed) again before the run that uses them.
void ref_predict_ray_jitters(unsigned seed, double dx, int count, float *out) {
    srand(seed);
    double jit = 0.001 * dx;
    for (int q = 0; q < count; q++) {
        vmath::vec3 jitter(MeshUtils::_randomDouble(jit, -jit), MeshUtils::_randomDouble(jit, -jit),
                           MeshUtils::_randomDouble(jit, -jit));
        out[3 * q] = jitter.x; out[3 * q + 1] = jitter.y; out[3 * q + 2] = jitter.z;
    }
}

void ref_seed_rand(unsigned seed) { srand(seed); }

// MeshFluidSource(isize, jsize, ksize, dx) + updateMeshStatic + setInflow +
// setVelocity, then FluidSimulation::addMeshFluidSource
// (meshfluidsource.cpp:32-44, 74-76, 114-116; fluidsimulation.cpp:1953-1965).
int ref_sim_add_inflow(void *hp, const float *verts, int nv, const int *tris, int nt, double vx,
                       double vy, double vz) {
    RefSim *h = (RefSim *)hp;
    FluidSimulation *sim = h->sim;
    MeshFluidSource *src = new MeshFluidSource(sim->_isize, sim->_jsize, sim->_ksize, sim->_dx);
    src->updateMeshStatic(make_tri_mesh(verts, nv, tris, nt));
    src->setInflow();
    src->setVelocity(vmath::vec3(vx, vy, vz));
    sim->addMeshFluidSource(src);
    h->sources.push_back(src);
    return (int)h->sources.size() - 1;
}

void ref_sim_set_inflow(void *hp, int id, int enabled, int constrain, int emissions) {
    MeshFluidSource *src = ((RefSim *)hp)->sources[id];
    if (enabled) src->enable(); else src->disable();
    if (constrain) src->enableConstrainedFluidVelocity(); else src->disableConstrainedFluidVelocity();
    src->setSubstepEmissions(emissions);
}

void ref_sim_set_outflow(void *hp, int id, int outflow, int inversed) {
    MeshFluidSource *src = ((RefSim *)hp)->sources[id];
    if (outflow) src->setOutflow(); else src->setInflow();
    if (inversed != (src->isOutflowInversed() ? 1 : 0)) src->outflowInverse();
}

void ref_sim_set_jitter(void *hp, double factor, int surface) {
    FluidSimulation *sim = ((RefSim *)hp)->sim;
    sim->setMarkerParticleJitterFactor(factor);
    if (surface) sim->enableJitterSurfaceMarkerParticles();
    else sim->disableJitterSurfaceMarkerParticles();
}

// The source's SDF after its first update: grid offset and node dims in
// dims[0..5], phi nodal.
int ref_sim_inflow_sdf(void *hp, int id, int *dims, float *phi) {
    MeshFluidSource *src = ((RefSim *)hp)->sources[id];
    MeshLevelSet *ls = src->getMeshLevelSet();
    dims[0] = src->_sourceSDFGridOffset.i; dims[1] = src->_sourceSDFGridOffset.j;
    dims[2] = src->_sourceSDFGridOffset.k;
    dims[3] = ls->_phi.width; dims[4] = ls->_phi.height; dims[5] = ls->_phi.depth;
    if (phi) copy_array(ls->_phi, phi);
    return 0;
}

// MeshFluidSource::getCells (meshfluidsource.cpp:228-234) with rand() seeded
// by `seed`, as a flat I*J*K byte grid.
void ref_sim_inflow_cells(void *hp, int id, unsigned seed, float *jitter_out, uint8_t *cells) {
    RefSim *h = (RefSim *)hp;
    MeshFluidSource *src = h->sources[id];
    vmath::vec3 j = predicted_ray_jitter(seed, h->sim->_dx);
    jitter_out[0] = j.x; jitter_out[1] = j.y; jitter_out[2] = j.z;
    std::vector<GridIndex> list;
    srand(seed);
    src->getCells(list);
    int I = h->sim->_isize, J = h->sim->_jsize, K = h->sim->_ksize;
    std::memset(cells, 0, (size_t)I * J * K);
    for (const GridIndex &g : list) cells[(size_t)g.i + (size_t)I * ((size_t)g.j + (size_t)J * g.k)] = 1;
}

// Domain-box solid SDF (nodal) and the derived weight grid
// (fluidsimulation.cpp:2918-2930, 3646-3730).
void ref_domain_solid(int isize, int jsize, int ksize, double dx, float *phi_nodal,
                      float *wU, float *wV, float *wW, float *wC) {
    FluidSimulation sim(isize, jsize, ksize, dx);
    sim.disableConsoleOutput();
    sim.disableSurfaceReconstruction();
    sim._initializeSimulationGrids(isize, jsize, ksize, dx);
    build_domain_solid(sim);
    sim._isWeightGridUpToDate = false;
    sim._updateWeightGrid();
    copy_array(sim._solidSDF._phi, phi_nodal);
    if (wU) copy_array(sim._weightGrid.U, wU);
    if (wV) copy_array(sim._weightGrid.V, wV);
    if (wW) copy_array(sim._weightGrid.W, wW);
    if (wC) copy_array(sim._weightGrid.center, wC);
}

// ParticleLevelSet::calculateSignedDistanceField + postProcess against a
// caller-supplied nodal solid phi (particlelevelset.cpp:161-196).
void ref_liquid_sdf(int isize, int jsize, int ksize, double dx, const float *pos, int n,
                    double radius, const float *solid_phi_nodal, float *phi_out) {
    FragmentedVector<MarkerParticle> particles;
    for (int i = 0; i < n; i++) {
        particles.push_back(MarkerParticle(vmath::vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
                                           vmath::vec3()));
    }
    ParticleLevelSet ls(isize, jsize, ksize, dx);
    ls.calculateSignedDistanceField(particles, radius);
    if (solid_phi_nodal) {
        MeshLevelSet solid(isize, jsize, ksize, dx);
        fill_array(solid._phi, solid_phi_nodal);
        ls.postProcessSignedDistanceField(solid);
    }
    copy_array(ls._phi, phi_out);
}

// VelocityAdvector::advect (velocityadvector.cpp:38-459) on a cleared MAC grid.
void ref_p2g(int isize, int jsize, int ksize, double dx, const float *pos, const float *vel,
             int n, double radius, float *u, float *v, float *w, uint8_t *vu, uint8_t *vv,
             uint8_t *vw) {
    FragmentedVector<MarkerParticle> particles;
    for (int i = 0; i < n; i++) {
        particles.push_back(MarkerParticle(vmath::vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
                                           vmath::vec3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2])));
    }
    MACVelocityField mac(isize, jsize, ksize, dx);
    ValidVelocityComponentGrid valid(isize, jsize, ksize);
    VelocityAdvectorParameters params;
    params.particles = &particles;
    params.vfield = &mac;
    params.validVelocities = &valid;
    params.particleRadius = radius;
    VelocityAdvector adv;
    adv.advect(params);
    copy_array(mac._u, u);
    copy_array(mac._v, v);
    copy_array(mac._w, w);
    copy_bool_array(valid.validU, vu);
    copy_bool_array(valid.validV, vv);
    copy_bool_array(valid.validW, vw);
}

// GridUtils::extrapolateGrid (gridutils.cpp:32-103) in place.
void ref_extrapolate(int ni, int nj, int nk, float *grid, uint8_t *valid, int layers) {
    Array3d<float> g(ni, nj, nk, 0.0f);
    Array3d<bool> vd(ni, nj, nk, false);
    fill_array(g, grid);
    fill_bool_array(vd, valid);
    GridUtils::extrapolateGrid(&g, &vd, layers);
    copy_array(g, grid);
}

// PressureSolver::solve (pressuresolver.cpp:44-77) with caller-supplied grids.
// The MAC and valid grids are updated in place.  Optionally returns the
// assembled right-hand side and solution in the reference's row order
// (k outer, j, i inner over fluid interior cells), and the PCG iterations /
// final infinity-norm residual.  Returns 1 on success, 0 on solver failure.
int ref_pressure_solve(int isize, int jsize, int ksize, double dx, double dt, double tol,
                       int maxiter, const float *liquid_phi, const float *solid_phi_nodal,
                       const float *wU, const float *wV, const float *wW, const float *wC,
                       float *u, float *v, float *w, uint8_t *vu, uint8_t *vv, uint8_t *vw,
                       float *su, float *sv, float *sw,
                       double *rhs_out, double *soln_out, int *n_rows, int *iters,
                       double *err) {
    MACVelocityField mac(isize, jsize, ksize, dx);
    fill_array(mac._u, u);
    fill_array(mac._v, v);
    fill_array(mac._w, w);
    ValidVelocityComponentGrid valid(isize, jsize, ksize);
    fill_bool_array(valid.validU, vu);
    fill_bool_array(valid.validV, vv);
    fill_bool_array(valid.validW, vw);
    ParticleLevelSet liquid(isize, jsize, ksize, dx);
    fill_array(liquid._phi, liquid_phi);
    MeshLevelSet solid(isize, jsize, ksize, dx);
    fill_array(solid._phi, solid_phi_nodal);
    // Solid face velocities (in/out: _conditionSolidVelocityField zeroes them
    // around isolated pockets); null = the zero field of a fresh level set.
    if (su) fill_array(solid._velocityData.field._u, (const float *)su);
    if (sv) fill_array(solid._velocityData.field._v, (const float *)sv);
    if (sw) fill_array(solid._velocityData.field._w, (const float *)sw);
    WeightGrid wg(isize, jsize, ksize);
    fill_array(wg.U, wU);
    fill_array(wg.V, wV);
    fill_array(wg.W, wW);
    fill_array(wg.center, wC);

    PressureSolverParameters params;
    params.cellwidth = dx;
    params.deltaTime = dt;
    params.tolerance = tol;
    params.acceptableTolerance = 1.0;
    params.maxIterations = maxiter;
    params.velocityField = &mac;
    params.validVelocities = &valid;
    params.liquidSDF = &liquid;
    params.solidSDF = &solid;
    params.weightGrid = &wg;
    params.isSurfaceTensionEnabled = false;

    // Same sequence as PressureSolver::solve, split so the intermediate
    // vectors can be returned.
    PressureSolver ps;
    ps._initialize(params);
    ps._conditionSolidVelocityField();
    ps._initializeSurfaceTensionClusterData();
    std::vector<double> rhs(ps._matSize, 0);
    ps._calculateNegativeDivergenceVector(rhs);
    if (n_rows) *n_rows = ps._matSize;
    if (rhs_out) {
        for (int i = 0; i < ps._matSize; i++) rhs_out[i] = rhs[i];
    }
    double maxAbsCoeff = 0.0;
    for (size_t i = 0; i < rhs.size(); i++) {
        if (std::fabs(rhs[i]) > maxAbsCoeff) maxAbsCoeff = std::fabs(rhs[i]);
    }
    int ok = 1;
    if (maxAbsCoeff < ps._pressureSolveTolerance) {
        if (iters) *iters = 0;
        if (err) *err = 0.0;
        if (soln_out) {
            for (int i = 0; i < ps._matSize; i++) soln_out[i] = 0.0;
        }
    } else {
        std::vector<double> soln(ps._matSize, 0);
        SparseMatrixd matrix(ps._matSize, 7);
        ps._calculateMatrixCoefficients(matrix);
        PCGSolver<double> solver;
        solver.setSolverParameters(ps._pressureSolveTolerance, ps._maxCGIterations);
        double estimatedError;
        int numIterations;
        bool success = solver.solve(matrix, rhs, soln, estimatedError, numIterations);
        if (!success && !(numIterations == ps._maxCGIterations &&
                          estimatedError < ps._pressureSolveAcceptableTolerance)) {
            ok = 0;
        }
        if (iters) *iters = numIterations;
        if (err) *err = estimatedError;
        if (soln_out) {
            for (int i = 0; i < ps._matSize; i++) soln_out[i] = soln[i];
        }
        if (ok) ps._applySolutionToVelocityField(soln);
    }
    if (su) copy_array(solid._velocityData.field._u, su);
    if (sv) copy_array(solid._velocityData.field._v, sv);
    if (sw) copy_array(solid._velocityData.field._w, sw);
    copy_array(mac._u, u);
    copy_array(mac._v, v);
    copy_array(mac._w, w);
    copy_bool_array(valid.validU, vu);
    copy_bool_array(valid.validV, vv);
    copy_bool_array(valid.validW, vw);
    return ok;
}

// MIC(0) factor + one preconditioner application z = M^-1 r on the system
// assembled by PressureSolver (pcgsolver.h:69-221).  Returns invdiag and z in
// reference row order.
void ref_mic0_apply(int isize, int jsize, int ksize, double dx, double dt,
                    const float *liquid_phi, const float *wU, const float *wV,
                    const float *wW, const float *wC, const double *r_in, double *invdiag_out,
                    double *z_out) {
    ParticleLevelSet liquid(isize, jsize, ksize, dx);
    fill_array(liquid._phi, liquid_phi);
    WeightGrid wg(isize, jsize, ksize);
    fill_array(wg.U, wU);
    fill_array(wg.V, wV);
    fill_array(wg.W, wW);
    fill_array(wg.center, wC);
    MACVelocityField mac(isize, jsize, ksize, dx);
    ValidVelocityComponentGrid valid(isize, jsize, ksize);
    MeshLevelSet solid(isize, jsize, ksize, dx);
    PressureSolverParameters params;
    params.cellwidth = dx;
    params.deltaTime = dt;
    params.tolerance = 1e-9;
    params.acceptableTolerance = 1.0;
    params.maxIterations = 1000;
    params.velocityField = &mac;
    params.validVelocities = &valid;
    params.liquidSDF = &liquid;
    params.solidSDF = &solid;
    params.weightGrid = &wg;
    PressureSolver ps;
    ps._initialize(params);
    SparseMatrixd matrix(ps._matSize, 7);
    ps._calculateMatrixCoefficients(matrix);
    PCGSolver<double> solver;
    solver.formPreconditioner(matrix);
    std::vector<double> r(r_in, r_in + ps._matSize), z(ps._matSize, 0.0);
    solver.applyPreconditioner(r, z);
    for (int i = 0; i < ps._matSize; i++) {
        invdiag_out[i] = solver.icfactor.invdiag[i];
        z_out[i] = z[i];
    }
}

// ParticleAdvector::advectParticlesRK{1,2,3,4} (particleadvector.cpp:137-336)
// through a caller-supplied MAC grid.
void ref_advect_particles(int isize, int jsize, int ksize, double dx, const float *u,
                          const float *v, const float *w, const float *pos, int n, double dt,
                          int order, float *out) {
    MACVelocityField mac(isize, jsize, ksize, dx);
    fill_array(mac._u, u);
    fill_array(mac._v, v);
    fill_array(mac._w, w);
    std::vector<vmath::vec3> pin(n), pout;
    for (int i = 0; i < n; i++) pin[i] = vmath::vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    ParticleAdvector adv;
    if (order == 4) adv.advectParticlesRK4(pin, &mac, dt, pout);
    else if (order == 3) adv.advectParticlesRK3(pin, &mac, dt, pout);
    else if (order == 2) adv.advectParticlesRK2(pin, &mac, dt, pout);
    else adv.advectParticlesRK1(pin, &mac, dt, pout);
    for (int i = 0; i < n; i++) {
        out[3 * i] = pout[i].x;
        out[3 * i + 1] = pout[i].y;
        out[3 * i + 2] = pout[i].z;
    }
}

// ParticleAdvector::trilinearInterpolate (particleadvector.cpp:338-410).
void ref_trilinear_velocity(int isize, int jsize, int ksize, double dx, const float *u,
                            const float *v, const float *w, const float *pos, int n,
                            float *out) {
    MACVelocityField mac(isize, jsize, ksize, dx);
    fill_array(mac._u, u);
    fill_array(mac._v, v);
    fill_array(mac._w, w);
    std::vector<vmath::vec3> pin(n), pout;
    for (int i = 0; i < n; i++) pin[i] = vmath::vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    ParticleAdvector adv;
    adv.trilinearInterpolate(pin, &mac, pout);
    for (int i = 0; i < n; i++) {
        out[3 * i] = pout[i].x;
        out[3 * i + 1] = pout[i].y;
        out[3 * i + 2] = pout[i].z;
    }
}

}  // extern "C"
