// Drives the C++ classes of include/flipmi_engine.hpp the way the reference
// engine's callers drive FluidSimulation, PressureSolver, VelocityAdvector and
// ParticleAdvector.  Inputs and outputs are raw little-endian arrays in a
// directory; the Python test writes the inputs and checks the outputs
// against the oracle.
//
//   wrapper_check <mode> <dir>
//   modes: levelset  pressure  p2g  advect  sim  sources  surface
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "flipmi_engine.hpp"

using namespace flipmi;

namespace {

std::string g_dir;

template <class T>
std::vector<T> read_raw(const std::string &name) {
    std::ifstream f(g_dir + "/" + name, std::ios::binary);
    if (!f) throw std::runtime_error("missing input " + name);
    f.seekg(0, std::ios::end);
    size_t bytes = (size_t)f.tellg();
    f.seekg(0);
    std::vector<T> v(bytes / sizeof(T));
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)bytes);
    return v;
}

template <class T>
void write_raw(const std::string &name, const T *p, size_t n) {
    std::ofstream f(g_dir + "/" + name, std::ios::binary);
    f.write(reinterpret_cast<const char *>(p), (std::streamsize)(n * sizeof(T)));
}

template <class T>
void load_into(Array3d<T> &a, const std::string &name) {
    std::vector<T> v = read_raw<T>(name);
    if ((int)v.size() != a.getNumElements()) throw std::runtime_error("size mismatch for " + name);
    std::copy(v.begin(), v.end(), a.getRawArray());
}

// bool grids travel as bytes (0 / 1).
void load_into(Array3d<bool> &a, const std::string &name) {
    std::vector<unsigned char> v = read_raw<unsigned char>(name);
    if ((int)v.size() != a.getNumElements()) throw std::runtime_error("size mismatch for " + name);
    for (size_t q = 0; q < v.size(); q++) a.getRawArray()[q] = v[q] != 0;
}

template <class T>
void save(const Array3d<T> &a, const std::string &name) {
    write_raw(name, a.getRawArray(), (size_t)a.getNumElements());
}

struct Dims {
    int I, J, K;
    double dx, dt;
    double extra;
};

Dims read_dims() {
    std::vector<double> d = read_raw<double>("dims.f64");
    return Dims{(int)d[0], (int)d[1], (int)d[2], d[3], d[4], d.size() > 5 ? d[5] : 0.0};
}

FragmentedVector<MarkerParticle> read_particles() {
    std::vector<float> p = read_raw<float>("particles.f32");  // n x 6: position, velocity
    FragmentedVector<MarkerParticle> out;
    for (size_t q = 0; q + 6 <= p.size(); q += 6)
        out.push_back(MarkerParticle(vmath::vec3(p[q], p[q + 1], p[q + 2]), vmath::vec3(p[q + 3], p[q + 4], p[q + 5])));
    return out;
}

// Domain walls, then ParticleLevelSet::calculateSignedDistanceField +
// postProcessSignedDistanceField as FluidSimulation calls them.
int run_levelset() {
    Dims d = read_dims();
    FragmentedVector<MarkerParticle> ps = read_particles();
    MeshLevelSet solid(d.I, d.J, d.K, d.dx);
    WeightGrid weights(d.I, d.J, d.K);
    solid.calculateDomainBoundary(&weights);
    ParticleLevelSet liquid(d.I, d.J, d.K, d.dx);
    liquid.calculateSignedDistanceField(ps, d.extra);
    save(*liquid.getPhiArray3d(), "phi_raw.f32");
    liquid.postProcessSignedDistanceField(solid);
    save(*liquid.getPhiArray3d(), "phi.f32");
    save(*solid.getPhiArray3d(), "solid.f32");
    save(weights.U, "wu.f32");
    save(weights.V, "wv.f32");
    save(weights.W, "ww.f32");
    save(weights.center, "wc.f32");
    return 0;
}

int run_pressure() {
    Dims d = read_dims();
    MACVelocityField mac(d.I, d.J, d.K, d.dx);
    load_into(*mac.getArray3dU(), "u.f32");
    load_into(*mac.getArray3dV(), "v.f32");
    load_into(*mac.getArray3dW(), "w.f32");
    ValidVelocityComponentGrid valid(d.I, d.J, d.K);
    load_into(valid.validU, "vu.u8");
    load_into(valid.validV, "vv.u8");
    load_into(valid.validW, "vw.u8");
    ParticleLevelSet liquid(d.I, d.J, d.K, d.dx);
    load_into(*liquid.getPhiArray3d(), "phi.f32");
    MeshLevelSet solid(d.I, d.J, d.K, d.dx);
    WeightGrid weights(d.I, d.J, d.K);
    load_into(weights.U, "wu.f32");
    load_into(weights.V, "wv.f32");
    load_into(weights.W, "ww.f32");
    load_into(weights.center, "wc.f32");

    PressureSolverParameters params;
    params.cellwidth = d.dx;
    params.deltaTime = d.dt;
    params.tolerance = 1e-9;
    params.acceptableTolerance = 1.0;
    params.maxIterations = 1000;
    params.velocityField = &mac;
    params.validVelocities = &valid;
    params.liquidSDF = &liquid;
    params.solidSDF = &solid;
    params.weightGrid = &weights;
    params.precision = (int)d.extra;
    PressureSolver solver;
    bool ok = solver.solve(params);
    save(*mac.getArray3dU(), "u_out.f32");
    save(*mac.getArray3dV(), "v_out.f32");
    save(*mac.getArray3dW(), "w_out.f32");
    save(valid.validU, "vu_out.u8");
    save(valid.validV, "vv_out.u8");
    save(valid.validW, "vw_out.u8");
    std::ofstream(g_dir + "/status.txt") << (ok ? 1 : 0) << "\n" << solver.getSolverStatus() << "\n";
    return 0;
}

int run_p2g() {
    Dims d = read_dims();
    FragmentedVector<MarkerParticle> ps = read_particles();
    MACVelocityField mac(d.I, d.J, d.K, d.dx);
    ValidVelocityComponentGrid valid(d.I, d.J, d.K);
    VelocityAdvectorParameters params;
    params.particles = &ps;
    params.vfield = &mac;
    params.validVelocities = &valid;
    params.particleRadius = d.extra;
    VelocityAdvector adv;
    adv.advect(params);
    save(*mac.getArray3dU(), "u.f32");
    save(*mac.getArray3dV(), "v.f32");
    save(*mac.getArray3dW(), "w.f32");
    save(valid.validU, "vu.u8");
    save(valid.validV, "vv.u8");
    save(valid.validW, "vw.u8");
    return 0;
}

int run_advect() {
    Dims d = read_dims();
    MACVelocityField mac(d.I, d.J, d.K, d.dx);
    load_into(*mac.getArray3dU(), "u.f32");
    load_into(*mac.getArray3dV(), "v.f32");
    load_into(*mac.getArray3dW(), "w.f32");
    std::vector<float> raw = read_raw<float>("points.f32");
    std::vector<vmath::vec3> pts(raw.size() / 3);
    for (size_t q = 0; q < pts.size(); q++) pts[q] = vmath::vec3(raw[3 * q], raw[3 * q + 1], raw[3 * q + 2]);
    ParticleAdvector adv;
    if (!adv.initialize()) throw std::runtime_error(adv.getInitializationErrorMessage());
    std::vector<vmath::vec3> out;
    adv.advectParticlesRK1(pts, &mac, d.dt, out);
    write_raw("rk1.f32", &out[0].x, 3 * out.size());
    adv.advectParticlesRK2(pts, &mac, d.dt, out);
    write_raw("rk2.f32", &out[0].x, 3 * out.size());
    adv.advectParticlesRK3(pts, &mac, d.dt, out);
    write_raw("rk3.f32", &out[0].x, 3 * out.size());
    adv.advectParticlesRK4(pts, &mac, d.dt, out);
    write_raw("rk4.f32", &out[0].x, 3 * out.size());
    adv.trilinearInterpolate(pts, &mac, out);
    write_raw("vel.f32", &out[0].x, 3 * out.size());
    std::vector<vmath::vec3> inplace = pts;
    adv.trilinearInterpolate(inplace, &mac);
    write_raw("vel_inplace.f32", &inplace[0].x, 3 * inplace.size());
    return 0;
}

// FluidSimulation set up as the dam-break workloads are, then `frames`
// updates of dt (extra = frames).
int run_sim() {
    Dims d = read_dims();
    std::vector<float> pos = read_raw<float>("positions.f32"), vel = read_raw<float>("velocities.f32");
    FluidSimulation sim(d.I, d.J, d.K, d.dx);
    sim.setMaxThreadCount(8);
    sim.disableSurfaceReconstruction();
    sim.disableConsoleOutput();
    sim.addBodyForce(0.0, -9.81, 0.0);
    sim.setPICFLIPRatio(0.05);
    sim.setCFLConditionNumber(5);
    sim.setMinTimeStepsPerFrame(1);
    sim.setMaxTimeStepsPerFrame(1);
    FluidSimulationMarkerParticleData data;
    data.size = (int)(pos.size() / 3);
    data.positions = reinterpret_cast<char *>(pos.data());
    data.velocities = reinterpret_cast<char *>(vel.data());
    sim.loadMarkerParticleData(data);
    sim.initialize();
    bool threw = false;
    try {
        sim.update(-1.0);
    } catch (const std::domain_error &) {
        threw = true;
    }
    if (!threw) throw std::runtime_error("update(-1) did not throw std::domain_error");
    int frames = (int)d.extra;
    std::ofstream st(g_dir + "/stats.txt");
    for (int f = 0; f < frames; f++) {
        sim.update(d.dt);
        FluidSimulationFrameStats fs = sim.getFrameStatsData();
        st << fs.frame << " " << fs.substeps << " " << fs.fluidParticles << " "
           << sim.getLastPressureSolverIterations() << " " << (sim.getLastPressureSolveSucceeded() ? 1 : 0) << "\n";
    }
    unsigned int n = sim.getNumMarkerParticles();
    std::vector<char> pblob(sim.getMarkerParticlePositionDataSize()), vblob(sim.getMarkerParticleVelocityDataSize());
    if (n) {
        sim.getMarkerParticlePositionData(pblob.data());
        sim.getMarkerParticleVelocityData(vblob.data());
    }
    write_raw("positions_out.f32", pblob.data(), pblob.size());
    write_raw("velocities_out.f32", vblob.data(), vblob.size());
    // fluidsimulation.cpp:2314-2338: a sub-range, and the domain_error of an
    // invalid one.
    if (n >= 8) {
        std::vector<char> pr(12 * (size_t)(n - 5)), vr(12 * (size_t)(n - 5));
        sim.getMarkerParticlePositionDataRange(2, (int)n - 3, pr.data());
        sim.getMarkerParticleVelocityDataRange(2, (int)n - 3, vr.data());
        write_raw("positions_range.f32", pr.data(), pr.size());
        write_raw("velocities_range.f32", vr.data(), vr.size());
    }
    bool range_threw = false;
    try {
        std::vector<char> tmp(24);
        sim.getMarkerParticlePositionDataRange(3, 1, tmp.data());
    } catch (const std::domain_error &) {
        range_threw = true;
    }
    if (!range_threw) throw std::runtime_error("an invalid range did not throw std::domain_error");
    MACVelocityField *mac = sim.getVelocityField();
    save(*mac->getArray3dU(), "u_out.f32");
    save(*mac->getArray3dV(), "v_out.f32");
    save(*mac->getArray3dW(), "w_out.f32");
    return 0;
}

}  // namespace

TriangleMesh read_mesh(const std::string &stem) {
    std::vector<float> v = read_raw<float>(stem + "_verts.f32");
    std::vector<int> t = read_raw<int>(stem + "_tris.i32");
    TriangleMesh m;
    for (size_t q = 0; q + 2 < v.size(); q += 3) m.vertices.push_back(vmath::vec3(v[q], v[q + 1], v[q + 2]));
    for (size_t q = 0; q + 2 < t.size(); q += 3) {
        Triangle tri;
        tri.tri[0] = t[q]; tri.tri[1] = t[q + 1]; tri.tri[2] = t[q + 2];
        m.triangles.push_back(tri);
    }
    return m;
}

// A pool with a static obstacle (MeshObject + addMeshObstacle) and an inflow
// source (MeshFluidSource + addMeshFluidSource), stepped for d.extra frames.
int run_sources() {
    Dims d = read_dims();
    std::vector<float> pos = read_raw<float>("positions.f32"), vel = read_raw<float>("velocities.f32");
    std::vector<double> src_vel = read_raw<double>("src_vel.f64");
    FluidSimulation sim(d.I, d.J, d.K, d.dx);
    sim.addBodyForce(0.0, -9.81, 0.0);
    sim.setMinTimeStepsPerFrame(1);
    sim.setMaxTimeStepsPerFrame(6);
    sim.setMarkerParticleJitterSeed(7);
    FluidSimulationMarkerParticleData data;
    data.size = (int)(pos.size() / 3);
    data.positions = reinterpret_cast<char *>(pos.data());
    data.velocities = reinterpret_cast<char *>(vel.data());
    sim.loadMarkerParticleData(data);
    MeshObject obstacle(d.I, d.J, d.K, d.dx);
    obstacle.updateMeshStatic(read_mesh("obst"));
    sim.addMeshObstacle(&obstacle);
    MeshFluidSource source(d.I, d.J, d.K, d.dx);
    source.updateMeshStatic(read_mesh("src"));
    source.setInflow();
    source.setVelocity(vmath::vec3(src_vel[0], src_vel[1], src_vel[2]));
    sim.addMeshFluidSource(&source);
    bool threw = false;
    try {
        sim.addMeshFluidSource(&source);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    if (!threw) throw std::runtime_error("adding a source twice did not throw");
    sim.initialize();
    std::ofstream st(g_dir + "/stats.txt");
    for (int f = 0; f < (int)d.extra; f++) {
        sim.update(d.dt);
        FluidSimulationFrameStats fs = sim.getFrameStatsData();
        st << fs.frame << " " << fs.substeps << " " << fs.fluidParticles << "\n";
    }
    std::vector<char> pblob(sim.getMarkerParticlePositionDataSize()), vblob(sim.getMarkerParticleVelocityDataSize());
    if (!pblob.empty()) {
        sim.getMarkerParticlePositionData(pblob.data());
        sim.getMarkerParticleVelocityData(vblob.data());
    }
    write_raw("positions_out.f32", pblob.data(), pblob.size());
    write_raw("velocities_out.f32", vblob.data(), vblob.size());
    return 0;
}

// FluidSimulation with the viewer's surface setting (subdivision level 2,
// FluidManager.cpp:52) for d.extra frames of one substep each; the isomesh
// is read and turned into vertex normals the way FluidManager::iGetVertices
// and iGetIndices do (FluidManager.cpp:104-170).  Then ParticleMesher::
// meshParticles on the loaded particles over the domain's solid SDF.
int run_surface() {
    Dims d = read_dims();
    std::vector<float> pos = read_raw<float>("positions.f32"), vel = read_raw<float>("velocities.f32");
    FluidSimulation sim(d.I, d.J, d.K, d.dx);
    if (!sim.isSurfaceReconstructionEnabled()) throw std::runtime_error("surface reconstruction is off by default");
    sim.setSurfaceSubdivisionLevel(2);
    sim.addBodyForce(0.0, -9.81, 0.0);
    sim.setMinTimeStepsPerFrame(1);
    sim.setMaxTimeStepsPerFrame(1);
    FluidSimulationMarkerParticleData data;
    data.size = (int)(pos.size() / 3);
    data.positions = reinterpret_cast<char *>(pos.data());
    data.velocities = reinterpret_cast<char *>(vel.data());
    sim.loadMarkerParticleData(data);
    sim.initialize();
    if (!sim.getIsomesh().vertices.empty()) throw std::runtime_error("a mesh before the first update");
    for (int f = 0; f < (int)d.extra; f++) sim.update(d.dt);
    using namespace vmath;
    TriangleMesh isomesh = sim.getIsomesh();
    std::vector<vec3> normal(isomesh.vertices.size(), vec3(0.0f, 0.0f, 0.0f));
    std::vector<unsigned int> indices;
    for (size_t i = 0; i < isomesh.triangles.size(); i++) {
        int i0 = isomesh.triangles[i].tri[0], i1 = isomesh.triangles[i].tri[1], i2 = isomesh.triangles[i].tri[2];
        vec3 e0 = isomesh.vertices[i1] - isomesh.vertices[i0];
        vec3 e1 = isomesh.vertices[i2] - isomesh.vertices[i0];
        vec3 faceN = cross(e0, e1);
        normal[i0] += faceN;
        normal[i1] += faceN;
        normal[i2] += faceN;
        for (int j = 0; j < 3; j++) indices.push_back(static_cast<unsigned int>(isomesh.triangles[i].tri[j]));
    }
    for (vec3 &n : normal) n = normalize(n);
    write_raw("iso_verts.f32", reinterpret_cast<const float *>(isomesh.vertices.data()), 3 * isomesh.vertices.size());
    write_raw("iso_tris.i32", indices.data(), indices.size());
    write_raw("iso_normals.f32", reinterpret_cast<const float *>(normal.data()), 3 * normal.size());
    std::vector<char> pblob(sim.getMarkerParticlePositionDataSize());
    if (!pblob.empty()) sim.getMarkerParticlePositionData(pblob.data());
    write_raw("positions_out.f32", pblob.data(), pblob.size());

    MeshLevelSet solid(d.I, d.J, d.K, d.dx);
    solid.calculateDomainBoundary();
    std::vector<vmath::vec3> particles;
    for (size_t q = 0; q + 2 < pos.size(); q += 3) particles.push_back(vmath::vec3(pos[q], pos[q + 1], pos[q + 2]));
    ParticleMesherParameters params;
    params.isize = d.I;
    params.jsize = d.J;
    params.ksize = d.K;
    params.dx = d.dx;
    params.subdivisions = 1;
    params.computechunks = 1;
    const double volume = d.dx * d.dx * d.dx / 8.0, pi = 3.141592653;
    params.radius = std::pow(3 * volume / (4 * pi), 1.0 / 3.0) * 3.0;
    params.particles = &particles;
    params.solidSDF = &solid;
    ParticleMesher mesher;
    TriangleMesh m = mesher.meshParticles(params);
    write_raw("mesher_verts.f32", reinterpret_cast<const float *>(m.vertices.data()), 3 * m.vertices.size());
    write_raw("mesher_tris.i32", reinterpret_cast<const int *>(m.triangles.data()), 3 * m.triangles.size());
    return 0;
}

int main(int argc, char **argv) {
    if (argc != 3) {
        std::cerr << "usage: wrapper_check <levelset|pressure|p2g|advect|sim> <dir>\n";
        return 2;
    }
    std::string mode = argv[1];
    g_dir = argv[2];
    std::cout << "wrapper_check " << mode << " " << g_dir << std::endl;
    try {
        if (mode == "levelset") return run_levelset();
        if (mode == "pressure") return run_pressure();
        if (mode == "p2g") return run_p2g();
        if (mode == "advect") return run_advect();
        if (mode == "sim") return run_sim();
        if (mode == "sources") return run_sources();
        if (mode == "surface") return run_surface();
    } catch (const std::exception &e) {
        std::cerr << "wrapper_check " << mode << ": " << e.what() << "\n";
        return 1;
    }
    std::cerr << "unknown mode " << mode << "\n";
    return 2;
}
