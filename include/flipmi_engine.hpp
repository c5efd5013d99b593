/*
 * flipmi_engine.hpp -- C++ classes with the reference engine's signatures
 * (jklae/FLIPEngine3D, src/engine), forwarding to the flipmi C ABI.
 *
 * A caller of the reference's FluidSimulation, PressureSolver,
 * VelocityAdvector or ParticleAdvector switches by including this header and
 * using namespace flipmi instead of the engine headers: class names, method
 * signatures, parameter structs and data layouts are the reference's.  All
 * computation runs on the HIP device; there is no CPU path.
 *
 * Errors: PressureSolver::solve returns false with a status text as the
 * reference does; FluidSimulation throws std::runtime_error /
 * std::domain_error where the reference throws; any device or argument error
 * from the ABI becomes std::runtime_error carrying flipmi_last_error().
 *
 * Header-only; link with -lflipmi.
 */
#ifndef FLIPMI_ENGINE_HPP
#define FLIPMI_ENGINE_HPP

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "flipmi.h"

namespace flipmi {

// ---------------------------------------------------------------------------
// Device context shared by all objects of the process (the reference sizes its
// worker threads process-wide, threadutils.cpp:43-52; here the shared resource
// is the device).
// ---------------------------------------------------------------------------

namespace detail {
inline int &device_index() {
    static int d = 0;
    return d;
}
inline std::mutex &device_mutex() {
    static std::mutex m;
    return m;
}
inline flipmi_ctx *&device_slot() {
    static flipmi_ctx *c = nullptr;
    return c;
}
inline flipmi_ctx *ctx() {
    std::lock_guard<std::mutex> g(device_mutex());
    flipmi_ctx *&c = device_slot();
    if (!c) {
        flipmi_ctx *made = nullptr;
        int rc = flipmi_ctx_create(device_index(), &made);
        if (rc != FLIPMI_OK)
            throw std::runtime_error("flipmi: no usable HIP device (status " + std::to_string(rc) + ")");
        c = made;
    }
    return c;
}
inline void check(int rc, const char *what) {
    if (rc != FLIPMI_OK) throw std::runtime_error(std::string(what) + ": " + flipmi_last_error(ctx()));
}
}  // namespace detail

// Selects the HIP device used by objects created afterwards (before the first
// call that touches the device).
inline void setDevice(int device) {
    std::lock_guard<std::mutex> g(detail::device_mutex());
    if (detail::device_slot()) throw std::runtime_error("flipmi: the device is already in use");
    detail::device_index() = device;
}

// ---------------------------------------------------------------------------
// Value types (vmath.h, array3d.h, gridindexkeymap / grid3d.h)
// ---------------------------------------------------------------------------

namespace vmath {
// 3 x float, 12 bytes (vmath.h:76-97).
struct vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    vec3() {}
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    vec3(double a, double b, double c) : x((float)a), y((float)b), z((float)c) {}
    float operator[](int q) const { return q == 0 ? x : (q == 1 ? y : z); }
    float &operator[](int q) { return q == 0 ? x : (q == 1 ? y : z); }
};
// vmath.cpp:33-123: componentwise float arithmetic; division multiplies by
// a float reciprocal formed in double.
inline vec3 operator+(const vec3 &a, const vec3 &b) { return vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline vec3 operator-(const vec3 &a, const vec3 &b) { return vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline vec3 operator-(const vec3 &a) { return vec3(-a.x, -a.y, -a.z); }
inline vec3 operator*(float s, const vec3 &v) { return vec3(v.x * s, v.y * s, v.z * s); }
inline vec3 operator*(const vec3 &v, float s) { return vec3(v.x * s, v.y * s, v.z * s); }
inline vec3 operator/(const vec3 &v, float s) {
    float inv = (float)(1.0 / s);
    return vec3(v.x * inv, v.y * inv, v.z * inv);
}
inline vec3 &operator+=(vec3 &a, const vec3 &b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline vec3 &operator-=(vec3 &a, const vec3 &b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
inline vec3 &operator*=(vec3 &v, float s) { v.x *= s; v.y *= s; v.z *= s; return v; }
inline vec3 &operator/=(vec3 &v, float s) { float inv = (float)(1.0 / s); v.x *= inv; v.y *= inv; v.z *= inv; return v; }
inline float dot(const vec3 &a, const vec3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3 cross(const vec3 &a, const vec3 &b) {
    return vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float lengthsq(const vec3 &v) { return dot(v, v); }
inline float length(const vec3 &v) { return (float)std::sqrt((double)lengthsq(v)); }
inline vec3 normalize(const vec3 &v) { return v / length(v); }
}  // namespace vmath
static_assert(sizeof(vmath::vec3) == 12, "vec3 must be 3 packed floats");

struct GridIndex {
    int i = 0, j = 0, k = 0;
    GridIndex() {}
    GridIndex(int ii, int jj, int kk) : i(ii), j(jj), k(kk) {}
};

// Dense 3-D array, i fastest (array3d.h:63-450); bool elements are one byte.
template <class T>
class Array3d {
public:
    int width = 0, height = 0, depth = 0;

    Array3d() {}
    Array3d(int i, int j, int k) : width(i), height(j), depth(k), _data(new T[count(i, j, k)]()) {}
    Array3d(int i, int j, int k, T v) : Array3d(i, j, k) { fill(v); }
    Array3d(const Array3d &o) : Array3d(o.width, o.height, o.depth) {
        std::copy(o._data.get(), o._data.get() + o.getNumElements(), _data.get());
    }
    Array3d &operator=(const Array3d &o) {
        if (this != &o) {
            Array3d t(o);
            swap(t);
        }
        return *this;
    }
    Array3d(Array3d &&) = default;
    Array3d &operator=(Array3d &&) = default;

    T get(int i, int j, int k) const { return _data[index(i, j, k)]; }
    T get(GridIndex g) const { return get(g.i, g.j, g.k); }
    T operator()(int i, int j, int k) const { return get(i, j, k); }
    T operator()(GridIndex g) const { return get(g); }
    void set(int i, int j, int k, T v) { _data[index(i, j, k)] = v; }
    void set(GridIndex g, T v) { set(g.i, g.j, g.k, v); }
    void add(int i, int j, int k, T v) { _data[index(i, j, k)] += v; }
    void fill(T v) { std::fill(_data.get(), _data.get() + getNumElements(), v); }
    bool isIndexInRange(int i, int j, int k) const {
        return i >= 0 && j >= 0 && k >= 0 && i < width && j < height && k < depth;
    }
    int getNumElements() const { return width * height * depth; }
    T *getRawArray() { return _data.get(); }
    const T *getRawArray() const { return _data.get(); }

private:
    static size_t count(int i, int j, int k) { return (size_t)i * (size_t)j * (size_t)k; }
    size_t index(int i, int j, int k) const {
        return (size_t)i + (size_t)width * ((size_t)j + (size_t)height * (size_t)k);
    }
    void swap(Array3d &o) {
        std::swap(width, o.width);
        std::swap(height, o.height);
        std::swap(depth, o.depth);
        std::swap(_data, o._data);
    }
    std::unique_ptr<T[]> _data;
};
static_assert(sizeof(bool) == 1, "bool grids are passed as bytes");

// MarkerParticle (markerparticle.h:30-44): position and velocity, 24 bytes.
struct MarkerParticle {
    vmath::vec3 position;
    vmath::vec3 velocity;
    MarkerParticle() {}
    MarkerParticle(vmath::vec3 p) : position(p) {}
    MarkerParticle(vmath::vec3 p, vmath::vec3 v) : position(p), velocity(v) {}
    MarkerParticle(double x, double y, double z) : position(x, y, z) {}
};
static_assert(sizeof(MarkerParticle) == 24, "MarkerParticle must be 6 packed floats");

// FragmentedVector<T> (fragmentedvector.h:32-328): the element interface the
// callers use, stored contiguously.
template <class T>
class FragmentedVector {
public:
    void push_back(const T &v) { _v.push_back(v); }
    void pop_back() { _v.pop_back(); }
    size_t size() const { return _v.size(); }
    bool empty() const { return _v.empty(); }
    T &operator[](size_t q) { return _v[q]; }
    const T &operator[](size_t q) const { return _v[q]; }
    T &at(size_t q) { return _v.at(q); }
    T &front() { return _v.front(); }
    T &back() { return _v.back(); }
    void clear() { _v.clear(); }
    void reserve(size_t n) { _v.reserve(n); }
    void shrink_to_fit() { _v.shrink_to_fit(); }
    void insert(const std::vector<T> &vs) { _v.insert(_v.end(), vs.begin(), vs.end()); }
    T *data() { return _v.data(); }
    const T *data() const { return _v.data(); }

private:
    std::vector<T> _v;
};

// ---------------------------------------------------------------------------
// Grids (macvelocityfield.h:36-164, pressuresolver.h:48-60)
// ---------------------------------------------------------------------------

struct ValidVelocityComponentGrid {
    Array3d<bool> validU, validV, validW;
    ValidVelocityComponentGrid() {}
    ValidVelocityComponentGrid(int i, int j, int k)
        : validU(i + 1, j, k, false), validV(i, j + 1, k, false), validW(i, j, k + 1, false) {}
    void reset() {
        validU.fill(false);
        validV.fill(false);
        validW.fill(false);
    }
};

struct WeightGrid {
    Array3d<float> center, U, V, W;
    WeightGrid() {}
    WeightGrid(int i, int j, int k)
        : center(i, j, k, 0.0f), U(i + 1, j, k, 0.0f), V(i, j + 1, k, 0.0f), W(i, j, k + 1, 0.0f) {}
};

class MACVelocityField {
public:
    MACVelocityField() {}
    MACVelocityField(int isize, int jsize, int ksize, double dx)
        : _isize(isize), _jsize(jsize), _ksize(ksize), _dx(dx), _u(isize + 1, jsize, ksize, 0.0f),
          _v(isize, jsize + 1, ksize, 0.0f), _w(isize, jsize, ksize + 1, 0.0f) {}

    void getGridDimensions(int *i, int *j, int *k) {
        *i = _isize;
        *j = _jsize;
        *k = _ksize;
    }
    double getGridCellSize() { return _dx; }

    float U(int i, int j, int k) { return _u(i, j, k); }
    float V(int i, int j, int k) { return _v(i, j, k); }
    float W(int i, int j, int k) { return _w(i, j, k); }
    float U(GridIndex g) { return _u(g); }
    float V(GridIndex g) { return _v(g); }
    float W(GridIndex g) { return _w(g); }
    void set(MACVelocityField &o) { *this = o; }
    void setU(int i, int j, int k, double val) { _u.set(i, j, k, (float)val); }
    void setV(int i, int j, int k, double val) { _v.set(i, j, k, (float)val); }
    void setW(int i, int j, int k, double val) { _w.set(i, j, k, (float)val); }
    void setU(GridIndex g, double val) { _u.set(g, (float)val); }
    void setV(GridIndex g, double val) { _v.set(g, (float)val); }
    void setW(GridIndex g, double val) { _w.set(g, (float)val); }
    void setU(Array3d<float> &g) { _u = g; }
    void setV(Array3d<float> &g) { _v = g; }
    void setW(Array3d<float> &g) { _w = g; }
    // addU(double) stores through float (macvelocityfield.cpp:247-293).
    void addU(int i, int j, int k, double val) { _u.set(i, j, k, (float)(_u(i, j, k) + val)); }
    void addV(int i, int j, int k, double val) { _v.set(i, j, k, (float)(_v(i, j, k) + val)); }
    void addW(int i, int j, int k, double val) { _w.set(i, j, k, (float)(_w(i, j, k) + val)); }
    Array3d<float> *getArray3dU() { return &_u; }
    Array3d<float> *getArray3dV() { return &_v; }
    Array3d<float> *getArray3dW() { return &_w; }
    float *getRawArrayU() { return _u.getRawArray(); }
    float *getRawArrayV() { return _v.getRawArray(); }
    float *getRawArrayW() { return _w.getRawArray(); }
    void clear() {
        clearU();
        clearV();
        clearW();
    }
    void clearU() { _u.fill(0.0f); }
    void clearV() { _v.fill(0.0f); }
    void clearW() { _w.fill(0.0f); }
    bool isIndexInRangeU(int i, int j, int k) { return _u.isIndexInRange(i, j, k); }
    bool isIndexInRangeV(int i, int j, int k) { return _v.isIndexInRange(i, j, k); }
    bool isIndexInRangeW(int i, int j, int k) { return _w.isIndexInRange(i, j, k); }

private:
    int _isize = 0, _jsize = 0, _ksize = 0;
    double _dx = 0.0;
    Array3d<float> _u, _v, _w;
};

// Solid signed distance, nodal (I+1)(J+1)(K+1), with the solid face
// velocities (meshlevelset.h:94-137).
class MeshLevelSet {
public:
    MeshLevelSet() {}
    MeshLevelSet(int isize, int jsize, int ksize, double dx)
        : _isize(isize), _jsize(jsize), _ksize(ksize), _dx(dx), _phi(isize + 1, jsize + 1, ksize + 1, 0.0f),
          _vu(isize + 1, jsize, ksize, 0.0f), _vv(isize, jsize + 1, ksize, 0.0f), _vw(isize, jsize, ksize + 1, 0.0f) {}

    float operator()(int i, int j, int k) { return _phi(i, j, k); }
    float operator()(GridIndex g) { return _phi(g); }
    float get(int i, int j, int k) { return _phi(i, j, k); }
    float get(GridIndex g) { return _phi(g); }
    void set(int i, int j, int k, float d) { _phi.set(i, j, k, d); }
    void set(GridIndex g, float d) { _phi.set(g, d); }
    // meshlevelset.cpp:152-162
    float getDistanceAtCellCenter(int i, int j, int k) {
        return 0.125f * (_phi(i, j, k) + _phi(i + 1, j, k) + _phi(i, j + 1, k) + _phi(i + 1, j + 1, k) +
                         _phi(i, j, k + 1) + _phi(i + 1, j, k + 1) + _phi(i, j + 1, k + 1) +
                         _phi(i + 1, j + 1, k + 1));
    }
    float getFaceVelocityU(int i, int j, int k) { return _vu(i, j, k); }
    float getFaceVelocityV(int i, int j, int k) { return _vv(i, j, k); }
    float getFaceVelocityW(int i, int j, int k) { return _vw(i, j, k); }
    void setFaceVelocityU(int i, int j, int k, float v) { _vu.set(i, j, k, v); }
    void setFaceVelocityV(int i, int j, int k, float v) { _vv.set(i, j, k, v); }
    void setFaceVelocityW(int i, int j, int k, float v) { _vw.set(i, j, k, v); }

    // The domain's wall box as FluidSimulation builds it at initialisation
    // (_addStaticObjectsToSDF and _updateWeightGrid, fluidsimulation.cpp:
    // 2918-2930, 3646-3730), computed on the device; weights may be null.
    void calculateDomainBoundary(WeightGrid *weights = nullptr) {
        detail::check(flipmi_domain_solid(detail::ctx(), _isize, _jsize, _ksize, _dx, _phi.getRawArray(),
                                          weights ? weights->U.getRawArray() : nullptr,
                                          weights ? weights->V.getRawArray() : nullptr,
                                          weights ? weights->W.getRawArray() : nullptr,
                                          weights ? weights->center.getRawArray() : nullptr),
                      "MeshLevelSet::calculateDomainBoundary");
    }

    Array3d<float> *getPhiArray3d() { return &_phi; }
    float *getRawFaceVelocityU() { return _vu.getRawArray(); }
    float *getRawFaceVelocityV() { return _vv.getRawArray(); }
    float *getRawFaceVelocityW() { return _vw.getRawArray(); }
    void getGridDimensions(int *i, int *j, int *k) {
        *i = _isize;
        *j = _jsize;
        *k = _ksize;
    }
    double getCellSize() { return _dx; }

private:
    int _isize = 0, _jsize = 0, _ksize = 0;
    double _dx = 0.0;
    Array3d<float> _phi, _vu, _vv, _vw;
};

// Liquid signed distance, cell-centred (particlelevelset.h:47-69).
class ParticleLevelSet {
public:
    ParticleLevelSet() {}
    ParticleLevelSet(int i, int j, int k, double dx) : _isize(i), _jsize(j), _ksize(k), _dx(dx), _phi(i, j, k, 0.0f) {}

    float operator()(int i, int j, int k) { return _phi(i, j, k); }
    float operator()(GridIndex g) { return _phi(g); }
    float get(int i, int j, int k) { return _phi(i, j, k); }
    float get(GridIndex g) { return _phi(g); }

    // particlelevelset.cpp:161-170
    void calculateSignedDistanceField(FragmentedVector<MarkerParticle> &particles, double radius) {
        std::vector<float> pos(3 * particles.size());
        for (size_t q = 0; q < particles.size(); q++) {
            pos[3 * q] = particles[q].position.x;
            pos[3 * q + 1] = particles[q].position.y;
            pos[3 * q + 2] = particles[q].position.z;
        }
        detail::check(flipmi_liquid_sdf(detail::ctx(), _isize, _jsize, _ksize, _dx, pos.data(),
                                        (int64_t)particles.size(), radius, nullptr, _phi.getRawArray()),
                      "ParticleLevelSet::calculateSignedDistanceField");
    }
    // particlelevelset.cpp:172-196
    void postProcessSignedDistanceField(MeshLevelSet &solidPhi) {
        detail::check(flipmi_liquid_sdf_postprocess(detail::ctx(), _isize, _jsize, _ksize, _dx,
                                                    solidPhi.getPhiArray3d()->getRawArray(), _phi.getRawArray()),
                      "ParticleLevelSet::postProcessSignedDistanceField");
    }

    Array3d<float> *getPhiArray3d() { return &_phi; }

private:
    int _isize = 0, _jsize = 0, _ksize = 0;
    double _dx = 0.0;
    Array3d<float> _phi;
};

// ---------------------------------------------------------------------------
// PressureSolver (pressuresolver.h:63-92)
// ---------------------------------------------------------------------------

struct PressureSolverParameters {
    double cellwidth;
    double deltaTime;
    double tolerance;
    double acceptableTolerance;
    int maxIterations;

    MACVelocityField *velocityField;
    ValidVelocityComponentGrid *validVelocities;
    ParticleLevelSet *liquidSDF;
    MeshLevelSet *solidSDF;
    WeightGrid *weightGrid;

    bool isSurfaceTensionEnabled = false;
    double surfaceTensionConstant;
    Array3d<float> *curvatureGrid;

    // Not in the reference: FLIPMI_PCG_FP64 keeps the reference's arithmetic.
    int precision = FLIPMI_PCG_FP64;
};

class PressureSolver {
public:
    PressureSolver() {}

    // pressuresolver.cpp:44-77: projects velocityField in place and marks
    // the projected faces valid; returns false (velocity untouched) when the
    // solve fails.  Surface tension is not available on this path.
    bool solve(PressureSolverParameters params) {
        if (params.isSurfaceTensionEnabled) {
            _status = "***Pressure Solver FAILED\nsurface tension is not supported";
            return false;
        }
        int I, J, K;
        params.velocityField->getGridDimensions(&I, &J, &K);
        flipmi_pressure_args a;
        std::memset(&a, 0, sizeof(a));
        a.isize = I;
        a.jsize = J;
        a.ksize = K;
        a.dx = params.cellwidth;
        a.dt = params.deltaTime;
        a.tolerance = params.tolerance;
        a.acceptable_tolerance = params.acceptableTolerance;
        a.max_iterations = params.maxIterations;
        a.precision = params.precision;
        a.liquid_phi = params.liquidSDF->getPhiArray3d()->getRawArray();
        a.weight_u = params.weightGrid->U.getRawArray();
        a.weight_v = params.weightGrid->V.getRawArray();
        a.weight_w = params.weightGrid->W.getRawArray();
        a.weight_c = params.weightGrid->center.getRawArray();
        if (params.solidSDF) {
            a.solid_u = params.solidSDF->getRawFaceVelocityU();
            a.solid_v = params.solidSDF->getRawFaceVelocityV();
            a.solid_w = params.solidSDF->getRawFaceVelocityW();
        }
        a.u = params.velocityField->getRawArrayU();
        a.v = params.velocityField->getRawArrayV();
        a.w = params.velocityField->getRawArrayW();
        a.valid_u = reinterpret_cast<uint8_t *>(params.validVelocities->validU.getRawArray());
        a.valid_v = reinterpret_cast<uint8_t *>(params.validVelocities->validV.getRawArray());
        a.valid_w = reinterpret_cast<uint8_t *>(params.validVelocities->validW.getRawArray());
        flipmi_pressure_report rep;
        std::memset(&rep, 0, sizeof(rep));
        int rc = flipmi_pressure_solve(detail::ctx(), &a, &rep);
        if (rc != FLIPMI_OK && rc != FLIPMI_ERR_NOT_CONVERGED) detail::check(rc, "PressureSolver::solve");
        // Status text of PressureSolver::_solveLinearSystem (pressuresolver.cpp:810-840).
        std::ostringstream ss;
        if (rc != FLIPMI_OK) ss << "***Pressure Solver FAILED\n";
        ss << "Pressure Solver Iterations: " << rep.iterations << "\nEstimated Error: " << rep.residual;
        if (rc == FLIPMI_OK && rep.iterations == 0) ss.str("Pressure Solver Iterations: 0\nEstimated Error: 0.0");
        _status = ss.str();
        return rc == FLIPMI_OK;
    }
    std::string getSolverStatus() { return _status; }

private:
    std::string _status;
};

// ---------------------------------------------------------------------------
// VelocityAdvector (velocityadvector.h:62-75)
// ---------------------------------------------------------------------------

struct VelocityAdvectorParameters {
    FragmentedVector<MarkerParticle> *particles;
    MACVelocityField *vfield;
    ValidVelocityComponentGrid *validVelocities;
    double particleRadius = 1.0;
};

class VelocityAdvector {
public:
    VelocityAdvector() {}

    // Splats particle velocities onto the faces of vfield and sets
    // validVelocities.  Every face is written: faces no particle reaches get
    // 0 and invalid, which is what the reference leaves on the cleared grids
    // its caller passes (fluidsimulation.cpp:3259-3260).
    void advect(VelocityAdvectorParameters params) {
        int I, J, K;
        params.vfield->getGridDimensions(&I, &J, &K);
        FragmentedVector<MarkerParticle> &ps = *params.particles;
        size_t n = ps.size();
        std::vector<float> pos(3 * n), vel(3 * n);
        for (size_t q = 0; q < n; q++) {
            for (int c = 0; c < 3; c++) {
                pos[3 * q + c] = ps[q].position[c];
                vel[3 * q + c] = ps[q].velocity[c];
            }
        }
        detail::check(
            flipmi_p2g(detail::ctx(), I, J, K, params.vfield->getGridCellSize(), pos.data(), vel.data(), (int64_t)n,
                       params.particleRadius, params.vfield->getRawArrayU(), params.vfield->getRawArrayV(),
                       params.vfield->getRawArrayW(),
                       reinterpret_cast<uint8_t *>(params.validVelocities->validU.getRawArray()),
                       reinterpret_cast<uint8_t *>(params.validVelocities->validV.getRawArray()),
                       reinterpret_cast<uint8_t *>(params.validVelocities->validW.getRawArray())),
            "VelocityAdvector::advect");
    }
};

// ---------------------------------------------------------------------------
// ParticleAdvector (particleadvector.h:37-79)
// ---------------------------------------------------------------------------

class ParticleAdvector {
public:
    ParticleAdvector() {}

    // The reference probes OpenCL here; this checks for a HIP device.
    bool initialize() {
        try {
            detail::ctx();
            _initialized = true;
        } catch (const std::exception &e) {
            _error = e.what();
            _initialized = false;
        }
        return _initialized;
    }
    bool isInitialized() { return _initialized; }
    std::string getInitializationErrorMessage() { return _error; }
    std::string getDeviceInfo() { return "HIP device " + std::to_string(detail::device_index()); }
    std::string getKernelInfo() { return "flipmi advect_particles"; }
    // The OpenCL switches of the reference: the device path is the only one.
    void disableOpenCL() {}
    void enableOpenCL() {}
    bool isOpenCLEnabled() { return false; }
    int getKernelWorkLoadSize() { return _workload; }
    void setKernelWorkLoadSize(int n) { _workload = n; }

    void advectParticlesRK4(std::vector<vmath::vec3> &particles, MACVelocityField *vfield, double dt,
                            std::vector<vmath::vec3> &output) {
        _advect(particles, vfield, dt, 4, output);
    }
    void advectParticlesRK3(std::vector<vmath::vec3> &particles, MACVelocityField *vfield, double dt,
                            std::vector<vmath::vec3> &output) {
        _advect(particles, vfield, dt, 3, output);
    }
    void advectParticlesRK2(std::vector<vmath::vec3> &particles, MACVelocityField *vfield, double dt,
                            std::vector<vmath::vec3> &output) {
        _advect(particles, vfield, dt, 2, output);
    }
    void advectParticlesRK1(std::vector<vmath::vec3> &particles, MACVelocityField *vfield, double dt,
                            std::vector<vmath::vec3> &output) {
        _advect(particles, vfield, dt, 1, output);
    }
    void trilinearInterpolate(std::vector<vmath::vec3> &particles, MACVelocityField *vfield,
                              std::vector<vmath::vec3> &output) {
        int I, J, K;
        vfield->getGridDimensions(&I, &J, &K);
        output.resize(particles.size());
        if (particles.empty()) return;
        detail::check(flipmi_trilinear_velocity(detail::ctx(), I, J, K, vfield->getGridCellSize(),
                                                vfield->getRawArrayU(), vfield->getRawArrayV(),
                                                vfield->getRawArrayW(), &particles[0].x, (int64_t)particles.size(),
                                                &output[0].x),
                      "ParticleAdvector::trilinearInterpolate");
    }
    // Overwrites particles with the sampled velocities.
    void trilinearInterpolate(std::vector<vmath::vec3> &particles, MACVelocityField *vfield) {
        std::vector<vmath::vec3> out;
        trilinearInterpolate(particles, vfield, out);
        particles.swap(out);
    }

private:
    void _advect(std::vector<vmath::vec3> &particles, MACVelocityField *vfield, double dt, int order,
                 std::vector<vmath::vec3> &output) {
        int I, J, K;
        vfield->getGridDimensions(&I, &J, &K);
        output.resize(particles.size());
        if (particles.empty()) return;
        detail::check(flipmi_advect_particles(detail::ctx(), I, J, K, vfield->getGridCellSize(),
                                              vfield->getRawArrayU(), vfield->getRawArrayV(),
                                              vfield->getRawArrayW(), &particles[0].x, (int64_t)particles.size(),
                                              dt, order, &output[0].x),
                      "ParticleAdvector::advectParticles");
    }
    bool _initialized = false;
    std::string _error;
    int _workload = 0;
};

// ---------------------------------------------------------------------------
// FluidSimulation (fluidsimulation.h:60-153, 304-1126)
// ---------------------------------------------------------------------------

struct FluidSimulationMeshStats {
    int enabled = 0;
    int vertices = -1;
    int triangles = -1;
    unsigned int bytes = 0;
};

struct FluidSimulationTimingStats {
    double total = 0.0;
    double mesh = 0.0;
    double advection = 0.0;
    double particles = 0.0;
    double pressure = 0.0;
    double diffuse = 0.0;
    double viscosity = 0.0;
    double objects = 0.0;
};

struct FluidSimulationFrameStats {
    int frame = 0;
    int substeps = 0;
    double deltaTime = 0.0;
    int fluidParticles = 0;
    int diffuseParticles = 0;
    FluidSimulationMeshStats surface, preview, surfaceblur, foam, bubble, spray, dust, foamblur, bubbleblur,
        sprayblur, dustblur, particles, obstacle, forcefield;
    FluidSimulationTimingStats timing;
};

struct FluidSimulationMarkerParticleData {
    int size = 0;
    char *positions;
    char *velocities;
};

// trianglemesh.h:28-84: vertex list and index triples.
struct Triangle {
    int tri[3] = {0, 0, 0};
    Triangle() {}
    Triangle(int p1, int p2, int p3) { tri[0] = p1; tri[1] = p2; tri[2] = p3; }
};
struct TriangleMesh {
    std::vector<vmath::vec3> vertices;
    std::vector<Triangle> triangles;
    int numVertices() { return (int)vertices.size(); }
    int numTriangles() { return (int)triangles.size(); }
    void clear() { vertices.clear(); triangles.clear(); }
};

// aabb.h:31-80: float corner, double extents.
struct AABB {
    vmath::vec3 position;
    double width = 0.0, height = 0.0, depth = 0.0;
    AABB() {}
    AABB(double x, double y, double z, double w, double h, double d)
        : position(x, y, z), width(w), height(h), depth(d) {}
    AABB(vmath::vec3 p, double w, double h, double d) : position(p), width(w), height(h), depth(d) {}
    vmath::vec3 getMinPoint() { return position; }
    vmath::vec3 getMaxPoint() { return position + vmath::vec3(width, height, depth); }
};

// particlemesher.h:36-48
struct ParticleMesherParameters {
    int isize = 0;
    int jsize = 0;
    int ksize = 0;
    double dx = 0.0;
    int subdivisions = 1;
    int computechunks = 1;
    double radius = 0.0;
    bool isPreviewMesherEnabled = false;
    double previewdx = 0.0;
    std::vector<vmath::vec3> *particles = nullptr;
    class MeshLevelSet *solidSDF = nullptr;
};

namespace detail {
inline void copy_mesh(flipmi_mesh *m, TriangleMesh &out) {
    int64_t nv = 0, nt = 0;
    detail::check(flipmi_mesh_counts(m, &nv, &nt), "mesh");
    out.vertices.assign((size_t)nv, vmath::vec3());
    out.triangles.assign((size_t)nt, Triangle());
    detail::check(flipmi_mesh_copy(m, reinterpret_cast<float *>(out.vertices.data()),
                                   reinterpret_cast<int32_t *>(out.triangles.data())),
                  "mesh");
}
}  // namespace detail
static_assert(sizeof(Triangle) == 12, "Triangle must be 3 packed ints");

// particlemesher.h:50-60, particlemesher.cpp:36-55: the surface of the
// particles over the solid SDF, before the simulation's post-processing
// (one compute chunk; no preview mesh).  Positions in grid coordinates,
// inside the grid.
class ParticleMesher {
public:
    ParticleMesher() {}
    TriangleMesh meshParticles(ParticleMesherParameters params) {
        if (params.computechunks != 1) throw std::runtime_error("flipmi: one polygonizer slice only");
        if (params.isPreviewMesherEnabled) throw std::runtime_error("flipmi: the preview mesh is not built");
        if (!params.particles || !params.solidSDF) throw std::runtime_error("flipmi: particles and a solid SDF are required");
        flipmi_surface_desc d;
        flipmi_surface_desc_default(&d);
        d.subdivisions = params.subdivisions;
        d.smoothing_iterations = 0;
        flipmi_mesh *m = nullptr;
        detail::check(flipmi_mesh_particles(detail::ctx(), params.isize, params.jsize, params.ksize, params.dx,
                                            reinterpret_cast<const float *>(params.particles->data()),
                                            (int64_t)params.particles->size(),
                                            params.solidSDF->getPhiArray3d()->getRawArray(), &d, params.radius,
                                            nullptr, 1.0, &m),
                      "ParticleMesher::meshParticles");
        TriangleMesh out;
        try {
            detail::copy_mesh(m, out);
        } catch (...) {
            flipmi_mesh_destroy(m);
            throw;
        }
        flipmi_mesh_destroy(m);
        return out;
    }
};

// meshobject.h: a static mesh obstacle (updateMeshStatic).  enable() and
// disable() after it joins an initialized simulation rebuild the solid at the
// next substep.
class MeshObject {
public:
    MeshObject(int isize, int jsize, int ksize, double dx) { (void)isize; (void)jsize; (void)ksize; (void)dx; }
    void updateMeshStatic(TriangleMesh mesh) {
        if (_scene) throw std::runtime_error("flipmi: the obstacle mesh is fixed once it is on the device");
        _mesh = mesh;
    }
    TriangleMesh getMesh() { return _mesh; }
    void enable() { _enabled = true; _push(); }
    void disable() { _enabled = false; _push(); }
    bool isEnabled() { return _enabled; }
    void inverse() { _inversed = !_inversed; _push(); }  // meshobject.cpp:293-295
    bool isInversed() { return _inversed; }

private:
    friend class FluidSimulation;
    void _attach(flipmi_scene *scene) {
        std::vector<float> v;
        std::vector<int> t;
        for (const vmath::vec3 &p : _mesh.vertices) { v.push_back(p.x); v.push_back(p.y); v.push_back(p.z); }
        for (const Triangle &q : _mesh.triangles) t.insert(t.end(), q.tri, q.tri + 3);
        detail::check(flipmi_scene_add_obstacle(scene, v.data(), (int)_mesh.vertices.size(), t.data(),
                                                (int)_mesh.triangles.size(), &_id),
                      "FluidSimulation::addMeshObstacle");
        _scene = scene;
        _push();
    }
    void _push() {
        if (!_scene) return;
        detail::check(flipmi_scene_set_obstacle(_scene, _id, _enabled), "MeshObject");
        detail::check(flipmi_scene_set_obstacle_inversed(_scene, _id, _inversed), "MeshObject");
    }
    TriangleMesh _mesh;
    bool _enabled = true, _inversed = false;
    flipmi_scene *_scene = nullptr;
    int _id = -1;
};

// meshfluidsource.h: a mesh inflow source with a static mesh
// (updateMeshStatic).  Settings made after the source joins an initialized
// simulation reach the device at once.
class MeshFluidSource {
public:
    MeshFluidSource(int isize, int jsize, int ksize, double dx) { (void)isize; (void)jsize; (void)ksize; (void)dx; }
    void updateMeshStatic(TriangleMesh mesh) {
        if (_scene) throw std::runtime_error("flipmi: the source mesh is fixed once the source is on the device");
        _mesh = mesh;
    }
    void setInflow() { _outflow = false; _push(); }
    void setOutflow() { _outflow = true; _push(); }
    bool isInflow() { return !_outflow; }
    bool isOutflow() { return _outflow; }
    void outflowInverse() { _inversed = !_inversed; _push(); }  // meshfluidsource.cpp:158-160
    bool isOutflowInversed() { return _inversed; }
    void setVelocity(vmath::vec3 v) {
        if (_scene) throw std::runtime_error("flipmi: the source velocity is fixed once the source is on the device");
        _velocity = v;
    }
    vmath::vec3 getVelocity() { return _velocity; }
    void enable() { _enabled = true; _push(); }
    void disable() { _enabled = false; _push(); }
    bool isEnabled() { return _enabled; }
    void enableConstrainedFluidVelocity() { _constrain = true; _push(); }
    void disableConstrainedFluidVelocity() { _constrain = false; _push(); }
    bool isConstrainedFluidVelocityEnabled() { return _constrain; }
    void setSubstepEmissions(int n) { _emissions = n < 0 ? 0 : n; _push(); }  // meshfluidsource.cpp:63-68
    int getSubstepEmissions() { return _emissions; }

private:
    friend class FluidSimulation;
    void _attach(flipmi_scene *scene) {
        std::vector<float> v;
        std::vector<int> t;
        for (const vmath::vec3 &p : _mesh.vertices) { v.push_back(p.x); v.push_back(p.y); v.push_back(p.z); }
        for (const Triangle &q : _mesh.triangles) t.insert(t.end(), q.tri, q.tri + 3);
        const double vel[3] = {_velocity.x, _velocity.y, _velocity.z};
        detail::check(flipmi_scene_add_inflow_source(scene, v.data(), (int)_mesh.vertices.size(), t.data(),
                                                     (int)_mesh.triangles.size(), vel, &_id),
                      "FluidSimulation::addMeshFluidSource");
        _scene = scene;
        _push();
    }
    void _push() {
        if (!_scene) return;
        detail::check(flipmi_scene_set_inflow_source(_scene, _id, _enabled, _constrain, _emissions),
                      "MeshFluidSource");
        detail::check(flipmi_scene_set_outflow_source(_scene, _id, _outflow, _inversed), "MeshFluidSource");
    }
    TriangleMesh _mesh;
    vmath::vec3 _velocity;
    bool _enabled = true, _constrain = true;
    int _emissions = 1;
    bool _outflow = false, _inversed = false;
    flipmi_scene *_scene = nullptr;
    int _id = -1;
};

class FluidSimulation {
public:
    FluidSimulation(int isize, int jsize, int ksize, double dx) : _isize(isize), _jsize(jsize), _ksize(ksize), _dx(dx) {
        flipmi_scene_desc_default(&_desc);
        _desc.isize = isize;
        _desc.jsize = jsize;
        _desc.ksize = ksize;
        _desc.dx = dx;
        flipmi_surface_desc_default(&_surface);
    }
    FluidSimulation(const FluidSimulation &) = delete;
    FluidSimulation &operator=(const FluidSimulation &) = delete;
    virtual ~FluidSimulation() {
        if (_scene) flipmi_scene_destroy(_scene);
    }

    // fluidsimulation.cpp:82-88: builds the device scene from the settings
    // and the particles loaded so far.
    virtual void initialize() {
        if (_initialized) return;
        flipmi_ctx *c = detail::ctx();
        _desc.max_particles = _maxParticles;
        detail::check(flipmi_scene_create(c, &_desc, &_scene), "FluidSimulation::initialize");
        detail::check(flipmi_scene_set_domain(_scene, _domainOffset.x, _domainOffset.y, _domainOffset.z, _domainScale),
                      "FluidSimulation::initialize");
        int64_t n = (int64_t)(_loadPos.size() / 3);
        if (n > 0)
            detail::check(flipmi_scene_load_particles(_scene, _loadPos.data(), _loadVel.data(), n),
                          "FluidSimulation::loadMarkerParticleData");
        _loadPos.clear();
        _loadVel.clear();
        detail::check(flipmi_scene_set_marker_jitter(_scene, _jitterFactor, _jitterSurface, _jitterSeed),
                      "FluidSimulation::initialize");
        detail::check(flipmi_scene_set_surface(_scene, &_surface), "FluidSimulation::initialize");
        for (MeshObject *ob : _obstacles) ob->_attach(_scene);
        for (MeshFluidSource *src : _sources) src->_attach(_scene);
        detail::check(flipmi_scene_initialize(_scene), "FluidSimulation::initialize");
        _initialized = true;
    }

    // fluidsimulation.cpp:1994-2008; the obstacle must outlive the simulation.
    void addMeshObstacle(MeshObject *obstacle) {
        for (MeshObject *o : _obstacles)
            if (o == obstacle) throw std::runtime_error("Error: mesh obstacle has already been added.");
        _obstacles.push_back(obstacle);
        if (_scene) obstacle->_attach(_scene);
    }
    // fluidsimulation.cpp:1953-1965; the source must outlive the simulation.
    void addMeshFluidSource(MeshFluidSource *source) {
        for (MeshFluidSource *s : _sources)
            if (s == source) throw std::runtime_error("Error: Mesh fluid source has already been added.");
        _sources.push_back(source);
        if (_scene) source->_attach(_scene);
    }
    // fluidsimulation.cpp:185-210 for the particles inflow sources add; the
    // seed keys this build's jitter draws (the reference calls rand()).
    void setMarkerParticleJitterFactor(double jit) {
        if (jit < 0.0) throw std::domain_error("Error: marker particle jitter must be greater than or equal to 0.");
        _jitterFactor = jit;
        _pushJitter();
    }
    double getMarkerParticleJitterFactor() { return _jitterFactor; }
    void enableJitterSurfaceMarkerParticles() { _jitterSurface = 1; _pushJitter(); }
    void disableJitterSurfaceMarkerParticles() { _jitterSurface = 0; _pushJitter(); }
    bool isJitterSurfaceMarkerParticlesEnabled() { return _jitterSurface != 0; }
    void setMarkerParticleJitterSeed(uint64_t seed) { _jitterSeed = seed; _pushJitter(); }
    bool isInitialized() { return _initialized; }

    // fluidsimulation.cpp:5755-5846
    void update(double dt) {
        if (!_initialized) throw std::runtime_error("Error: FluidSimulation must be initialized before update.");
        if (dt < 0.0) throw std::domain_error("Error: delta time must be greater than or equal to 0.");
        auto t0 = std::chrono::steady_clock::now();
        flipmi_frame_stats st;
        detail::check(flipmi_scene_update(_scene, dt, &st), "FluidSimulation::update");
        _isomeshStale = true;
        auto t1 = std::chrono::steady_clock::now();
        _stats = FluidSimulationFrameStats();
        _stats.frame = _currentFrame;
        _stats.substeps = st.substeps;
        _stats.deltaTime = dt;
        _stats.fluidParticles = (int)st.num_particles;
        _stats.timing.total = std::chrono::duration<double>(t1 - t0).count();
        _lastPcgIterations = st.pcg_iterations;
        _lastPcgResidual = st.pcg_residual;
        _lastPressureOk = st.pressure_ok != 0;
        _velocityStale = true;
        _currentFrame++;
    }

    // fluidsimulation.cpp:303-330: particles are loaded and read back in
    // domain coordinates (p * scale + offset).
    vmath::vec3 getDomainOffset() { return _domainOffset; }
    void setDomainOffset(double x, double y, double z) { setDomainOffset(vmath::vec3(x, y, z)); }
    void setDomainOffset(vmath::vec3 offset) {
        _domainOffset = offset;
        if (_scene) detail::check(flipmi_scene_set_domain(_scene, offset.x, offset.y, offset.z, _domainScale),
                                  "FluidSimulation::setDomainOffset");
    }
    double getDomainScale() { return _domainScale; }
    void setDomainScale(double scale) {
        _domainScale = scale;
        if (_scene) detail::check(flipmi_scene_set_domain(_scene, _domainOffset.x, _domainOffset.y, _domainOffset.z, scale),
                                  "FluidSimulation::setDomainScale");
    }

    int getCurrentFrame() { return _currentFrame; }
    double getCellSize() { return _dx; }
    void getGridDimensions(int *i, int *j, int *k) {
        *i = _isize;
        *j = _jsize;
        *k = _ksize;
    }
    void getSimulationDimensions(double *width, double *height, double *depth) {
        *width = _isize * _dx;
        *height = _jsize * _dx;
        *depth = _ksize * _dx;
    }
    double getSimulationWidth() { return _isize * _dx; }
    double getSimulationHeight() { return _jsize * _dx; }
    double getSimulationDepth() { return _ksize * _dx; }

    // Settings: fixed when initialize() builds the device scene.
    void setMaxThreadCount(int n) {
        if (n < 1) throw std::domain_error("Error: thread count must be greater than or equal to 1.");
        _maxThreads = n;  // host threads play no part in the device step
    }
    int getMaxThreadCount() { return _maxThreads; }
    void addBodyForce(double fx, double fy, double fz) {
        _settable("addBodyForce");
        _desc.body_force[0] += fx;
        _desc.body_force[1] += fy;
        _desc.body_force[2] += fz;
    }
    void addBodyForce(vmath::vec3 f) { addBodyForce(f.x, f.y, f.z); }
    void setBoundaryFriction(double f) {
        if (f < 0.0 || f > 1.0) throw std::domain_error("Error: boundary friction must be in range [0.0, 1.0].");
        if (f != 0.0) throw std::domain_error("flipmi: only frictionless boundaries (0.0) are available");
    }
    double getBoundaryFriction() { return 0.0; }
    void setCFLConditionNumber(int n) {
        if (n < 1) throw std::domain_error("Error: CFL must be greater than or equal to 1.");
        _settable("setCFLConditionNumber");
        _desc.cfl = n;
    }
    int getCFLConditionNumber() { return (int)_desc.cfl; }
    void setMinTimeStepsPerFrame(int n) {
        if (n < 1) throw std::domain_error("Error: min step count must be greater than or equal to 1.");
        _settable("setMinTimeStepsPerFrame");
        _desc.min_substeps = n;
    }
    void setMaxTimeStepsPerFrame(int n) {
        if (n < 1) throw std::domain_error("Error: max step count must be greater than or equal to 1.");
        _settable("setMaxTimeStepsPerFrame");
        _desc.max_substeps = n;
    }
    void setPICFLIPRatio(double r) {
        if (r < 0.0 || r > 1.0) throw std::domain_error("Error: PICFLIP ratio must be in range [0.0, 1.0].");
        _settable("setPICFLIPRatio");
        _desc.pic_ratio = r;
    }
    double getPICFLIPRatio() { return _desc.pic_ratio; }
    // Nothing is printed on this path.
    void disableConsoleOutput() {}
    void enableConsoleOutput() {}

    // Surface reconstruction (fluidsimulation.cpp:168-540): on by default, as
    // in the reference; the mesh of the particles after the first substep of
    // each frame, read with getIsomesh().  One polygonizer slice and a
    // minimum polyhedron size of 0 are built; other values throw at the
    // setter (or at initialize()).
    void enableSurfaceReconstruction() { _surface.enabled = 1; _pushSurface(); }
    void disableSurfaceReconstruction() { _surface.enabled = 0; _pushSurface(); }
    bool isSurfaceReconstructionEnabled() { return _surface.enabled != 0; }
    int getSurfaceSubdivisionLevel() { return _surface.subdivisions; }
    void setSurfaceSubdivisionLevel(int n) {
        if (n < 1) throw std::domain_error("Error: subdivision level must be greater than or equal to 1.");
        _surface.subdivisions = n;
        _pushSurface();
    }
    int getNumPolygonizerSlices() { return _surface.polygonizer_slices; }
    void setNumPolygonizerSlices(int n) {
        if (n < 1) throw std::domain_error("Error: number of polygonizer slices must be greater than or equal to 1.");
        _surface.polygonizer_slices = n;
        _pushSurface();
    }
    double getSurfaceSmoothingValue() { return _surface.smoothing_value; }
    void setSurfaceSmoothingValue(double s) { _surface.smoothing_value = s; _pushSurface(); }
    int getSurfaceSmoothingIterations() { return _surface.smoothing_iterations; }
    void setSurfaceSmoothingIterations(int n) {
        if (n < 0) throw std::domain_error("Error: smoothing iterations must be greater than or equal to 0.");
        _surface.smoothing_iterations = n;
        _pushSurface();
    }
    int getMinPolyhedronTriangleCount() { return _surface.min_polyhedron_triangles; }
    void setMinPolyhedronTriangleCount(int n) {
        if (n < 0) throw std::domain_error("Error: min triangle count must be greater than or equal to 0.");
        _surface.min_polyhedron_triangles = n;
        _pushSurface();
    }
    double getMarkerParticleScale() { return _surface.particle_scale; }
    void setMarkerParticleScale(double s) {
        if (s < 0.0) throw std::domain_error("Error: marker particle scale must be greater than or equal to 0.");
        _surface.particle_scale = s;
        _pushSurface();
    }
    double getObstacleMeshingOffset() { return _surface.obstacle_meshing_offset; }
    void setObstacleMeshingOffset(double s) { _surface.obstacle_meshing_offset = s; _pushSurface(); }
    void enableRemoveSurfaceNearDomain() { _surface.remove_near_domain = 1; _pushSurface(); }
    void disableRemoveSurfaceNearDomain() { _surface.remove_near_domain = 0; _pushSurface(); }
    bool isRemoveSurfaceNearDomainEnabled() { return _surface.remove_near_domain != 0; }
    int getRemoveSurfaceNearDomainDistance() { return _surface.remove_near_domain_distance; }
    void setRemoveSurfaceNearDomainDistance(int n) {
        if (n < 0) throw std::domain_error("Error: distance must be greater than or equal to 0.");
        _surface.remove_near_domain_distance = n;
        _pushSurface();
    }
    void enableInvertedContactNormals() { _surface.invert_contact_normals = 1; _pushSurface(); }
    void disableInvertedContactNormals() { _surface.invert_contact_normals = 0; _pushSurface(); }
    bool isInvertedContactNormalsEnabled() { return _surface.invert_contact_normals != 0; }
    // fluidsimulation.cpp:2543-2546: the last surface mesh, domain coordinates.
    TriangleMesh &getIsomesh() {
        if (_scene && _isomeshStale) {
            int64_t nv = 0, nt = 0;
            detail::check(flipmi_scene_get_isomesh(_scene, &nv, &nt, nullptr, nullptr), "FluidSimulation::getIsomesh");
            _isomesh.vertices.assign((size_t)nv, vmath::vec3());
            _isomesh.triangles.assign((size_t)nt, Triangle());
            detail::check(flipmi_scene_get_isomesh(_scene, &nv, &nt, reinterpret_cast<float *>(_isomesh.vertices.data()),
                                                   reinterpret_cast<int32_t *>(_isomesh.triangles.data())),
                          "FluidSimulation::getIsomesh");
            _isomeshStale = false;
        }
        return _isomesh;
    }

    // Not in the reference: PCG arithmetic (FLIPMI_PCG_FP64 = the reference's).
    void setPressureSolverPrecision(int precision) {
        _settable("setPressureSolverPrecision");
        _desc.pcg_precision = precision;
    }
    // Not in the reference: particle capacity of the device scene (0 = size
    // of the loaded set).
    void setMaxMarkerParticles(int64_t n) {
        _settable("setMaxMarkerParticles");
        _maxParticles = n;
    }

    // fluidsimulation.cpp:2488-2509: particles queue up until initialize().
    void loadMarkerParticleData(FluidSimulationMarkerParticleData data) {
        if (data.size == 0) return;
        if (_initialized) throw std::runtime_error("flipmi: marker particles must be loaded before initialize()");
        const float *p = reinterpret_cast<const float *>(data.positions);
        const float *v = reinterpret_cast<const float *>(data.velocities);
        _loadPos.insert(_loadPos.end(), p, p + 3 * (size_t)data.size);
        if (v) _loadVel.insert(_loadVel.end(), v, v + 3 * (size_t)data.size);
        else _loadVel.insert(_loadVel.end(), 3 * (size_t)data.size, 0.0f);
    }

    // Loaded particles join at initialize(); none are counted before.
    unsigned int getNumMarkerParticles() {
        if (!_initialized) return 0;
        int64_t n = 0;
        detail::check(flipmi_scene_num_particles(_scene, &n), "FluidSimulation::getNumMarkerParticles");
        return (unsigned int)n;
    }
    std::vector<MarkerParticle> getMarkerParticles() {
        std::vector<vmath::vec3> p = getMarkerParticlePositions(), v = getMarkerParticleVelocities();
        std::vector<MarkerParticle> out(p.size());
        for (size_t q = 0; q < p.size(); q++) out[q] = MarkerParticle(p[q], v[q]);
        return out;
    }
    std::vector<vmath::vec3> getMarkerParticlePositions() {
        std::vector<vmath::vec3> out(getNumMarkerParticles());
        if (!out.empty()) getMarkerParticlePositionData(reinterpret_cast<char *>(out.data()));
        return out;
    }
    std::vector<vmath::vec3> getMarkerParticleVelocities() {
        std::vector<vmath::vec3> out(getNumMarkerParticles());
        if (!out.empty()) getMarkerParticleVelocityData(reinterpret_cast<char *>(out.data()));
        return out;
    }
    // fluidsimulation.cpp:2408-2420: vec3 blobs, 12 bytes per particle.
    void getMarkerParticlePositionData(char *data) {
        _requireScene();
        detail::check(flipmi_scene_get_particles(_scene, reinterpret_cast<float *>(data), nullptr),
                      "FluidSimulation::getMarkerParticlePositionData");
    }
    void getMarkerParticleVelocityData(char *data) {
        _requireScene();
        detail::check(flipmi_scene_get_particles(_scene, nullptr, reinterpret_cast<float *>(data)),
                      "FluidSimulation::getMarkerParticleVelocityData");
    }
    // fluidsimulation.cpp:2314-2338: particles [start_idx, end_idx).
    void getMarkerParticlePositionDataRange(int start_idx, int end_idx, char *data) {
        _range(start_idx, end_idx, reinterpret_cast<float *>(data), nullptr);
    }
    void getMarkerParticleVelocityDataRange(int start_idx, int end_idx, char *data) {
        _range(start_idx, end_idx, nullptr, reinterpret_cast<float *>(data));
    }
    unsigned int getMarkerParticlePositionDataSize() { return getNumMarkerParticles() * 3 * sizeof(float); }
    unsigned int getMarkerParticleVelocityDataSize() { return getNumMarkerParticles() * 3 * sizeof(float); }

    // fluidsimulation.cpp:2242: a host copy of the device MAC grid, refreshed
    // on each call after an update.
    MACVelocityField *getVelocityField() {
        _requireScene();
        if (!_mac || _velocityStale) {
            if (!_mac) _mac.reset(new MACVelocityField(_isize, _jsize, _ksize, _dx));
            detail::check(flipmi_scene_get_grid(_scene, FLIPMI_GRID_U, _mac->getRawArrayU()), "getVelocityField");
            detail::check(flipmi_scene_get_grid(_scene, FLIPMI_GRID_V, _mac->getRawArrayV()), "getVelocityField");
            detail::check(flipmi_scene_get_grid(_scene, FLIPMI_GRID_W, _mac->getRawArrayW()), "getVelocityField");
            _velocityStale = false;
        }
        return _mac.get();
    }

    FluidSimulationFrameStats getFrameStatsData() { return _stats; }

    // Not in the reference: the last substep's pressure solve.
    int getLastPressureSolverIterations() { return _lastPcgIterations; }
    double getLastPressureSolverError() { return _lastPcgResidual; }
    bool getLastPressureSolveSucceeded() { return _lastPressureOk; }

private:
    void _settable(const char *what) {
        if (_initialized) throw std::runtime_error(std::string("flipmi: ") + what + " must be called before initialize()");
    }
    void _requireScene() {
        if (!_initialized) throw std::runtime_error("Error: FluidSimulation must be initialized first.");
    }
    void _pushSurface() {
        if (_scene) detail::check(flipmi_scene_set_surface(_scene, &_surface), "FluidSimulation surface setting");
    }
    void _pushJitter() {
        if (_scene)
            detail::check(flipmi_scene_set_marker_jitter(_scene, _jitterFactor, _jitterSurface, _jitterSeed),
                          "FluidSimulation::setMarkerParticleJitterFactor");
    }
    void _range(int start, int end, float *pos, float *vel) {
        if (start < 0 || end > (int)getNumMarkerParticles() || start > end)
            throw std::domain_error("Error: invalid range.");
        if (start == end) return;
        detail::check(flipmi_scene_get_particles_range(_scene, start, end, pos, vel),
                      "FluidSimulation::getMarkerParticleDataRange");
    }

    int _isize, _jsize, _ksize;
    double _dx;
    flipmi_scene_desc _desc;
    flipmi_scene *_scene = nullptr;
    bool _initialized = false;
    int _currentFrame = 0;
    int _maxThreads = 1;
    int64_t _maxParticles = 0;
    vmath::vec3 _domainOffset;
    double _domainScale = 1.0;
    std::vector<float> _loadPos, _loadVel;
    FluidSimulationFrameStats _stats;
    std::unique_ptr<MACVelocityField> _mac;
    bool _velocityStale = true;
    int _lastPcgIterations = 0;
    double _lastPcgResidual = 0.0;
    bool _lastPressureOk = true;
    std::vector<MeshFluidSource *> _sources;
    std::vector<MeshObject *> _obstacles;
    flipmi_surface_desc _surface;
    TriangleMesh _isomesh;
    bool _isomeshStale = true;
    double _jitterFactor = 1.0;  // fluidsimulation.h:1553
    int _jitterSurface = 0;
    uint64_t _jitterSeed = 0;
};

}  // namespace flipmi

#endif  // FLIPMI_ENGINE_HPP
