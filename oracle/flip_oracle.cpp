// CPU oracle for the FLIP step -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// Single-threaded C++17 restatement of the reference FLIP engine step
// (jklae/FLIPEngine3D, src/engine).  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this library, as a checker or as the
// timed CPU baseline.  It must be built with -ffp-contract=off and without
// -ffast-math: every float/double operation below mirrors the reference's
// types and evaluation order so results are bit-identical wherever the
// reference's result does not depend on thread scheduling.
//
// All citations are paths relative to the reference repository root.
//
// Parity status (pinned by tests/test_oracle_vs_ref.py against the compiled
// reference in oracle/_ref, and by closed-form checks in
// tests/test_oracle_math.py):
//   domain solid SDF + weights ........ bit-exact vs reference
//   liquid SDF / labels / counts ...... bit-exact vs reference
//   P2G (velocity + valid) ............ bit-exact vs reference
//   extrapolation ..................... bit-exact vs reference
//   pressure RHS / MIC(0) / apply ..... bit-exact vs reference
//   PCG (ref-exact mode) .............. bit-exact vs reference at 1 thread
//   PIC/FLIP, RK1-4, collision, removal bit-exact vs reference
//   whole update(dt) .................. bit-exact vs reference at 1 thread
//   surface mesh (getIsomesh) ......... bit-exact vs reference (one polygonizer
//                                       slice, minimum polyhedron size 0)

#include "flip_oracle.h"

#include "../flipengine3d_amd/csrc/mc_tables.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Float 3-vector with the reference's vmath::vec3 semantics
// (src/engine/vmath.h:38-97, vmath.cpp:33-123): every component is float and
// scalar products take a float scalar; division multiplies by a float
// reciprocal formed in double.
// ---------------------------------------------------------------------------
struct V3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    V3() {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
};
static inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 operator*(float s, V3 v) { return V3(v.x * s, v.y * s, v.z * s); }
static inline V3 operator*(V3 v, float s) { return V3(v.x * s, v.y * s, v.z * s); }
static inline V3 operator/(V3 v, float s) {
    float inv = 1.0 / s;
    return V3(v.x * inv, v.y * inv, v.z * inv);
}
static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float length(V3 v) { return (float)std::sqrt((double)dot(v, v)); }
static inline V3 normalize(V3 v) { return v / length(v); }

struct Idx {
    int i, j, k;
};

// grid3d.h:52-57 -- floor of position times 1/dx, in double.
static inline Idx pos_to_index(double x, double y, double z, double dx) {
    double inv = 1.0 / dx;
    return Idx{(int)std::floor(x * inv), (int)std::floor(y * inv), (int)std::floor(z * inv)};
}
static inline Idx pos_to_index(V3 p, double dx) { return pos_to_index(p.x, p.y, p.z, dx); }
// grid3d.h:86-92 -- ((float)i)*dx rounded to float.
static inline V3 index_to_pos(int i, int j, int k, double dx) {
    return V3((float)((double)(float)i * dx), (float)((double)(float)j * dx),
              (float)((double)(float)k * dx));
}
// grid3d.h:106-109 -- (float)i*dx + 0.5*dx rounded to float.
static inline V3 index_to_center(int i, int j, int k, double dx) {
    double hw = 0.5 * dx;
    return V3((float)((double)(float)i * dx + hw), (float)((double)(float)j * dx + hw),
              (float)((double)(float)k * dx + hw));
}

template <class T>
struct Grid {
    int ni = 0, nj = 0, nk = 0;
    std::vector<T> a;
    Grid() {}
    Grid(int i, int j, int k, T v) : ni(i), nj(j), nk(k), a((size_t)i * j * k, v) {}
    bool in(int i, int j, int k) const {
        return i >= 0 && j >= 0 && k >= 0 && i < ni && j < nj && k < nk;
    }
    size_t id(int i, int j, int k) const { return (size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * k); }
    T &operator()(int i, int j, int k) { return a[id(i, j, k)]; }
    const T &operator()(int i, int j, int k) const { return a[id(i, j, k)]; }
    void fill(T v) { std::fill(a.begin(), a.end(), v); }
};
typedef Grid<float> GridF;
typedef Grid<uint8_t> GridB;

template <class T>
static T clampv(T v, T lo, T hi) { return std::max(lo, std::min(v, hi)); }

// ---------------------------------------------------------------------------
// Axis-aligned box with the reference's AABB semantics (aabb.cpp:32-133,
// 489-518): float min corner, double extents.
// ---------------------------------------------------------------------------
struct Box {
    V3 pos;
    double w = 0, h = 0, d = 0;
    void expand(double v) {  // aabb.cpp:122-128
        double hh = 0.5 * v;
        pos = pos - V3((float)hh, (float)hh, (float)hh);
        w += v; h += v; d += v;
    }
    bool inside(V3 p) const {  // aabb.cpp:130-133
        return p.x >= pos.x && p.y >= pos.y && p.z >= pos.z && p.x < pos.x + w &&
               p.y < pos.y + h && p.z < pos.z + d;
    }
    V3 maxp() const { return pos + V3((float)w, (float)h, (float)d); }  // aabb.cpp:493-495
    V3 nearest_inside(V3 p) const {                                      // aabb.cpp:501-518
        const double eps = 1e-6;
        if (inside(p)) return p;
        V3 mn = pos, mx = maxp();
        p.x = (float)std::fmax(p.x, mn.x);
        p.y = (float)std::fmax(p.y, mn.y);
        p.z = (float)std::fmax(p.z, mn.z);
        p.x = (float)std::fmin((double)p.x, (double)mx.x - eps);
        p.y = (float)std::fmin((double)p.y, (double)mx.y - eps);
        p.z = (float)std::fmin((double)p.z, (double)mx.z - eps);
        return p;
    }
};

// FluidSimulation::_getBoundaryAABB (fluidsimulation.cpp:2834-2839).
static Box boundary_box(int isize, int jsize, int ksize, double dx) {
    Box b;
    b.pos = V3(0, 0, 0);
    b.w = isize * dx; b.h = jsize * dx; b.d = ksize * dx;
    b.expand(-3 * dx - 1e-4);
    return b;
}

// ---------------------------------------------------------------------------
// Domain solid SDF (nodal, (I+1)(J+1)(K+1)).  The domain is a box mesh
// (fluidsimulation.cpp:2805-2845) whose exact-band unsigned distance is
// computed per 10^3 node block in block-local float coordinates
// (meshlevelset.cpp:773-1010), signed by inside/outside (meshlevelset.cpp:
// 1298-1317) and negated (meshlevelset.cpp:702-708, fluidsimulation.cpp:
// 2918-2930), so the fluid side is positive.
// ---------------------------------------------------------------------------

// Collision::findClosestPointOnTriangle (collision.cpp:161-270).
static V3 closest_point_on_triangle(V3 p0, V3 v0, V3 v1, V3 v2) {
    V3 edge0 = v1 - v0, edge1 = v2 - v0, pv = v0 - p0;
    double a = dot(edge0, edge0), b = dot(edge0, edge1), c = dot(edge1, edge1);
    double d = dot(edge0, pv), e = dot(edge1, pv);
    double det = a * c - b * b;
    if (std::abs(det) < 1e-9) {
        float feps = 1e-6f;
        float len1 = length(v0 - v1), len2 = length(v1 - v2), len3 = length(v2 - v0);
        if (len1 < feps && len2 < feps && len3 < feps) return v0;
        V3 dp, dq;
        if (len1 < feps) { dp = v0; dq = v2; }
        else if (len2 < feps) { dp = v1; dq = v0; }
        else { dp = v2; dq = v1; }
        float num = dot(p0 - dp, dq - dp), den = dot(dq - dp, dq - dp);
        if (std::abs(den) < feps) return v0;
        float lam = num / den;
        if (lam <= 0.0f) return dp;
        if (lam >= 1.0f) return dq;
        return dp + lam * (dq - dp);
    }
    double s = b * e - c * d, t = b * d - a * e;
    if (s + t < det) {
        if (s < 0.0) {
            if (t < 0.0) {
                if (d < 0.0) { s = clampv(-d / a, 0.0, 1.0); t = 0.0; }
                else { s = 0.0; t = clampv(-e / c, 0.0, 1.0); }
            } else { s = 0.0; t = clampv(-e / c, 0.0, 1.0); }
        } else if (t < 0.0) { s = clampv(-d / a, 0.0, 1.0); t = 0.0; }
        else { double inv = 1.0 / det; s *= inv; t *= inv; }
    } else {
        if (s < 0.0) {
            double tmp0 = b + d, tmp1 = c + e;
            if (tmp1 > tmp0) {
                double numer = tmp1 - tmp0, denom = a - 2 * b + c;
                s = clampv(numer / denom, 0.0, 1.0); t = 1 - s;
            } else { t = clampv(-e / c, 0.0, 1.0); s = 0.0; }
        } else if (t < 0.0) {
            if (a + d > b + e) {
                double numer = c + e - b - d, denom = a - 2 * b + c;
                s = clampv(numer / denom, 0.0, 1.0); t = 1 - s;
            } else { s = clampv(-e / c, 0.0, 1.0); t = 0.0; }
        } else {
            double numer = c + e - b - d, denom = a - 2 * b + c;
            s = clampv(numer / denom, 0.0, 1.0); t = 1.0 - s;
        }
    }
    return v0 + (float)s * edge0 + (float)t * edge1;
}

struct Tri {
    V3 v[3];
    Idx gmin, gmax;
};

// MeshLevelSet::_initializeTriangleData for one triangle (meshlevelset.cpp:
// 878-915): the node range [floor(min/dx) - band, floor(max/dx) + band + 1]
// of the triangle's AABB (aabb.cpp:84-112), clamped to the node grid; false
// when the range misses the grid.
static bool tri_band(Tri &T, int isize, int jsize, int ksize, double dx, int band) {
    double mnx = T.v[0].x, mny = T.v[0].y, mnz = T.v[0].z;
    double mxx = mnx, mxy = mny, mxz = mnz;
    for (int c = 0; c < 3; c++) {
        mnx = std::fmin((double)T.v[c].x, mnx); mny = std::fmin((double)T.v[c].y, mny);
        mnz = std::fmin((double)T.v[c].z, mnz); mxx = std::fmax((double)T.v[c].x, mxx);
        mxy = std::fmax((double)T.v[c].y, mxy); mxz = std::fmax((double)T.v[c].z, mxz);
    }
    Box tb;
    tb.pos = V3((float)mnx, (float)mny, (float)mnz);
    tb.w = mxx - mnx + 1e-9; tb.h = mxy - mny + 1e-9; tb.d = mxz - mnz + 1e-9;
    V3 pmax = tb.maxp();
    Idx gmin = pos_to_index(tb.pos, dx), gmax = pos_to_index(pmax, dx);
    gmin.i -= band; gmin.j -= band; gmin.k -= band;
    gmax.i += band + 1; gmax.j += band + 1; gmax.k += band + 1;
    if (gmax.i < 0 || gmin.i >= isize + 1 || gmax.j < 0 || gmin.j >= jsize + 1 || gmax.k < 0 ||
        gmin.k >= ksize + 1)
        return false;
    gmin.i = std::max(gmin.i, 0); gmin.j = std::max(gmin.j, 0); gmin.k = std::max(gmin.k, 0);
    gmax.i = std::min(gmax.i, isize); gmax.j = std::min(gmax.j, jsize);
    gmax.k = std::min(gmax.k, ksize);
    T.gmin = gmin; T.gmax = gmax;
    return true;
}

// Box mesh (FluidSimulation::_getTriangleMeshFromAABB, fluidsimulation.cpp:
// 2805-2832) with per-triangle exact-band node ranges
// (MeshLevelSet::_initializeTriangleData, meshlevelset.cpp:874-915).
static std::vector<Tri> box_triangles(const Box &bb, int isize, int jsize, int ksize, double dx,
                                      int band) {
    V3 p = bb.pos;
    V3 vert[8] = {
        V3(p.x, p.y, p.z),
        V3((float)(p.x + bb.w), p.y, p.z),
        V3((float)(p.x + bb.w), p.y, (float)(p.z + bb.d)),
        V3(p.x, p.y, (float)(p.z + bb.d)),
        V3(p.x, (float)(p.y + bb.h), p.z),
        V3((float)(p.x + bb.w), (float)(p.y + bb.h), p.z),
        V3((float)(p.x + bb.w), (float)(p.y + bb.h), (float)(p.z + bb.d)),
        V3(p.x, (float)(p.y + bb.h), (float)(p.z + bb.d)),
    };
    const int tri[12][3] = {{0, 1, 2}, {0, 2, 3}, {4, 7, 6}, {4, 6, 5}, {0, 3, 7}, {0, 7, 4},
                            {1, 5, 6}, {1, 6, 2}, {0, 4, 5}, {0, 5, 1}, {3, 2, 6}, {3, 6, 7}};
    std::vector<Tri> out;
    for (int t = 0; t < 12; t++) {
        Tri T;
        for (int c = 0; c < 3; c++) T.v[c] = vert[tri[t][c]];
        if (tri_band(T, isize, jsize, ksize, dx, band)) out.push_back(T);
    }
    return out;
}

static void domain_solid_phi(int isize, int jsize, int ksize, double dx, GridF &phi) {
    const int band = 3;     // fluidsimulation.h:1580 (_solidLevelSetExactBand)
    const int bw = 10;      // meshlevelset.h:370 (_blockwidth)
    phi = GridF(isize + 1, jsize + 1, ksize + 1, 0.0f);
    float upper = (float)((phi.ni + phi.nj + phi.nk) * dx);  // meshlevelset.cpp:760-762
    Box bb = boundary_box(isize, jsize, ksize, dx);
    std::vector<Tri> tris = box_triangles(bb, isize, jsize, ksize, dx, band);
    for (int k = 0; k < phi.nk; k++)
        for (int j = 0; j < phi.nj; j++)
            for (int i = 0; i < phi.ni; i++) {
                // Each node is owned by one 10^3 block; distances are taken in
                // that block's local frame (meshlevelset.cpp:967-1000).
                int bi = i / bw, bj = j / bw, bk = k / bw;
                V3 boff = index_to_pos(bi, bj, bk, bw * dx);
                V3 gpos = index_to_pos(i - bi * bw, j - bj * bw, k - bk * bw, dx);
                float best = upper;
                for (const Tri &t : tris) {
                    if (i < t.gmin.i || j < t.gmin.j || k < t.gmin.k || i > t.gmax.i ||
                        j > t.gmax.j || k > t.gmax.k)
                        continue;
                    V3 cp = closest_point_on_triangle(gpos, t.v[0] - boff, t.v[1] - boff,
                                                      t.v[2] - boff);
                    float dist = length(cp - gpos);
                    if (dist < best) best = dist;
                }
                // Sign: nodes strictly inside the closed box mesh are inside
                // (MeshUtils::getGridNodesInsideTriangleMesh, meshutils.cpp:
                // 347-428); the field is then negated (meshlevelset.cpp:702).
                V3 np = index_to_pos(i, j, k, dx);
                bool inside = bb.inside(np);
                float signed_d = inside ? -best : best;
                phi(i, j, k) = -signed_d;
            }
}

// LevelsetUtils::fractionInside(float, float) (levelsetutils.cpp:39-52).
static float fraction_inside2(float l, float r) {
    if (l < 0 && r < 0) return 1;
    if (l < 0 && r >= 0) return l / (l - r);
    if (l >= 0 && r < 0) return r / (r - l);
    return 0;
}

static void cycle4(float *a) {
    float t = a[0];
    a[0] = a[1]; a[1] = a[2]; a[2] = a[3]; a[3] = t;
}

// LevelsetUtils::fractionInside(float x4) (levelsetutils.cpp:62-141).
static float fraction_inside4(float bl, float br, float tl, float tr) {
    int n = (bl < 0 ? 1 : 0) + (tl < 0 ? 1 : 0) + (br < 0 ? 1 : 0) + (tr < 0 ? 1 : 0);
    float L[4] = {bl, br, tr, tl};
    if (n == 4) return 1;
    if (n == 3) {
        while (L[0] < 0) cycle4(L);
        float s0 = 1 - fraction_inside2(L[0], L[3]);
        float s1 = 1 - fraction_inside2(L[0], L[1]);
        return 1.0f - 0.5f * s0 * s1;
    }
    if (n == 2) {
        while (L[0] >= 0 || !(L[1] < 0 || L[2] < 0)) cycle4(L);
        if (L[1] < 0) {
            float sl = fraction_inside2(L[0], L[3]);
            float sr = fraction_inside2(L[1], L[2]);
            return 0.5f * (sl + sr);
        }
        float mid = 0.25f * (L[0] + L[1] + L[2] + L[3]);
        if (mid < 0) {
            float area = 0;
            float s1 = 1 - fraction_inside2(L[0], L[3]);
            float s3 = 1 - fraction_inside2(L[2], L[3]);
            area += 0.5f * s1 * s3;
            float s2 = 1 - fraction_inside2(L[2], L[1]);
            float s0 = 1 - fraction_inside2(L[0], L[1]);
            area += 0.5f * s0 * s2;
            return 1.0f - area;
        }
        float area = 0;
        float s0 = fraction_inside2(L[0], L[1]);
        float s1 = fraction_inside2(L[0], L[3]);
        area += 0.5f * s0 * s1;
        float s2 = fraction_inside2(L[2], L[1]);
        float s3 = fraction_inside2(L[2], L[3]);
        area += 0.5f * s2 * s3;
        return area;
    }
    if (n == 1) {
        while (L[0] >= 0) cycle4(L);
        float s0 = fraction_inside2(L[0], L[3]);
        float s1 = fraction_inside2(L[0], L[1]);
        return 0.5f * s0 * s1;
    }
    return 0;
}

// Tetrahedron volume fractions (levelsetutils.h:62-102, levelsetutils.cpp:
// 213-259) and the two-decomposition cube average (levelsetutils.cpp:245-259).
static inline void sort4(float &a, float &b, float &c, float &d) {
    if (a > b) std::swap(a, b);
    if (c > d) std::swap(c, d);
    if (a > c) std::swap(a, c);
    if (b > d) std::swap(b, d);
    if (b > c) std::swap(b, c);
}
static inline float tet_frac(float p0, float p1, float p2, float p3) {
    return p0 * p0 * p0 / ((p0 - p1) * (p0 - p2) * (p0 - p3));
}
static inline float prism_frac(float p0, float p1, float p2, float p3) {
    float a = p0 / (p0 - p2), b = p0 / (p0 - p3), c = p1 / (p1 - p3), d = p1 / (p1 - p2);
    return a * b * (1 - d) + b * (1 - c) * d + c * d;
}
static float volume_fraction4(float p0, float p1, float p2, float p3) {
    sort4(p0, p1, p2, p3);
    if (p3 <= 0) return 1;
    if (p2 <= 0) return 1 - tet_frac(p3, p2, p1, p0);
    if (p1 <= 0) return prism_frac(p0, p1, p2, p3);
    if (p0 <= 0) return tet_frac(p0, p1, p2, p3);
    return 0;
}
static float volume_fraction8(float p000, float p100, float p010, float p110, float p001,
                              float p101, float p011, float p111) {
    return (volume_fraction4(p000, p001, p101, p011) + volume_fraction4(p000, p101, p100, p110) +
            volume_fraction4(p000, p010, p011, p110) + volume_fraction4(p101, p011, p111, p110) +
            2 * volume_fraction4(p000, p011, p101, p110) +
            volume_fraction4(p100, p101, p001, p111) + volume_fraction4(p100, p001, p000, p010) +
            volume_fraction4(p100, p110, p111, p010) + volume_fraction4(p001, p111, p011, p010) +
            2 * volume_fraction4(p100, p111, p001, p010)) /
           12.0f;
}

struct Weights {
    GridF U, V, W, C;
};

// FluidSimulation::_updateWeightGridThread (fluidsimulation.cpp:3688-3730)
// with MeshLevelSet::getFaceWeightU/V/W and _getCellWeight
// (meshlevelset.cpp:357-391, 1490-1513).
static void compute_weights(const GridF &phi, int isize, int jsize, int ksize, Weights &w) {
    w.U = GridF(isize + 1, jsize, ksize, 0.0f);
    w.V = GridF(isize, jsize + 1, ksize, 0.0f);
    w.W = GridF(isize, jsize, ksize + 1, 0.0f);
    w.C = GridF(isize, jsize, ksize, 0.0f);
    for (int k = 0; k < ksize; k++)
        for (int j = 0; j < jsize; j++)
            for (int i = 0; i < isize + 1; i++) {
                float f = fraction_inside4(phi(i, j, k), phi(i, j + 1, k), phi(i, j, k + 1),
                                           phi(i, j + 1, k + 1));
                w.U(i, j, k) = clampv(1.0f - f, 0.0f, 1.0f);
            }
    for (int k = 0; k < ksize; k++)
        for (int j = 0; j < jsize + 1; j++)
            for (int i = 0; i < isize; i++) {
                float f = fraction_inside4(phi(i, j, k), phi(i, j, k + 1), phi(i + 1, j, k),
                                           phi(i + 1, j, k + 1));
                w.V(i, j, k) = clampv(1.0f - f, 0.0f, 1.0f);
            }
    for (int k = 0; k < ksize + 1; k++)
        for (int j = 0; j < jsize; j++)
            for (int i = 0; i < isize; i++) {
                float f = fraction_inside4(phi(i, j, k), phi(i, j + 1, k), phi(i + 1, j, k),
                                           phi(i + 1, j + 1, k));
                w.W(i, j, k) = clampv(1.0f - f, 0.0f, 1.0f);
            }
    for (int k = 0; k < ksize; k++)
        for (int j = 0; j < jsize; j++)
            for (int i = 0; i < isize; i++) {
                float p000 = phi(i, j, k), p001 = phi(i, j, k + 1), p010 = phi(i, j + 1, k),
                      p011 = phi(i, j + 1, k + 1), p100 = phi(i + 1, j, k),
                      p101 = phi(i + 1, j, k + 1), p110 = phi(i + 1, j + 1, k),
                      p111 = phi(i + 1, j + 1, k + 1);
                float cw;
                if (p000 < 0 && p001 < 0 && p010 < 0 && p011 < 0 && p100 < 0 && p101 < 0 &&
                    p110 < 0 && p111 < 0)
                    cw = 1.0f;
                else if (p000 >= 0 && p001 >= 0 && p010 >= 0 && p011 >= 0 && p100 >= 0 &&
                         p101 >= 0 && p110 >= 0 && p111 >= 0)
                    cw = 0.0f;
                else
                    cw = volume_fraction8(p000, p100, p010, p110, p001, p101, p011, p111);
                w.C(i, j, k) = clampv(1.0f - cw, 0.0f, 1.0f);
            }
}

// MeshLevelSet::getDistanceAtCellCenter (meshlevelset.cpp:152-162).
static inline float solid_at_center(const GridF &phi, int i, int j, int k) {
    return 0.125f * (phi(i, j, k) + phi(i + 1, j, k) + phi(i, j + 1, k) + phi(i + 1, j + 1, k) +
                     phi(i, j, k + 1) + phi(i + 1, j, k + 1) + phi(i, j + 1, k + 1) +
                     phi(i + 1, j + 1, k + 1));
}

// Interpolation::trilinearInterpolate(p[8], x, y, z) (interpolation.cpp:62-71).
static inline double trilerp(const double *p, double x, double y, double z) {
    return p[0] * (1 - x) * (1 - y) * (1 - z) + p[1] * x * (1 - y) * (1 - z) +
           p[2] * (1 - x) * y * (1 - z) + p[3] * (1 - x) * (1 - y) * z +
           p[4] * x * (1 - y) * z + p[5] * (1 - x) * y * z + p[6] * x * y * (1 - z) +
           p[7] * x * y * z;
}

// Interpolation::trilinearInterpolate(vec3, dx, Array3d<float>) (interpolation.cpp:73-112)
// -- used for the nodal solid SDF (meshlevelset.cpp:261-263).
static float grid_trilinear(const GridF &g, V3 p, double dx) {
    Idx c = pos_to_index(p, dx);
    V3 gp = index_to_pos(c.i, c.j, c.k, dx);
    double inv = 1.0 / dx;
    double ix = (double)(p.x - gp.x) * inv, iy = (double)(p.y - gp.y) * inv,
           iz = (double)(p.z - gp.z) * inv;
    double pts[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                           {1, 0, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 1}};
    for (int q = 0; q < 8; q++) {
        int a = c.i + off[q][0], b = c.j + off[q][1], d = c.k + off[q][2];
        if (g.in(a, b, d)) pts[q] = g(a, b, d);
    }
    return (float)trilerp(pts, ix, iy, iz);
}

// Interpolation::bilinearInterpolate / trilinearInterpolateGradient
// (interpolation.cpp:193-258).
static inline double bilerp(double v00, double v10, double v01, double v11, double ix, double iy) {
    double l1 = (1 - ix) * v00 + ix * v10;
    double l2 = (1 - ix) * v01 + ix * v11;
    return (1 - iy) * l1 + iy * l2;
}
static V3 grid_gradient(const GridF &g, V3 p, double dx) {
    Idx c = pos_to_index(p, dx);
    V3 gp = index_to_pos(c.i, c.j, c.k, dx);
    double inv = 1.0 / dx;
    double ix = (double)(p.x - gp.x) * inv, iy = (double)(p.y - gp.y) * inv,
           iz = (double)(p.z - gp.z) * inv;
    auto at = [&](int a, int b, int d) -> float { return g.in(a, b, d) ? g(a, b, d) : 0.0f; };
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
    return V3(dvdx, dvdy, dvdz);
}

// ---------------------------------------------------------------------------
// Block assignment shared by the liquid SDF and P2G: the reference bins each
// point into every 10^3 block its search box [p-sr, p+sr] overlaps
// (particlelevelset.cpp:487-536, velocityadvector.cpp:270-320) and then, per
// block, visits the grid points of the clamped local box
// (particlelevelset.cpp:582-630, velocityadvector.cpp:390-446).
// The oracle visits the same (point, block) pairs in ascending point order,
// which is the order the reference's per-block lists hold.
// ---------------------------------------------------------------------------
template <class F>
static void for_each_block_of_point(V3 p, float sr, float blockdx, int nbi, int nbj, int nbk,
                                    F &&fn) {
    Idx b = pos_to_index(p, blockdx);
    V3 bp = index_to_pos(b.i, b.j, b.k, blockdx);
    auto in_range = [&](int i, int j, int k) {
        return i >= 0 && j >= 0 && k >= 0 && i < nbi && j < nbj && k < nbk;
    };
    if (p.x - sr > bp.x && p.y - sr > bp.y && p.z - sr > bp.z && p.x + sr < bp.x + blockdx &&
        p.y + sr < bp.y + blockdx && p.z + sr < bp.z + blockdx) {
        if (in_range(b.i, b.j, b.k)) fn(b.i, b.j, b.k);
        return;
    }
    Idx gmin = pos_to_index(p.x - sr, p.y - sr, p.z - sr, blockdx);
    Idx gmax = pos_to_index(p.x + sr, p.y + sr, p.z + sr, blockdx);
    for (int k = gmin.k; k <= gmax.k; k++)
        for (int j = gmin.j; j <= gmax.j; j++)
            for (int i = gmin.i; i <= gmax.i; i++)
                if (in_range(i, j, k)) fn(i, j, k);
}

// ParticleLevelSet::calculateSignedDistanceField (particlelevelset.cpp:161-170,
// 299-630): phi(c) = min(3 dx, min_p |c - p| - r) in block-local float
// arithmetic.
static void liquid_sdf(int isize, int jsize, int ksize, double dx, const std::vector<V3> &pts,
                       double radius, GridF &phi) {
    const int bw = 10;  // particlelevelset.h:140
    float maxd = (float)(3.0 * dx);
    phi = GridF(isize, jsize, ksize, maxd);
    if (pts.empty()) return;
    float r = (float)radius;
    float sr = 2.0f * r;  // particlelevelset.h:142 (_searchRadiusFactor)
    float blockdx = (float)(bw * dx);
    int nbi = (isize + bw - 1) / bw, nbj = (jsize + bw - 1) / bw, nbk = (ksize + bw - 1) / bw;
    for (const V3 &p0 : pts) {
        for_each_block_of_point(p0, (float)(2.0f * (float)radius), blockdx, nbi, nbj, nbk,
                                [&](int bi, int bj, int bk) {
            V3 boff = index_to_pos(bi, bj, bk, bw * dx);
            V3 p = p0 - boff;
            V3 pmin(p.x - sr, p.y - sr, p.z - sr), pmax(p.x + sr, p.y + sr, p.z + sr);
            Idx gmin = pos_to_index(pmin, dx), gmax = pos_to_index(pmax, dx);
            gmin.i = std::max(gmin.i, 0); gmin.j = std::max(gmin.j, 0); gmin.k = std::max(gmin.k, 0);
            gmax.i = std::min(gmax.i, bw - 1); gmax.j = std::min(gmax.j, bw - 1);
            gmax.k = std::min(gmax.k, bw - 1);
            for (int k = gmin.k; k <= gmax.k; k++)
                for (int j = gmin.j; j <= gmax.j; j++)
                    for (int i = gmin.i; i <= gmax.i; i++) {
                        int gi = bi * bw + i, gj = bj * bw + j, gk = bk * bw + k;
                        V3 gpos = index_to_center(i, j, k, dx);
                        float dist = length(gpos - p) - r;
                        if (!phi.in(gi, gj, gk)) continue;
                        float &dst = phi(gi, gj, gk);
                        if (dist < dst) dst = dist;
                    }
        });
    }
}

// ParticleLevelSet::postProcessSignedDistanceField (particlelevelset.cpp:172-196).
static void liquid_postprocess(GridF &phi, const GridF &solid, double dx) {
    float eps = (float)(0.005 * dx);
    for (int k = 0; k < phi.nk; k++)
        for (int j = 0; j < phi.nj; j++)
            for (int i = 0; i < phi.ni; i++) {
                if ((double)phi(i, j, k) < 0.5 * dx) {
                    if (solid_at_center(solid, i, j, k) < 0) phi(i, j, k) = (float)(-0.5f * dx);
                }
                float val = phi(i, j, k);
                if (std::abs(val) < eps) phi(i, j, k) = val > 0 ? eps : -eps;
            }
}

struct MAC {
    GridF U, V, W;
};
struct Valid {
    GridB U, V, W;
};

static void mac_init(MAC &m, int isize, int jsize, int ksize) {
    m.U = GridF(isize + 1, jsize, ksize, 0.0f);
    m.V = GridF(isize, jsize + 1, ksize, 0.0f);
    m.W = GridF(isize, jsize, ksize + 1, 0.0f);
}
static void valid_init(Valid &v, int isize, int jsize, int ksize) {
    v.U = GridB(isize + 1, jsize, ksize, 0);
    v.V = GridB(isize, jsize + 1, ksize, 0);
    v.W = GridB(isize, jsize, ksize + 1, 0);
}

// VelocityAdvector::_advectGrid for one component (velocityadvector.cpp:
// 64-151, 166-459): Wyvill-kernel weighted average of particle velocity on the
// faces of one MAC component, accumulated in ascending particle order per face.
static void p2g_component(int comp, double dx, const std::vector<V3> &pts,
                          const std::vector<V3> &vel, double radius, GridF &out, GridB &valid) {
    const int bw = 10;  // velocityadvector.h:163 (_chunkWidth)
    double chunkdx = dx * bw;
    V3 offset;  // velocityadvector.cpp:153-164
    if (comp == 0) offset = V3(0.0f, (float)(0.5 * dx), (float)(0.5 * dx));
    else if (comp == 1) offset = V3((float)(0.5 * dx), 0.0f, (float)(0.5 * dx));
    else offset = V3((float)(0.5 * dx), (float)(0.5 * dx), 0.0f);
    int ni = out.ni, nj = out.nj, nk = out.nk;
    int nbi = (ni + bw - 1) / bw, nbj = (nj + bw - 1) / bw, nbk = (nk + bw - 1) / bw;
    float eps = 1e-6f;
    float sr = (float)(radius + eps);
    float blockdx = (float)chunkdx;
    float r = (float)radius, rsq = r * r;
    float coef1 = (4.0f / 9.0f) * (1.0f / (r * r * r * r * r * r));
    float coef2 = (17.0f / 9.0f) * (1.0f / (r * r * r * r));
    float coef3 = (22.0f / 9.0f) * (1.0f / (r * r));
    GridF sum(ni, nj, nk, 0.0f), wsum(ni, nj, nk, 0.0f);
    std::vector<uint8_t> touched_block((size_t)nbi * nbj * nbk, 0);
    for (size_t pi = 0; pi < pts.size(); pi++) {
        V3 pp = pts[pi] - offset;
        float pv = comp == 0 ? vel[pi].x : (comp == 1 ? vel[pi].y : vel[pi].z);
        for_each_block_of_point(pp, sr, blockdx, nbi, nbj, nbk, [&](int bi, int bj, int bk) {
            touched_block[(size_t)bi + (size_t)nbi * (bj + (size_t)nbj * bk)] = 1;
            V3 boff = index_to_pos(bi, bj, bk, bw * dx);
            V3 p = pp - boff;
            V3 pmin(p.x - sr, p.y - sr, p.z - sr), pmax(p.x + sr, p.y + sr, p.z + sr);
            Idx gmin = pos_to_index(pmin, dx), gmax = pos_to_index(pmax, dx);
            gmin.i = std::max(gmin.i, 0); gmin.j = std::max(gmin.j, 0); gmin.k = std::max(gmin.k, 0);
            gmax.i = std::min(gmax.i, bw - 1); gmax.j = std::min(gmax.j, bw - 1);
            gmax.k = std::min(gmax.k, bw - 1);
            for (int k = gmin.k; k <= gmax.k; k++)
                for (int j = gmin.j; j <= gmax.j; j++)
                    for (int i = gmin.i; i <= gmax.i; i++) {
                        V3 gpos = index_to_pos(i, j, k, dx);
                        V3 v = gpos - p;
                        float d2 = dot(v, v);
                        if (d2 < rsq) {
                            float wt = 1.0f - coef1 * d2 * d2 * d2 + coef2 * d2 * d2 - coef3 * d2;
                            int gi = bi * bw + i, gj = bj * bw + j, gk = bk * bw + k;
                            if (!sum.in(gi, gj, gk)) continue;
                            sum(gi, gj, gk) += wt * pv;
                            wsum(gi, gj, gk) += wt;
                        }
                    }
        });
    }
    // Normalise and write back every face of each processed block
    // (velocityadvector.cpp:116-144, 448-453).
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                if (!touched_block[(size_t)(i / bw) + (size_t)nbi * ((j / bw) + (size_t)nbj * (k / bw))])
                    continue;
                float s = sum(i, j, k), w = wsum(i, j, k);
                if (w > eps) s /= w;
                out(i, j, k) = s;
                if (w > eps) valid(i, j, k) = 1;
            }
}

static void p2g(int isize, int jsize, int ksize, double dx, const std::vector<V3> &pts,
                const std::vector<V3> &vel, double radius, MAC &m, Valid &v) {
    (void)isize; (void)jsize; (void)ksize;
    p2g_component(0, dx, pts, vel, radius, m.U, v.U);
    p2g_component(1, dx, pts, vel, radius, m.V, v.V);
    p2g_component(2, dx, pts, vel, radius, m.W, v.W);
}

// GridUtils::extrapolateGrid (gridutils.cpp:32-228).  Layer-synchronous
// breadth-first fill: each new face takes the mean of its finished 6
// neighbours in the order +i,-i,+j,-j,+k,-k; border faces count as finished.
static void extrapolate(GridF &g, const GridB &valid, int layers) {
    const uint8_t UNKNOWN = 0, WAITING = 1, KNOWN = 2, DONE = 3;
    GridB st(g.ni, g.nj, g.nk, UNKNOWN);
    for (int k = 0; k < g.nk; k++)
        for (int j = 0; j < g.nj; j++)
            for (int i = 0; i < g.ni; i++) {
                if (i == 0 || j == 0 || k == 0 || i == g.ni - 1 || j == g.nj - 1 || k == g.nk - 1)
                    st(i, j, k) = DONE;
                else if (valid(i, j, k)) st(i, j, k) = KNOWN;
            }
    const int nb[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    std::vector<Idx> cells;
    for (int layer = 0; layer < layers; layer++) {
        cells.clear();
        for (int k = 0; k < g.nk; k++)
            for (int j = 0; j < g.nj; j++)
                for (int i = 0; i < g.ni; i++) {
                    if (st(i, j, k) != KNOWN) continue;
                    for (int q = 0; q < 6; q++) {
                        int a = i + nb[q][0], b = j + nb[q][1], c = k + nb[q][2];
                        if (st(a, b, c) == UNKNOWN) {
                            st(a, b, c) = WAITING;
                            cells.push_back(Idx{a, b, c});
                        }
                    }
                    st(i, j, k) = DONE;
                }
        for (const Idx &c : cells) {
            float s = 0.0f;
            int n = 0;
            for (int q = 0; q < 6; q++) {
                int a = c.i + nb[q][0], b = c.j + nb[q][1], d = c.k + nb[q][2];
                if (st(a, b, d) == DONE) { s += g(a, b, d); n++; }
            }
            g(c.i, c.j, c.k) = s / (float)n;
        }
        if (layer != layers - 1)
            for (const Idx &c : cells) st(c.i, c.j, c.k) = KNOWN;
    }
}

// MACVelocityField::_interpolateLinearU/V/W + evaluateVelocityAtPositionLinear
// (macvelocityfield.cpp:519-645): trilinear sampling in double.
static double interp_component(const GridF &g, int comp, double x, double y, double z,
                               double dx, int isize, int jsize, int ksize) {
    if (!(x >= 0 && y >= 0 && z >= 0 && x < dx * isize && y < dx * jsize && z < dx * ksize))
        return 0.0;
    if (comp != 0) x -= 0.5 * dx;
    if (comp != 1) y -= 0.5 * dx;
    if (comp != 2) z -= 0.5 * dx;
    double inv = 1.0 / dx;
    int i = (int)std::floor(x * inv), j = (int)std::floor(y * inv), k = (int)std::floor(z * inv);
    double gx = (double)i * dx, gy = (double)j * dx, gz = (double)k * dx;
    double invdx = 1 / dx;
    double ix = (x - gx) * invdx, iy = (y - gy) * invdx, iz = (z - gz) * invdx;
    double pts[8];
    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                           {1, 0, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 1}};
    for (int q = 0; q < 8; q++) {
        int a = i + off[q][0], b = j + off[q][1], c = k + off[q][2];
        pts[q] = g.in(a, b, c) ? (double)g(a, b, c) : 0.0;
    }
    return trilerp(pts, ix, iy, iz);
}

static V3 mac_velocity(const MAC &m, V3 p, double dx, int isize, int jsize, int ksize) {
    double x = p.x, y = p.y, z = p.z;
    if (!(x >= 0 && y >= 0 && z >= 0 && x < dx * isize && y < dx * jsize && z < dx * ksize))
        return V3();
    double u = interp_component(m.U, 0, x, y, z, dx, isize, jsize, ksize);
    double v = interp_component(m.V, 1, x, y, z, dx, isize, jsize, ksize);
    double w = interp_component(m.W, 2, x, y, z, dx, isize, jsize, ksize);
    return V3((float)u, (float)v, (float)w);
}

// ---------------------------------------------------------------------------
// Pressure projection (pressuresolver.cpp:44-1047) and MIC(0)-PCG
// (pcgsolver/pcgsolver.h:69-302, sparsematrix.h:175-216, blaswrapper.h:35-158).
// ---------------------------------------------------------------------------
struct PressureInputs {
    int isize, jsize, ksize;
    double dx, dt, tol;
    int maxiter, mode;
    const GridF *liquid;
    const Weights *w;
    GridF *solidU, *solidV, *solidW;  // solid face velocities (mutable)
    MAC *mac;
    Valid *valid;
};

struct PressureResult {
    int n = 0, iters = 0, ok = 1;
    double err = 0.0;
    bool solved = false;
    std::vector<double> rhs, soln;
};

// Row layout: fluid interior cells in k, j, i order (pressuresolver.cpp:
// 101-110).  For each row the couplings to the +x/+y/+z neighbouring rows
// (-1 if that neighbour is not a row) are stored; couplings to -x/-y/-z are
// the neighbours' +x/+y/+z couplings by symmetry.
struct Rows {
    std::vector<Idx> cell;
    Grid<int> key;  // GridIndexKeyMap (gridindexkeymap.cpp:49-69)
    std::vector<double> diag, offx, offy, offz;
    std::vector<int> nxp, nyp, nzp, nxm, nym, nzm;
};

static const double kMaxTheta = 25.0;  // pressuresolver.h (_maxtheta)

// PressureSolver::_conditionSolidVelocityField (pressuresolver.cpp:124-268):
// zero the solid face velocities around fluid pockets that never touch air.
static void condition_solid_velocity(const PressureInputs &in) {
    int I = in.isize, J = in.jsize, K = in.ksize;
    const GridF &phi = *in.liquid;
    const Weights &w = *in.w;
    float eps = 1e-6f;
    GridB bordersAir(I, J, K, 0);
    for (int k = 1; k < K - 1; k++)
        for (int j = 1; j < J - 1; j++)
            for (int i = 1; i < I - 1; i++) {
                if ((w.U(i, j, k) >= 0.0f + eps && phi(i - 1, j, k) >= 0.0f) ||
                    (w.U(i + 1, j, k) >= 0.0f + eps && phi(i + 1, j, k) >= 0.0f) ||
                    (w.V(i, j, k) >= 0.0f + eps && phi(i, j - 1, k) >= 0.0f) ||
                    (w.V(i, j + 1, k) >= 0.0f + eps && phi(i, j + 1, k) >= 0.0f) ||
                    (w.W(i, j, k) >= 0.0f + eps && phi(i, j, k - 1) >= 0.0f) ||
                    (w.W(i, j, k + 1) >= 0.0f + eps && phi(i, j, k + 1) >= 0.0f))
                    bordersAir(i, j, k) = 1;
            }
    GridB done(I, J, K, 0);
    std::vector<Idx> queue, group;
    for (int k = 1; k < K - 1; k++)
        for (int j = 1; j < J - 1; j++)
            for (int i = 1; i < I - 1; i++) {
                if (phi(i, j, k) >= 0.0f) { done(i, j, k) = 1; continue; }
                if (done(i, j, k)) continue;
                queue.push_back(Idx{i, j, k});
                done(i, j, k) = 1;
                group.clear();
                while (!queue.empty()) {
                    Idx g = queue.back();
                    queue.pop_back();
                    // A region can reach border cells; their out-of-grid
                    // neighbours are not cells and are never visited.
                    auto visit = [&](int a, int b, int c, float wf) {
                        if (a < 0 || b < 0 || c < 0 || a >= I || b >= J || c >= K) return;
                        if (!done(a, b, c) && phi(a, b, c) < 0.0f && wf >= 0.0f + eps) {
                            queue.push_back(Idx{a, b, c});
                            done(a, b, c) = 1;
                        }
                    };
                    visit(g.i - 1, g.j, g.k, w.U(g.i, g.j, g.k));
                    visit(g.i + 1, g.j, g.k, w.U(g.i + 1, g.j, g.k));
                    visit(g.i, g.j - 1, g.k, w.V(g.i, g.j, g.k));
                    visit(g.i, g.j + 1, g.k, w.V(g.i, g.j + 1, g.k));
                    visit(g.i, g.j, g.k - 1, w.W(g.i, g.j, g.k));
                    visit(g.i, g.j, g.k + 1, w.W(g.i, g.j, g.k + 1));
                    group.push_back(g);
                }
                if (group.size() == 1) continue;
                bool isolated = true;
                for (const Idx &g : group)
                    if (bordersAir(g.i, g.j, g.k)) { isolated = false; break; }
                if (!isolated) continue;
                for (const Idx &g : group) {
                    (*in.solidU)(g.i, g.j, g.k) = 0.0f; (*in.solidU)(g.i + 1, g.j, g.k) = 0.0f;
                    (*in.solidV)(g.i, g.j, g.k) = 0.0f; (*in.solidV)(g.i, g.j + 1, g.k) = 0.0f;
                    (*in.solidW)(g.i, g.j, g.k) = 0.0f; (*in.solidW)(g.i, g.j, g.k + 1) = 0.0f;
                }
            }
}

static void build_rows(const PressureInputs &in, Rows &R) {
    int I = in.isize, J = in.jsize, K = in.ksize;
    const GridF &phi = *in.liquid;
    R.key = Grid<int>(I, J, K, -1);
    R.cell.clear();
    for (int k = 1; k < K - 1; k++)
        for (int j = 1; j < J - 1; j++)
            for (int i = 1; i < I - 1; i++)
                if (phi(i, j, k) < 0) {
                    R.key(i, j, k) = (int)R.cell.size();
                    R.cell.push_back(Idx{i, j, k});
                }
}

// PressureSolver::_calculateNegativeDivergenceVectorThread (pressuresolver.cpp:580-669).
static void assemble_rhs(const PressureInputs &in, const Rows &R, std::vector<double> &rhs) {
    const Weights &w = *in.w;
    const MAC &m = *in.mac;
    double factor = 1.0 / in.dx;
    rhs.assign(R.cell.size(), 0.0);
    for (size_t r = 0; r < R.cell.size(); r++) {
        int i = R.cell[r].i, j = R.cell[r].j, k = R.cell[r].k;
        double volCenter = w.C(i, j, k);
        double volRight = w.U(i + 1, j, k), volLeft = w.U(i, j, k);
        double volTop = w.V(i, j + 1, k), volBottom = w.V(i, j, k);
        double volFront = w.W(i, j, k + 1), volBack = w.W(i, j, k);
        double div = 0.0;
        div += -factor * volRight * m.U(i + 1, j, k);
        div += factor * volLeft * m.U(i, j, k);
        div += -factor * volTop * m.V(i, j + 1, k);
        div += factor * volBottom * m.V(i, j, k);
        div += -factor * volFront * m.W(i, j, k + 1);
        div += factor * volBack * m.W(i, j, k);
        div += factor * (volRight - volCenter) * (*in.solidU)(i + 1, j, k);
        div += -factor * (volLeft - volCenter) * (*in.solidU)(i, j, k);
        div += factor * (volTop - volCenter) * (*in.solidV)(i, j + 1, k);
        div += -factor * (volBottom - volCenter) * (*in.solidV)(i, j, k);
        div += factor * (volFront - volCenter) * (*in.solidW)(i, j, k + 1);
        div += -factor * (volBack - volCenter) * (*in.solidW)(i, j, k);
        rhs[r] = div;
    }
}

// PressureSolver::_calculateMatrixCoefficientsThread (pressuresolver.cpp:723-808).
static void assemble_matrix(const PressureInputs &in, Rows &R) {
    const Weights &w = *in.w;
    const GridF &phi = *in.liquid;
    double factor = in.dt / (in.dx * in.dx);
    double eps = 1e-9;
    size_t n = R.cell.size();
    R.diag.assign(n, 0); R.offx.assign(n, 0); R.offy.assign(n, 0); R.offz.assign(n, 0);
    R.nxp.assign(n, -1); R.nyp.assign(n, -1); R.nzp.assign(n, -1);
    R.nxm.assign(n, -1); R.nym.assign(n, -1); R.nzm.assign(n, -1);
    for (size_t r = 0; r < n; r++) {
        int i = R.cell[r].i, j = R.cell[r].j, k = R.cell[r].k;
        double vR = w.U(i + 1, j, k), vL = w.U(i, j, k), vT = w.V(i, j + 1, k),
               vB = w.V(i, j, k), vF = w.W(i, j, k + 1), vK = w.W(i, j, k);
        double pc = phi(i, j, k);
        double diag = (vR + vL + vT + vB + vF + vK) * factor;
        auto side = [&](double vol, int a, int b, int c, int &nbr, double *off) {
            double pn = phi(a, b, c);
            if (pn < 0.0) {
                int key = R.key(a, b, c);
                if (key != -1) {
                    nbr = key;
                    if (off) *off = -vol * factor;
                }
            } else {
                double theta = clampv(pn / (pc + eps), -kMaxTheta, kMaxTheta);
                diag -= vol * factor * theta;
            }
        };
        side(vR, i + 1, j, k, R.nxp[r], &R.offx[r]);
        side(vL, i - 1, j, k, R.nxm[r], nullptr);
        side(vT, i, j + 1, k, R.nyp[r], &R.offy[r]);
        side(vB, i, j - 1, k, R.nym[r], nullptr);
        side(vF, i, j, k + 1, R.nzp[r], &R.offz[r]);
        side(vK, i, j, k - 1, R.nzm[r], nullptr);
        R.diag[r] = std::max(diag, 0.0);
    }
}

// Modified incomplete Cholesky, level zero, tau = 0.97, sigma = 0.25
// (factorModifiedIncompleteColesky0, pcgsolver.h:69-185).  For the 7-point
// stencil in k, j, i row order the reference's column-oriented elimination
// reduces to a per-row recurrence over the lower neighbours taken in
// increasing row order (-z, -y, -x): for each such neighbour n with a
// nonzero original diagonal,
//     L    = A(n, row) * invdiag(n)
//     d   -= L * L
//     d   -= (tau * L) * missing(n, row)
// where missing(n, row) sums, in increasing row order, the scaled couplings
// of n to its other upper neighbours (which are never coupled to row).  The
// pivot is then replaced by A(row,row) if it fell below sigma * A(row,row)
// and inverted as 1/sqrt.  Rows with A(row,row) == 0 keep their accumulated
// value and their column is left unscaled, as in the reference.
struct MIC {
    std::vector<double> invdiag, Lx, Ly, Lz;  // Lx[n] = L(n+x, n) etc. (column n)
};

static void mic0_factor(const Rows &R, MIC &M) {
    const double tau = 0.97, sigma = 0.25;
    size_t n = R.cell.size();
    M.invdiag.assign(n, 0.0);
    M.Lx.assign(n, 0.0); M.Ly.assign(n, 0.0); M.Lz.assign(n, 0.0);
    for (size_t r = 0; r < n; r++) {
        double d = R.diag[r];
        // Contributions from lower neighbours in increasing row order.
        int lower[3] = {R.nzm[r], R.nym[r], R.nxm[r]};
        for (int q = 0; q < 3; q++) {
            int nb = lower[q];
            if (nb < 0 || R.diag[nb] == 0) continue;
            double Lcx = R.nxp[nb] >= 0 ? M.Lx[nb] : 0.0;
            double Lcy = R.nyp[nb] >= 0 ? M.Ly[nb] : 0.0;
            double Lcz = R.nzp[nb] >= 0 ? M.Lz[nb] : 0.0;
            double L;
            double missing = 0;
            if (q == 0) {         // row is nb's +z neighbour: others are +x, +y
                L = Lcz;
                if (R.nxp[nb] >= 0) missing += Lcx;
                if (R.nyp[nb] >= 0) missing += Lcy;
            } else if (q == 1) {  // row is nb's +y neighbour: others are +x, +z
                L = Lcy;
                if (R.nxp[nb] >= 0) missing += Lcx;
                if (R.nzp[nb] >= 0) missing += Lcz;
            } else {              // row is nb's +x neighbour: others are +y, +z
                L = Lcx;
                if (R.nyp[nb] >= 0) missing += Lcy;
                if (R.nzp[nb] >= 0) missing += Lcz;
            }
            d -= L * L;
            d -= tau * L * missing;
        }
        if (R.diag[r] == 0) {
            M.invdiag[r] = d;
            M.Lx[r] = R.offx[r]; M.Ly[r] = R.offy[r]; M.Lz[r] = R.offz[r];
            continue;
        }
        if (d < sigma * R.diag[r]) M.invdiag[r] = 1 / std::sqrt(R.diag[r]);
        else M.invdiag[r] = 1 / std::sqrt(d);
        M.Lx[r] = R.offx[r] * M.invdiag[r];
        M.Ly[r] = R.offy[r] * M.invdiag[r];
        M.Lz[r] = R.offz[r] * M.invdiag[r];
    }
}

// solveLower + solveLowerTransposeInPlace (pcgsolver.h:191-221) written as
// row gathers with the same operation order.
static void mic0_apply(const Rows &R, const MIC &M, const std::vector<double> &r,
                       std::vector<double> &z) {
    size_t n = R.cell.size();
    z.assign(n, 0.0);
    for (size_t c = 0; c < n; c++) {
        double q = r[c];
        if (R.nzm[c] >= 0) q -= M.Lz[R.nzm[c]] * z[R.nzm[c]];
        if (R.nym[c] >= 0) q -= M.Ly[R.nym[c]] * z[R.nym[c]];
        if (R.nxm[c] >= 0) q -= M.Lx[R.nxm[c]] * z[R.nxm[c]];
        z[c] = q * M.invdiag[c];
    }
    for (size_t cc = n; cc-- > 0;) {
        double q = z[cc];
        if (R.nxp[cc] >= 0) q -= M.Lx[cc] * z[R.nxp[cc]];
        if (R.nyp[cc] >= 0) q -= M.Ly[cc] * z[R.nyp[cc]];
        if (R.nzp[cc] >= 0) q -= M.Lz[cc] * z[R.nzp[cc]];
        z[cc] = q * M.invdiag[cc];
    }
}

// FixedSparseMatrix multiply (sparsematrix.h:187-216): row entries summed in
// increasing column order (-z, -y, -x, diagonal, +x, +y, +z).
static void spmv(const Rows &R, const std::vector<double> &x, std::vector<double> &y) {
    size_t n = R.cell.size();
    y.assign(n, 0.0);
    for (size_t c = 0; c < n; c++) {
        double s = 0;
        if (R.nzm[c] >= 0) s += R.offz[R.nzm[c]] * x[R.nzm[c]];
        if (R.nym[c] >= 0) s += R.offy[R.nym[c]] * x[R.nym[c]];
        if (R.nxm[c] >= 0) s += R.offx[R.nxm[c]] * x[R.nxm[c]];
        s += R.diag[c] * x[c];
        if (R.nxp[c] >= 0) s += R.offx[c] * x[R.nxp[c]];
        if (R.nyp[c] >= 0) s += R.offy[c] * x[R.nyp[c]];
        if (R.nzp[c] >= 0) s += R.offz[c] * x[R.nzp[c]];
        y[c] = s;
    }
}

static double dot_seq(const std::vector<double> &a, const std::vector<double> &b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
}
static double absmax(const std::vector<double> &a) {
    double m = 0;
    for (double v : a)
        if (std::fabs(v) > m) m = std::fabs(v);
    return m;
}

// Dot product with Neumaier's compensated summation (a sum whose error does
// not grow with n), for the converged FP64 error oracle.
static double dot_compensated(const std::vector<double> &a, const std::vector<double> &b) {
    double s = 0, c = 0;
    for (size_t i = 0; i < a.size(); i++) {
        double p = a[i] * b[i];
        double t = s + p;
        if (std::fabs(s) >= std::fabs(p)) c += (s - t) + p;
        else c += (p - t) + s;
        s = t;
    }
    return s + c;
}

// The converged FP64 error oracle (mode ORC_PCG_FP64_TRUE): the same MIC(0)
// preconditioned CG with alpha/beta in double and compensated dots; when the
// recursive residual reaches the tolerance the true residual b - Ax is
// formed, and the solve ends only if that one is below the tolerance too
// (otherwise CG restarts from it).  This is the reference's iteration made
// exact enough to measure other solvers' error against, not the reference's
// arithmetic.
static bool pcg_true_residual(const Rows &R, const std::vector<double> &b, std::vector<double> &x, double tolf,
                              int maxit, double &resid, int &iters) {
    size_t n = R.cell.size();
    x.assign(n, 0.0);
    std::vector<double> r = b, z, s, q, ax;
    resid = absmax(r);
    if (resid == 0) { iters = 0; return true; }
    if (tolf < 1e-30) tolf = 1e-30;
    const double tol = tolf * resid;
    MIC M;
    mic0_factor(R, M);
    int it = 0;
    while (it < maxit) {
        mic0_apply(R, M, r, z);
        double rho = dot_compensated(z, r);
        if (rho == 0 || rho != rho) break;
        s = z;
        bool restart = false;
        for (; it < maxit && !restart; it++) {
            spmv(R, s, q);
            double alpha = rho / dot_compensated(s, q);
            for (size_t i = 0; i < n; i++) x[i] += alpha * s[i];
            for (size_t i = 0; i < n; i++) r[i] -= alpha * q[i];
            if (absmax(r) <= tol) {
                spmv(R, x, ax);
                for (size_t i = 0; i < n; i++) r[i] = b[i] - ax[i];
                resid = absmax(r);
                if (resid <= tol) { iters = it + 1; return true; }
                restart = true;
                continue;
            }
            mic0_apply(R, M, r, z);
            double rhoNew = dot_compensated(z, r);
            double beta = rhoNew / rho;
            for (size_t i = 0; i < n; i++) s[i] = z[i] + beta * s[i];
            rho = rhoNew;
        }
    }
    spmv(R, x, ax);
    for (size_t i = 0; i < n; i++) r[i] = b[i] - ax[i];
    resid = absmax(r);
    iters = it;
    return resid <= tol;
}

// PCGSolver<double>::solve (pcgsolver.h:248-302).  mode 0 rounds alpha/beta
// to float as BLAS::addScaled's float parameter does (blaswrapper.h:135-158).
static bool pcg(const Rows &R, const std::vector<double> &b, std::vector<double> &x, double tolf,
                int maxit, int mode, double &resid, int &iters) {
    if (mode == ORC_PCG_FP64_TRUE) return pcg_true_residual(R, b, x, tolf, maxit, resid, iters);
    size_t n = R.cell.size();
    x.assign(n, 0.0);
    std::vector<double> r = b, z, s, q;
    resid = absmax(r);
    if (resid == 0) { iters = 0; return true; }
    if (tolf < 1e-30) tolf = 1e-30;
    double tol = tolf * resid;
    MIC M;
    mic0_factor(R, M);
    mic0_apply(R, M, r, z);
    double rho = dot_seq(z, r);
    if (rho == 0 || rho != rho) { iters = 0; return false; }
    s = z;
    int it;
    for (it = 0; it < maxit; it++) {
        spmv(R, s, z);
        double alpha = rho / dot_seq(s, z);
        if (mode == ORC_PCG_REF_EXACT) {
            float af = (float)alpha, naf = (float)(-alpha);
            for (size_t i = 0; i < n; i++) x[i] += af * s[i];
            for (size_t i = 0; i < n; i++) r[i] += naf * z[i];
        } else {
            for (size_t i = 0; i < n; i++) x[i] += alpha * s[i];
            for (size_t i = 0; i < n; i++) r[i] -= alpha * z[i];
        }
        resid = absmax(r);
        if (resid <= tol) { iters = it + 1; return true; }
        mic0_apply(R, M, r, z);
        double rhoNew = dot_seq(z, r);
        double beta = rhoNew / rho;
        if (mode == ORC_PCG_REF_EXACT) {
            float bf = (float)beta;
            for (size_t i = 0; i < n; i++) z[i] += bf * s[i];
        } else {
            for (size_t i = 0; i < n; i++) z[i] += beta * s[i];
        }
        s.swap(z);
        rho = rhoNew;
    }
    iters = it;
    return false;
}

// PressureSolver::_applySolutionToVelocityField + _applyPressureToVelocityFieldThread
// (pressuresolver.cpp:842-1047) without surface tension.
static void apply_pressure(const PressureInputs &in, const Rows &R, const std::vector<double> &p) {
    int I = in.isize, J = in.jsize, K = in.ksize;
    const GridF &phi = *in.liquid;
    const Weights &w = *in.w;
    MAC &m = *in.mac;
    Valid &v = *in.valid;
    GridF pg(I, J, K, 0.0f);
    for (size_t r = 0; r < R.cell.size(); r++) pg(R.cell[r].i, R.cell[r].j, R.cell[r].k) = (float)p[r];
    auto fluid = [&](int i, int j, int k) { return phi(i, j, k) < 0.0; };
    std::fill(v.U.a.begin(), v.U.a.end(), 0);
    std::fill(v.V.a.begin(), v.V.a.end(), 0);
    std::fill(v.W.a.begin(), v.W.a.end(), 0);
    float factor = (float)(in.dt / in.dx);
    float eps = 1e-6f;
    // Ghost-fluid pressure difference across a face between cells a and b.
    auto diff = [&](int ai, int aj, int ak, int bi, int bj, int bk) -> float {
        float p1 = 0.0f, p2 = 0.0f;
        if (fluid(ai, aj, ak) && fluid(bi, bj, bk)) {
            p1 = pg(ai, aj, ak);
            p2 = pg(bi, bj, bk);
        } else {
            float phi1 = phi(ai, aj, ak), phi2 = phi(bi, bj, bk);
            double tension = 0.0;
            if (fluid(ai, aj, ak)) {
                float tp = phi2 / (phi1 + eps);
                tp = (float)clampv((double)tp, -kMaxTheta, kMaxTheta);
                float tt = (phi1 - phi2) / (phi1 + eps);
                tt = (float)clampv((double)tt, -kMaxTheta, kMaxTheta);
                p1 = pg(ai, aj, ak);
                p2 = (float)(tt * tension + tp * p1);
            } else {
                float tp = phi1 / (phi2 + eps);
                tp = (float)clampv((double)tp, -kMaxTheta, kMaxTheta);
                float tt = (phi2 - phi1) / (phi2 + eps);
                tt = (float)clampv((double)tt, -kMaxTheta, kMaxTheta);
                p2 = pg(bi, bj, bk);
                p1 = (float)(tt * tension + tp * p2);
            }
        }
        return p2 - p1;
    };
    // U faces (skip i == 0 and i == I-1, pressuresolver.cpp:918-920).
    for (int k = 0; k < K; k++)
        for (int j = 0; j < J; j++)
            for (int i = 0; i < I + 1; i++) {
                if (i == 0 || i == I - 1) continue;
                bool border = (i == I) ? fluid(i - 1, j, k) : (fluid(i, j, k) || fluid(i - 1, j, k));
                if (w.U(i, j, k) > 0 && border) {
                    m.U(i, j, k) += (float)(-factor * diff(i - 1, j, k, i, j, k));
                    v.U(i, j, k) = 1;
                } else {
                    m.U(i, j, k) = 0.0f;
                }
            }
    for (int k = 0; k < K; k++)
        for (int j = 0; j < J + 1; j++)
            for (int i = 0; i < I; i++) {
                if (j == 0 || j == J - 1) continue;
                bool border = (j == J) ? fluid(i, j - 1, k) : (fluid(i, j, k) || fluid(i, j - 1, k));
                if (w.V(i, j, k) > 0 && border) {
                    m.V(i, j, k) += (float)(-factor * diff(i, j - 1, k, i, j, k));
                    v.V(i, j, k) = 1;
                } else {
                    m.V(i, j, k) = 0.0f;
                }
            }
    for (int k = 0; k < K + 1; k++)
        for (int j = 0; j < J; j++)
            for (int i = 0; i < I; i++) {
                if (k == 0 || k == K - 1) continue;
                bool border = (k == K) ? fluid(i, j, k - 1) : (fluid(i, j, k) || fluid(i, j, k - 1));
                if (w.W(i, j, k) > 0 && border) {
                    m.W(i, j, k) += (float)(-factor * diff(i, j, k - 1, i, j, k));
                    v.W(i, j, k) = 1;
                } else {
                    m.W(i, j, k) = 0.0f;
                }
            }
}

// PressureSolver::solve (pressuresolver.cpp:44-77) + _solveLinearSystem (:810-840).
static PressureResult pressure_solve(const PressureInputs &in) {
    PressureResult res;
    Rows R;
    build_rows(in, R);
    res.n = (int)R.cell.size();
    condition_solid_velocity(in);
    assemble_rhs(in, R, res.rhs);
    double maxAbs = 0.0;
    for (double v : res.rhs)
        if (std::fabs(v) > maxAbs) maxAbs = std::fabs(v);
    res.soln.assign(R.cell.size(), 0.0);
    if (maxAbs < in.tol) {
        res.iters = 0; res.err = 0.0; res.ok = 1;
        return res;
    }
    assemble_matrix(in, R);
    double err;
    int it;
    bool success = pcg(R, res.rhs, res.soln, in.tol, in.maxiter, in.mode, err, it);
    res.iters = it; res.err = err;
    res.ok = (success || (it == in.maxiter && err < 1.0)) ? 1 : 0;
    if (res.ok) {
        apply_pressure(in, R, res.soln);
        res.solved = true;
    }
    return res;
}

// ---------------------------------------------------------------------------
// Simulation driver (FluidSimulation, fluidsimulation.cpp)
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Mesh inflow sources: MeshFluidSource with a static mesh
// (meshfluidsource.cpp:41-44, 166-233; MeshObject::updateMeshStatic,
// meshobject.cpp:50-59), inflow, without appended object velocity.
// ---------------------------------------------------------------------------
struct TriMesh {
    std::vector<V3> v;
    std::vector<int> t;  // three vertex indices per triangle
    size_t ntri() const { return t.size() / 3; }
};

static inline V3 cross(V3 a, V3 b) {  // vmath.h:80-84
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// AABB(std::vector<vec3>) (aabb.cpp:56-80): float corner, extents + 1e-9.
static Box points_box(const std::vector<V3> &pts) {
    double mnx = pts[0].x, mny = pts[0].y, mnz = pts[0].z;
    double mxx = mnx, mxy = mny, mxz = mnz;
    for (const V3 &p : pts) {
        mnx = std::fmin((double)p.x, mnx); mny = std::fmin((double)p.y, mny);
        mnz = std::fmin((double)p.z, mnz); mxx = std::fmax((double)p.x, mxx);
        mxy = std::fmax((double)p.y, mxy); mxz = std::fmax((double)p.z, mxz);
    }
    Box b;
    b.pos = V3((float)mnx, (float)mny, (float)mnz);
    b.w = mxx - mnx + 1e-9; b.h = mxy - mny + 1e-9; b.d = mxz - mnz + 1e-9;
    return b;
}

// Collision::lineIntersectsTriangle (collision.cpp:79-118) for the +z
// direction; *z receives the crossing's z (p + (float)t * dir).
static bool ray_hit_z(V3 p, V3 v0, V3 v1, V3 v2, float *z) {
    const V3 dir(0.0f, 0.0f, 1.0f);
    V3 e1 = v1 - v0, e2 = v2 - v0;
    V3 h = cross(dir, e2);
    float a = dot(e1, h);
    double eps = 10e-9;
    if (std::fabs(a) < eps) return false;
    float f = 1 / a;
    V3 s = p - v0;
    float u = f * dot(s, h);
    if (u < 0.0 || u > 1.0) return false;
    V3 q = cross(s, e1);
    float v = f * dot(dir, q);
    if (v < 0.0 || u + v > 1.0) return false;
    double t = f * dot(e2, q);
    *z = p.z + (float)t * dir.z;
    return true;
}

// MeshUtils::getCellsInsideTriangleMesh (meshutils.cpp:99-214) on an
// ni x nj x nk cell grid.  Each (i, j) column casts a +z ray from its
// jittered cell centre at k = -1; a cell is inside when its column has an
// even number of crossings of which an odd number lie below the cell centre.
// The reference first bins triangles into the columns their bounding boxes
// cover; a crossing of a column's ray lies inside that column, so testing
// every triangle finds the same crossings.
static void cells_inside(const TriMesh &m, int ni, int nj, int nk, double dx, V3 jitter,
                         GridB &inside) {
    inside = GridB(ni, nj, nk, 0);
    std::vector<float> zs;
    for (int j = 0; j < nj; j++)
        for (int i = 0; i < ni; i++) {
            V3 gp = index_to_center(i, j, -1, dx) + jitter;
            zs.clear();
            for (size_t t = 0; t < m.ntri(); t++) {
                float z;
                if (ray_hit_z(gp, m.v[m.t[3 * t]], m.v[m.t[3 * t + 1]], m.v[m.t[3 * t + 2]], &z))
                    zs.push_back(z);
            }
            if (zs.size() % 2 != 0) continue;
            for (int k = 0; k < nk; k++) {
                float zc = index_to_center(i, j, k, dx).z;
                int below = 0;
                for (float z : zs) below += z < zc;
                if (below % 2 == 1) inside(i, j, k) = 1;
            }
        }
}

// MeshUtils::getGridNodesInsideTriangleMesh(mesh, dx, Array3d<bool>)
// (meshutils.cpp:347-420) for meshes inside the node grid's box -- the only
// case handled; false otherwise (the reference then splits the mesh).
static bool nodes_inside(const TriMesh &m0, double dx, int nni, int nnj, int nnk, V3 jitter,
                         GridB &nodes) {
    nodes = GridB(nni, nnj, nnk, 0);
    Box grid;
    grid.w = (nni - 1) * dx; grid.h = (nnj - 1) * dx; grid.d = (nnk - 1) * dx;
    for (const V3 &p : m0.v)
        if (!grid.inside(p)) return false;
    if (m0.v.empty()) return true;
    // getGridNodesInsideTriangleMesh(mesh, dx, std::vector) (meshutils.cpp:347-372):
    // nodes are the cell centres of the grid shifted by half a cell.
    TriMesh m = m0;
    float hw = (float)(0.5 * dx);
    for (V3 &p : m.v) p = p + V3(hw, hw, hw);
    Box bb = points_box(m.v);
    Idx goff = pos_to_index(bb.pos, dx);
    V3 off = index_to_pos(goff.i, goff.j, goff.k, dx);
    for (V3 &p : m.v) p = p - off;
    bb.pos = bb.pos - off;
    Idx gmax = pos_to_index(bb.pos + V3((float)bb.w, (float)bb.h, (float)bb.d), dx);
    GridB local;
    cells_inside(m, gmax.i + 1, gmax.j + 1, gmax.k + 1, dx, jitter, local);
    for (int k = 0; k < local.nk; k++)
        for (int j = 0; j < local.nj; j++)
            for (int i = 0; i < local.ni; i++)
                if (local(i, j, k) && nodes.in(i + goff.i, j + goff.j, k + goff.k))
                    nodes(i + goff.i, j + goff.j, k + goff.k) = 1;
    return true;
}

// Exact-band unsigned distances of a triangle list on a node grid, each node
// taking its minimum in the frame of the 10^3 block that owns it
// (meshlevelset.cpp:781-876, 1115-1163); other nodes keep `upper`.
static void exact_band(const std::vector<Tri> &tris, GridF &phi, double dx, float upper) {
    const int bw = 10;  // meshlevelset.h:370 (_blockwidth)
    phi.fill(upper);
    for (int k = 0; k < phi.nk; k++)
        for (int j = 0; j < phi.nj; j++)
            for (int i = 0; i < phi.ni; i++) {
                int bi = i / bw, bj = j / bw, bk = k / bw;
                V3 boff = index_to_pos(bi, bj, bk, bw * dx);
                V3 gpos = index_to_pos(i - bi * bw, j - bj * bw, k - bk * bw, dx);
                float best = upper;
                for (const Tri &t : tris) {
                    if (i < t.gmin.i || j < t.gmin.j || k < t.gmin.k || i > t.gmax.i ||
                        j > t.gmax.j || k > t.gmax.k)
                        continue;
                    V3 cp = closest_point_on_triangle(gpos, t.v[0] - boff, t.v[1] - boff,
                                                      t.v[2] - boff);
                    float dist = length(cp - gpos);
                    if (dist < best) best = dist;
                }
                phi(i, j, k) = best;
            }
}

// MeshLevelSet::fastCalculateSignedDistanceField (meshlevelset.cpp:577-601)
// without velocity data: exact band, then negated inside the mesh
// (_computeDistanceFieldSigns, meshlevelset.cpp:1298-1317).  phi must be
// sized to the node grid already.  False if the mesh leaves the grid box.
static bool mesh_sdf(const TriMesh &m, double dx, int band, V3 jitter, GridF &phi) {
    int isize = phi.ni - 1, jsize = phi.nj - 1, ksize = phi.nk - 1;
    float upper = (float)((phi.ni + phi.nj + phi.nk) * dx);  // meshlevelset.cpp:769-771
    std::vector<Tri> tris;
    for (size_t t = 0; t < m.ntri(); t++) {
        Tri T;
        for (int c = 0; c < 3; c++) T.v[c] = m.v[m.t[3 * t + c]];
        if (tri_band(T, isize, jsize, ksize, dx, band)) tris.push_back(T);
    }
    exact_band(tris, phi, dx, upper);
    GridB nodes;
    if (!nodes_inside(m, dx, phi.ni, phi.nj, phi.nk, jitter, nodes)) return false;
    for (size_t q = 0; q < phi.a.size(); q++)
        if (nodes.a[q]) phi.a[q] = -phi.a[q];
    return true;
}

struct InflowSource {
    TriMesh mesh;
    V3 velocity;              // setVelocity (meshfluidsource.cpp:114-116)
    bool enabled = true;      // meshfluidsource.h
    bool constrain = true;    // _isConstrainedFluidVelocity (meshfluidsource.h:124)
    int emissions = 1;        // _substepEmissions (meshfluidsource.h:119)
    bool outflow = false;     // setOutflow() (meshfluidsource.cpp:82-84)
    bool inversed = false;    // outflowInverse() (meshfluidsource.cpp:158-160)
    bool ready = false;       // static mesh: the SDF and cells are built once
    Idx gridOffset;           // _sourceSDFGridOffset
    V3 sdfOffset;             // _sourceSDFOffset
    GridF sdf;                // _sourceSDF, nodal on the source sub-grid
    std::vector<Idx> cells;   // MeshObject::getCells
    GridB isCell;
};

// Splitmix64, the generator behind the inflow jitter: the reference draws
// rand() (fluidsimulation.h:1501-1503, meshutils.cpp:95-97), from several
// threads at once; the value of a draw here is keyed by what it jitters.
static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline double draw_unit(uint64_t seed, uint64_t stream, uint64_t item) {
    uint64_t h = splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + item);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
// The ray jitter of one inside test (meshutils.cpp:107-110):
// _randomDouble(jit, -jit) per axis with jit = 0.001 dx.  Stream 0 keys the
// source SDF's signs, stream 1 the source cells.
static V3 ray_jitter(uint64_t seed, int source, int which, double dx) {
    double jit = 0.001 * dx;
    uint64_t st = ((uint64_t)(source + 1) << 8) | (uint64_t)which;
    double r[3];
    for (int a = 0; a < 3; a++) r[a] = jit + draw_unit(seed, st, a) * (-jit - jit);
    return V3((float)r[0], (float)r[1], (float)r[2]);
}

// MeshFluidSource::update (meshfluidsource.cpp:178-222) with
// _getGridBoundsFromTriangleMesh (:318-333, grid pad 4, meshfluidsource.h:
// 133) and exact band 2 (meshfluidsource.h:116), plus the source cells of
// MeshObject::getCells (meshobject.cpp:101-140).  The mesh must lie inside
// the domain grid's box.
static bool source_build(InflowSource &S, int I, int J, int K, double dx, uint64_t seed, int id) {
    Box bb = points_box(S.mesh.v);
    bb.expand(4.0 * dx);
    Idx gmin = pos_to_index(bb.pos, dx), gmax = pos_to_index(bb.maxp(), dx);
    gmin.i = (int)std::fmax(gmin.i, 0); gmin.j = (int)std::fmax(gmin.j, 0);
    gmin.k = (int)std::fmax(gmin.k, 0);
    gmax.i = (int)std::fmin(gmax.i, I - 1); gmax.j = (int)std::fmin(gmax.j, J - 1);
    gmax.k = (int)std::fmin(gmax.k, K - 1);
    S.gridOffset = gmin;
    S.sdfOffset = index_to_pos(gmin.i, gmin.j, gmin.k, dx);
    int isdf = std::max(gmax.i - gmin.i + 1, 1), jsdf = std::max(gmax.j - gmin.j + 1, 1),
        ksdf = std::max(gmax.k - gmin.k + 1, 1);
    TriMesh local = S.mesh;
    V3 neg = V3(-S.sdfOffset.x, -S.sdfOffset.y, -S.sdfOffset.z);
    for (V3 &p : local.v) p = p + neg;  // TriangleMesh::translate (trianglemesh.cpp:732-736)
    S.sdf = GridF(isdf + 1, jsdf + 1, ksdf + 1, 0.0f);
    if (!mesh_sdf(local, dx, 2, ray_jitter(seed, id, 0, dx), S.sdf)) return false;
    GridB nodes;
    if (!nodes_inside(S.mesh, dx, I + 1, J + 1, K + 1, ray_jitter(seed, id, 1, dx), nodes))
        return false;
    S.isCell = GridB(I, J, K, 0);
    for (int k = 0; k < nodes.nk; k++)
        for (int j = 0; j < nodes.nj; j++)
            for (int i = 0; i < nodes.ni; i++) {
                if (!nodes(i, j, k)) continue;
                for (int c = 0; c < 8; c++) {  // Grid3d::getVertexGridIndexNeighbours
                    int a = i - ((c & 1) ? 1 : 0), b = j - ((c & 2) ? 1 : 0), d = k - ((c & 4) ? 1 : 0);
                    if (S.isCell.in(a, b, d)) S.isCell(a, b, d) = 1;
                }
            }
    S.cells.clear();
    for (int k = 0; k < K; k++)
        for (int j = 0; j < J; j++)
            for (int i = 0; i < I; i++)
                if (S.isCell(i, j, k)) S.cells.push_back(Idx{i, j, k});
    S.ready = true;
    return true;
}

// ParticleMaskGrid (particlemaskgrid.cpp:54-105): one bit per half-cell
// subcell, set from the particle's cell and its subcell's parities.
struct MaskGrid {
    GridB m;
    double dx = 0;
    static int bit(Idx s) {
        int c = 0;
        if (s.i % 2 == 1) c |= 1;
        if (s.j % 2 == 1) c |= 2;
        if (s.k % 2 == 1) c |= 4;
        return 1 << c;
    }
    void add(V3 p) {
        Idx s = pos_to_index(p, 0.5 * dx);
        Idx g = pos_to_index(p, dx);
        m(g.i, g.j, g.k) |= (uint8_t)bit(s);
    }
    bool is_set(V3 p) const {
        Idx s = pos_to_index(p, 0.5 * dx);
        return (m(s.i / 2, s.j / 2, s.k / 2) & bit(s)) != 0;
    }
};

// ---------------------------------------------------------------------------
// Static mesh obstacles: MeshObject::getMeshLevelSet (meshobject.cpp:239-262)
// for meshes of fewer than 25 islands, where the whole mesh is one level set
// on the sub-grid around its box (_getMeshIslandLevelSet, meshobject.cpp:
// 519-553), unioned into the solid SDF (MeshLevelSet::calculateUnion,
// meshlevelset.cpp:603-638, 1758-1790).
// ---------------------------------------------------------------------------
struct Obstacle {
    TriMesh mesh;          // referenced vertices only (removeExtraneousVertices)
    bool enabled = true;
    bool inversed = false; // MeshObject::inverse() (meshobject.cpp:293-295)
    bool hasJitter = false;
    V3 jitter;             // ray jitter of its inside test, if given
};

// TriangleMesh::removeExtraneousVertices (trianglemesh.cpp): keeps the
// vertices some triangle uses, in their order.
static TriMesh referenced_only(const TriMesh &m) {
    std::vector<int> map(m.v.size(), -1);
    for (int idx : m.t) map[idx] = 0;
    TriMesh out;
    for (size_t q = 0; q < m.v.size(); q++)
        if (map[q] == 0) { map[q] = (int)out.v.size(); out.v.push_back(m.v[q]); }
    for (int idx : m.t) out.t.push_back(map[idx]);
    return out;
}

// Islands of a mesh: components of vertices joined by triangle edges
// (MeshUtils::_splitIntoMeshIslands, meshutils.cpp:455-563).
static int island_count(const TriMesh &m) {
    std::vector<int> parent(m.v.size());
    for (size_t q = 0; q < parent.size(); q++) parent[q] = (int)q;
    auto find = [&](int a) {
        while (parent[a] != a) a = parent[a] = parent[parent[a]];
        return a;
    };
    for (size_t t = 0; t < m.ntri(); t++) {
        int a = find(m.t[3 * t]), b = find(m.t[3 * t + 1]), c = find(m.t[3 * t + 2]);
        parent[b] = a;
        parent[find(c)] = find(a);
    }
    int n = 0;
    for (size_t q = 0; q < parent.size(); q++) n += find((int)q) == (int)q;
    return n;
}

// MeshLevelSet::_initializeTriangleData with a grid offset (meshlevelset.cpp:
// 878-915): the triangle's box is taken from its untranslated vertices and
// then moved by the offset; the vertices are moved likewise.
static bool tri_band_offset(const V3 orig[3], V3 off, int isize, int jsize, int ksize, double dx,
                            int band, Tri &T) {
    Tri tmp;
    for (int c = 0; c < 3; c++) tmp.v[c] = orig[c];
    double mnx = tmp.v[0].x, mny = tmp.v[0].y, mnz = tmp.v[0].z, mxx = mnx, mxy = mny, mxz = mnz;
    for (int c = 0; c < 3; c++) {
        mnx = std::fmin((double)tmp.v[c].x, mnx); mny = std::fmin((double)tmp.v[c].y, mny);
        mnz = std::fmin((double)tmp.v[c].z, mnz); mxx = std::fmax((double)tmp.v[c].x, mxx);
        mxy = std::fmax((double)tmp.v[c].y, mxy); mxz = std::fmax((double)tmp.v[c].z, mxz);
    }
    Box tb;
    tb.pos = V3((float)mnx, (float)mny, (float)mnz) - off;
    tb.w = mxx - mnx + 1e-9; tb.h = mxy - mny + 1e-9; tb.d = mxz - mnz + 1e-9;
    V3 pmax = tb.maxp();
    Idx gmin = pos_to_index(tb.pos, dx), gmax = pos_to_index(pmax, dx);
    gmin.i -= band; gmin.j -= band; gmin.k -= band;
    gmax.i += band + 1; gmax.j += band + 1; gmax.k += band + 1;
    if (gmax.i < 0 || gmin.i >= isize + 1 || gmax.j < 0 || gmin.j >= jsize + 1 || gmax.k < 0 ||
        gmin.k >= ksize + 1)
        return false;
    gmin.i = std::max(gmin.i, 0); gmin.j = std::max(gmin.j, 0); gmin.k = std::max(gmin.k, 0);
    gmax.i = std::min(gmax.i, isize); gmax.j = std::min(gmax.j, jsize); gmax.k = std::min(gmax.k, ksize);
    for (int c = 0; c < 3; c++) T.v[c] = orig[c] - off;
    T.gmin = gmin; T.gmax = gmax;
    return true;
}

// _getMeshIslandLevelSet for the whole mesh (meshobject.cpp:519-553): the
// sub-grid is the mesh box widened by the band, clamped to the domain; gmin
// receives its offset.  Returns 0 when the sub-grid is empty (nothing to
// union), -1 when the mesh leaves the sub-grid's box, 1 otherwise.
static int obstacle_sdf(const TriMesh &m, int I, int J, int K, double dx, int band, V3 jitter,
                        Idx &gmin, GridF &phi) {
    Box bb = points_box(m.v);
    gmin = pos_to_index(bb.pos, dx);
    Idx gmax = pos_to_index(bb.maxp(), dx);
    gmin.i = (int)std::fmax(gmin.i - band, 0); gmin.j = (int)std::fmax(gmin.j - band, 0);
    gmin.k = (int)std::fmax(gmin.k - band, 0);
    gmax.i = (int)std::fmin(gmax.i + band + 1, I - 1); gmax.j = (int)std::fmin(gmax.j + band + 1, J - 1);
    gmax.k = (int)std::fmin(gmax.k + band + 1, K - 1);
    int gw = gmax.i - gmin.i, gh = gmax.j - gmin.j, gd = gmax.k - gmin.k;
    if (gw <= 0 || gh <= 0 || gd <= 0) return 0;
    V3 off = index_to_pos(gmin.i, gmin.j, gmin.k, dx);  // MeshLevelSet::setGridOffset
    phi = GridF(gw + 1, gh + 1, gd + 1, 0.0f);
    float upper = (float)((phi.ni + phi.nj + phi.nk) * dx);
    std::vector<Tri> tris;
    for (size_t t = 0; t < m.ntri(); t++) {
        V3 v[3] = {m.v[m.t[3 * t]], m.v[m.t[3 * t + 1]], m.v[m.t[3 * t + 2]]};
        Tri T;
        if (tri_band_offset(v, off, gw, gh, gd, dx, band, T)) tris.push_back(T);
    }
    exact_band(tris, phi, dx, upper);
    TriMesh local = m;
    V3 neg(-off.x, -off.y, -off.z);
    for (V3 &p : local.v) p = p + neg;
    GridB nodes;
    if (!nodes_inside(local, dx, phi.ni, phi.nj, phi.nk, jitter, nodes)) return -1;
    for (size_t q = 0; q < phi.a.size(); q++)
        if (nodes.a[q]) phi.a[q] = -phi.a[q];
    return 1;
}

// Union into the solid SDF: the smaller value wins on the sub-grid's nodes.
static void union_into(GridF &solid, const GridF &phi, Idx go) {
    for (int k = 0; k < phi.nk; k++)
        for (int j = 0; j < phi.nj; j++)
            for (int i = 0; i < phi.ni; i++) {
                int a = i + go.i, b = j + go.j, c = k + go.k;
                if (!solid.in(a, b, c)) continue;
                if (phi(i, j, k) < solid(a, b, c)) solid(a, b, c) = phi(i, j, k);
            }
}

// Surface reconstruction settings and output (see the surface section below).
struct SurfaceParams {
    int subdivisions = 1;                // fluidsimulation.h:1605
    double smoothValue = 0.5;            // fluidsimulation.h:1607
    int smoothIterations = 2;            // fluidsimulation.h:1608
    double particleScale = 3.0;          // fluidsimulation.h:1611 (_markerParticleScale)
    double obstacleOffset = 0.0;         // fluidsimulation.h:1593, in cells
    bool removeNearDomain = false;       // fluidsimulation.h:1594
    int removeNearDomainDistance = 0;    // fluidsimulation.h:1595
    bool invertContactNormals = false;   // fluidsimulation.h:1588
    double contactThreshold = 0.08;      // fluidsimulation.h:1591, in cells
};

struct SurfaceMesh {
    std::vector<V3> verts;
    std::vector<int> tris;  // three vertex indices per triangle
};

struct Sim {
    int I, J, K;
    double dx;
    double cfl = 5.0;                 // fluidsimulation.h:1545
    int minSteps = 1, maxSteps = 6;   // fluidsimulation.h:1541-1542
    double picRatio = 0.05;           // fluidsimulation.h:1678
    int pcgMode = ORC_PCG_REF_EXACT;
    double pcgTol = 1e-9;             // fluidsimulation.h:1657
    int pcgMaxIter = 1000;            // fluidsimulation.h:1659
    std::vector<V3> bodyForces;
    std::vector<V3> pos, vel;
    std::vector<V3> loadPos, loadVel;  // loadMarkerParticleData queue (applied by initialize)
    V3 domainOffset;                   // fluidsimulation.h:1612-1613
    double domainScale = 1.0;
    bool initialized = false, solidUpToDate = false, weightsUpToDate = false;
    bool outputInitialized = false;
    int currentFrame = 0, stepNumber = 0, lastSubsteps = 0;
    double frameDt = 0, remaining = 0, curStep = 0;
    double radius = 0;
    GridF liquid, solid;
    GridF solidU, solidV, solidW;
    Weights w;
    MAC mac, saved;
    Valid valid;
    Grid<uint8_t> nearSolid;
    double nearSolidCellSize = 0;
    int lastIters = 0, lastOk = 1;
    double lastErr = 0;
    // Count grid of the last removal pass (fluidsimulation.cpp:4325-4343) and
    // how many particles that pass removed for their speed.
    Grid<int> removalCount;
    int removedBySpeed = 0, removedBySpeedTotal = 0;
    // Mesh inflow sources (FluidSimulation::addMeshFluidSource,
    // fluidsimulation.cpp:1953-1965) and the marker-particle jitter
    // (fluidsimulation.h:1553-1555).
    std::vector<InflowSource> sources;
    double jitterFactor = 1.0;
    bool jitterSurface = false;
    uint64_t seed = 0;
    uint64_t emissions = 0;  // inflow emission passes so far (keys the jitter)
    std::vector<Obstacle> obstacles;  // FluidSimulation::addMeshObstacle (fluidsimulation.cpp:1994-2008)
    // Surface reconstruction (fluidsimulation.h:1586, 1605-1611): the mesh
    // of the particles after the first substep of each frame.
    bool surfaceEnabled = false;
    SurfaceParams surface;
    SurfaceMesh isomesh;

    V3 body_force() const {  // fluidsimulation.cpp:3363-3370
        V3 bf;
        for (const V3 &f : bodyForces) bf = bf + f;
        return bf;
    }
};

// FluidSimulation::_updateNearSolidGrid (fluidsimulation.cpp:3094-3126).
static void update_near_solid(Sim &S) {
    S.nearSolidCellSize = 3 * S.dx;
    int gi = (int)std::ceil((double)S.I * S.dx / S.nearSolidCellSize);
    int gj = (int)std::ceil((double)S.J * S.dx / S.nearSolidCellSize);
    int gk = (int)std::ceil((double)S.K * S.dx / S.nearSolidCellSize);
    S.nearSolid = GridB(gi, gj, gk, 0);
    float maxd = (float)(3 * S.dx);
    for (int k = 0; k < S.K; k++)
        for (int j = 0; j < S.J; j++)
            for (int i = 0; i < S.I; i++)
                if (std::abs(S.solid(i, j, k)) < maxd) S.nearSolid(i / 3, j / 3, k / 3) = 1;
    int layers = (int)std::ceil((float)S.cfl / (float)3);
    const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int l = 0; l < layers; l++) {  // GridUtils::featherGrid6 (gridutils.cpp:230-265)
        GridB src = S.nearSolid;
        for (int k = 0; k < gk; k++)
            for (int j = 0; j < gj; j++)
                for (int i = 0; i < gi; i++) {
                    if (!src(i, j, k)) continue;
                    for (int q = 0; q < 6; q++) {
                        int a = i + nb[q][0], b = j + nb[q][1], c = k + nb[q][2];
                        if (S.nearSolid.in(a, b, c)) S.nearSolid(a, b, c) = 1;
                    }
                }
    }
}

// FluidSimulation::_resolveCollision (fluidsimulation.cpp:4221-4296).
static V3 resolve_collision(const Sim &S, V3 oldp, V3 newp, const Box &boundary) {
    Idx g = pos_to_index(newp, S.dx);
    if (!(g.i >= 0 && g.j >= 0 && g.k >= 0 && g.i < S.I && g.j < S.J && g.k < S.K))
        newp = boundary.nearest_inside(newp);
    Idx og = pos_to_index(oldp, S.nearSolidCellSize), ng = pos_to_index(newp, S.nearSolidCellSize);
    auto near = [&](Idx c) { return S.nearSolid.in(c.i, c.j, c.k) && S.nearSolid(c.i, c.j, c.k); };
    if (!near(og) && !near(ng)) return newp;
    float eps = 1e-6f;
    float stepDistance = 0.5f * (float)S.dx;  // fluidsimulation.h:1691
    float travel = length(newp - oldp);
    if (travel < eps) return newp;
    int numSteps = (int)std::ceil(travel / stepDistance);
    V3 stepdir = normalize(newp - oldp);
    V3 last = oldp, cur;
    bool found = false;
    float collisionPhi = 0.0f;
    for (int s = 0; s < numSteps; s++) {
        if (s == numSteps - 1) cur = newp;
        else cur = oldp + ((float)(s + 1) * stepDistance) * stepdir;
        float phi = grid_trilinear(S.solid, cur, S.dx);
        if (phi < 0.0f || !boundary.inside(cur)) {
            collisionPhi = phi;
            found = true;
            break;
        }
        last = cur;
    }
    if (!found) return newp;
    V3 resolved;
    float maxResolved = (float)(S.cfl * S.dx);
    V3 grad = grid_gradient(S.solid, cur, S.dx);
    if (length(grad) > eps) {
        grad = normalize(grad);
        resolved = cur - (float)(collisionPhi - 0.1f * S.dx) * grad;
        float rphi = grid_trilinear(S.solid, resolved, S.dx);
        float rdist = length(resolved - cur);
        if (rphi < 0 || rdist > maxResolved) resolved = last;
    } else {
        resolved = last;
    }
    if (!boundary.inside(resolved)) {
        V3 orig = resolved;
        resolved = boundary.nearest_inside(resolved);
        float rphi = grid_trilinear(S.solid, resolved, S.dx);
        float rdist = length(resolved - orig);
        if (rphi < 0.0f || rdist > maxResolved) resolved = last;
    }
    return resolved;
}

// FluidSimulation::_RK3 (fluidsimulation.cpp:4191-4198).
static V3 rk3(const MAC &m, V3 p0, double dt, double dx, int I, int J, int K) {
    V3 k1 = mac_velocity(m, p0, dx, I, J, K);
    V3 k2 = mac_velocity(m, p0 + (float)(0.5 * dt) * k1, dx, I, J, K);
    V3 k3 = mac_velocity(m, p0 + (float)(0.75 * dt) * k2, dx, I, J, K);
    return p0 + (float)(dt / 9.0f) * (2.0f * k1 + 3.0f * k2 + 4.0f * k3);
}

// FluidSimulation::_getMarkerParticleSpeedLimit (fluidsimulation.cpp:4298-4322).
static float speed_limit(const Sim &S, double dt) {
    double step = S.cfl * S.dx / dt;
    std::vector<int> counts(S.maxSteps, 0);
    for (const V3 &v : S.vel) {
        double speed = (double)length(v);
        int idx = (int)std::fmin(std::floor(speed / step), S.maxSteps - 1);
        counts[idx]++;
    }
    double maxpct = 0.0005;  // fluidsimulation.h:1689
    int maxabs = 35;         // fluidsimulation.h:1690
    int maxRemoval = (int)std::fmin((int)((double)S.vel.size() * maxpct), maxabs);
    double maxspeed = S.maxSteps * step;
    int cur = 0;
    for (int i = (int)counts.size() - 1; i > 0; i--) {
        if (cur + counts[i] > maxRemoval) break;
        cur += counts[i];
        maxspeed = i * step;
    }
    return (float)maxspeed;
}

// FluidSimulation::_advanceMarkerParticles + _removeMarkerParticles
// (fluidsimulation.cpp:4200-4391).
static void advance_particles(Sim &S, double dt) {
    Box boundary = boundary_box(S.I, S.J, S.K, S.dx);
    boundary.expand(-0.1f * S.dx);  // fluidsimulation.h:1685 (_solidBufferWidth)
    std::vector<V3> out(S.pos.size());
    for (size_t i = 0; i < S.pos.size(); i++) {
        out[i] = rk3(S.mac, S.pos[i], dt, S.dx, S.I, S.J, S.K);
        out[i] = resolve_collision(S, S.pos[i], out[i], boundary);
    }
    S.pos = out;
    // Removal (fluidsimulation.cpp:4324-4353) uses the frame step.
    float maxspeed = speed_limit(S, S.frameDt);
    double maxsq = maxspeed * maxspeed;
    Grid<int> &count = S.removalCount;
    count = Grid<int>(S.I, S.J, S.K, 0);
    std::vector<uint8_t> removed(S.pos.size(), 0);
    for (size_t i = 0; i < S.pos.size(); i++)
        removed[i] = grid_trilinear(S.solid, S.pos[i], S.dx) < 0.0f;
    S.removedBySpeed = 0;
    for (size_t i = 0; i < S.pos.size(); i++) {
        if (removed[i]) continue;
        Idx g = pos_to_index(S.pos[i], S.dx);
        if (count(g.i, g.j, g.k) >= 250) { removed[i] = 1; continue; }
        count(g.i, g.j, g.k) += 1;
        if (dot(S.vel[i], S.vel[i]) > maxsq) {
            removed[i] = 1;
            S.removedBySpeed++;
            S.removedBySpeedTotal++;
            continue;
        }
    }
    size_t cur = 0;
    for (size_t i = 0; i < S.pos.size(); i++) {
        if (!removed[i]) { S.pos[cur] = S.pos[i]; S.vel[cur] = S.vel[i]; cur++; }
    }
    S.pos.resize(cur);
    S.vel.resize(cur);
}

// FluidSimulation::_constrainVelocityFieldThread (fluidsimulation.cpp:3873-3935)
// with the domain boundary's friction (0 by default, fluidsimulation.h:1524).
static void constrain(MAC &m, const Sim &S) {
    const float f = 0.0f;
    auto con = [&](GridF &g, const GridF &w, const GridF &sv) {
        for (size_t q = 0; q < g.a.size(); q++) {
            if (w.a[q] == 0) g.a[q] = sv.a[q];
            else if (w.a[q] < 1.0f) g.a[q] = f * sv.a[q] + (1.0f - f) * g.a[q];
        }
    };
    con(m.U, S.w.U, S.solidU);
    con(m.V, S.w.V, S.solidV);
    con(m.W, S.w.W, S.solidW);
}

// FluidSimulation::_addNewFluidCells / _addNewFluidCellsThread
// (fluidsimulation.cpp:4397-4431, 4523-4562) for one emission pass of one
// source: the eight subcell centres of every source cell inside the source
// SDF, jittered where deeper than a cell (or always with surface jitter),
// kept where the solid SDF is positive and the subcell is still empty.
static void emit_source(Sim &S, InflowSource &src, MaskGrid &mask) {
    const double dx = S.dx;
    const double q = 0.25 * dx;
    const V3 offs[8] = {V3((float)-q, (float)-q, (float)-q), V3((float)q, (float)-q, (float)-q),
                        V3((float)-q, (float)q, (float)-q),  V3((float)q, (float)q, (float)-q),
                        V3((float)-q, (float)-q, (float)q),  V3((float)q, (float)-q, (float)q),
                        V3((float)-q, (float)q, (float)q),   V3((float)q, (float)q, (float)q)};
    const double jitter = 0.25 * (S.jitterFactor - 1e-3) * dx;  // fluidsimulation.cpp:2623-2626
    const uint64_t stream = 0x1000000ull + S.emissions++;
    std::vector<V3> cand;
    for (size_t c = 0; c < src.cells.size(); c++) {
        const Idx g = src.cells[c];
        const V3 ctr = index_to_center(g.i, g.j, g.k, dx);
        for (int o = 0; o < 8; o++) {
            V3 p = ctr + offs[o];
            double d = grid_trilinear(src.sdf, p - src.sdfOffset, dx);
            if (d > 0) continue;
            if (S.jitterSurface || d < -dx) {
                // _jitterMarkerParticlePosition (fluidsimulation.cpp:2628-2635)
                uint64_t item = 3 * (8 * (uint64_t)c + o);
                p.x += -jitter + draw_unit(S.seed, stream, item) * (jitter - -jitter);
                p.y += -jitter + draw_unit(S.seed, stream, item + 1) * (jitter - -jitter);
                p.z += -jitter + draw_unit(S.seed, stream, item + 2) * (jitter - -jitter);
            }
            if (grid_trilinear(S.solid, p, dx) > 0) cand.push_back(p);
        }
    }
    for (const V3 &p : cand) {
        if (mask.is_set(p)) continue;
        Idx g = pos_to_index(p, dx);  // _addMarkerParticle (fluidsimulation.cpp:2637-2642)
        if (g.i >= 0 && g.j >= 0 && g.k >= 0 && g.i < S.I && g.j < S.J && g.k < S.K) {
            S.pos.push_back(p);
            S.vel.push_back(src.velocity);
        }
        mask.add(p);
    }
}

static void ensure_sources(Sim &S) {
    for (size_t i = 0; i < S.sources.size(); i++)
        if (!S.sources[i].ready && !source_build(S.sources[i], S.I, S.J, S.K, S.dx, S.seed, (int)i))
            throw std::runtime_error("inflow mesh must lie inside the domain grid");
}

// FluidSimulation::_updateInflowMeshFluidSources / _updateInflowMeshFluidSource
// (fluidsimulation.cpp:4564-4606, 4675-4697), static meshes.
static void update_inflow_sources(Sim &S) {
    int active = 0;
    for (const InflowSource &s : S.sources) active += s.enabled && !s.outflow;
    if (active == 0) return;
    ensure_sources(S);
    MaskGrid mask;
    mask.dx = S.dx;
    mask.m = GridB(S.I, S.J, S.K, 0);
    for (const V3 &p : S.pos) mask.add(p);
    for (InflowSource &src : S.sources) {
        if (!src.enabled || src.outflow) continue;
        int passes = src.emissions;
        if (passes == 0) {
            passes = 1;
            if (S.stepNumber != 0) continue;
        }
        for (int e = 0; e < passes; e++) emit_source(S, src, mask);
    }
}

// FluidSimulation::_updateOutflowMeshFluidSources / _updateOutflowMeshFluidSource
// (fluidsimulation.cpp:4608-4673, 4699-4718), fluid outflow (no diffuse
// particles exist): particles in the source's cells (all other cells when
// inversed) are removed where the source SDF, sampled with its offset, is
// negative (not negative when inversed); the rest keep their order.
static void update_outflow_sources(Sim &S) {
    bool any = false;
    for (const InflowSource &s : S.sources) any = any || (s.enabled && s.outflow);
    if (!any) return;
    ensure_sources(S);
    for (const InflowSource &src : S.sources) {
        if (!src.enabled || !src.outflow) continue;
        size_t cur = 0;
        for (size_t i = 0; i < S.pos.size(); i++) {
            Idx g = pos_to_index(S.pos[i], S.dx);
            bool cell = src.isCell(g.i, g.j, g.k) != 0;
            bool removed = false;
            if (src.inversed ? !cell : cell) {
                float d = grid_trilinear(src.sdf, S.pos[i] - src.sdfOffset, S.dx);
                removed = src.inversed ? d >= 0.0f : d < 0.0f;
            }
            if (!removed) { S.pos[cur] = S.pos[i]; S.vel[cur] = S.vel[i]; cur++; }
        }
        S.pos.resize(cur);
        S.vel.resize(cur);
    }
}

// FluidSimulation::_constrainMarkerParticleVelocities (fluidsimulation.cpp:
// 4109-4170): particles in a source cell where the source SDF, sampled at the
// particle's domain position (the reference passes no sub-grid offset
// here), is not positive take the source velocity.  Every emission pass of a
// static source repeats the same assignment, so one pass is applied.
static void constrain_inflow_particles(Sim &S) {
    for (InflowSource &src : S.sources) {
        if (!src.enabled || src.outflow || !src.constrain) continue;
        for (size_t i = 0; i < S.pos.size(); i++) {
            Idx g = pos_to_index(S.pos[i], S.dx);
            if (!src.isCell(g.i, g.j, g.k)) continue;
            if (grid_trilinear(src.sdf, S.pos[i], S.dx) > 0.0f) continue;
            S.vel[i] = src.velocity;
        }
    }
}

// FluidSimulation::_getInflowConstrainedVelocityComponents (fluidsimulation.cpp:
// 3372-3438): valid faces where a source SDF, sampled at the face's domain
// position (again without the sub-grid offset), is negative take no body
// force.
static void inflow_face_mask(const Sim &S, Valid &ex) {
    valid_init(ex, S.I, S.J, S.K);
    for (const InflowSource &src : S.sources) {
        if (!src.enabled || src.outflow || !src.constrain) continue;
        const double dx = S.dx;
        for (int k = 0; k < S.K; k++)
            for (int j = 0; j < S.J; j++)
                for (int i = 0; i < S.I + 1; i++) {
                    if (!S.valid.U(i, j, k)) continue;
                    V3 p((float)((float)i * dx), (float)(((float)j + 0.5) * dx),
                         (float)(((float)k + 0.5) * dx));
                    if (grid_trilinear(src.sdf, p, dx) < 0.0f) ex.U(i, j, k) = 1;
                }
        for (int k = 0; k < S.K; k++)
            for (int j = 0; j < S.J + 1; j++)
                for (int i = 0; i < S.I; i++) {
                    if (!S.valid.V(i, j, k)) continue;
                    V3 p((float)(((float)i + 0.5) * dx), (float)((float)j * dx),
                         (float)(((float)k + 0.5) * dx));
                    if (grid_trilinear(src.sdf, p, dx) < 0.0f) ex.V(i, j, k) = 1;
                }
        for (int k = 0; k < S.K + 1; k++)
            for (int j = 0; j < S.J; j++)
                for (int i = 0; i < S.I; i++) {
                    if (!S.valid.W(i, j, k)) continue;
                    V3 p((float)(((float)i + 0.5) * dx), (float)(((float)j + 0.5) * dx),
                         (float)((float)k * dx));
                    if (grid_trilinear(src.sdf, p, dx) < 0.0f) ex.W(i, j, k) = 1;
                }
    }
}

// FluidSimulation::_stepFluid (fluidsimulation.cpp:5471-5508) for scenes with
// domain walls only and a constant body force.
static void step_fluid(Sim &S, double dt) {
    // _updateObstacleObjects (fluidsimulation.cpp:3196-3210, 3028-3072).
    // The solid is built once particles or mesh fluid sources exist
    // (fluidsimulation.cpp:3038-3043).
    if (!S.solidUpToDate && (!S.pos.empty() || !S.sources.empty())) {
        domain_solid_phi(S.I, S.J, S.K, S.dx, S.solid);
        // _addStaticObjectsToSDF (fluidsimulation.cpp:2932-2985): enabled
        // static obstacles unioned in order, exact band 3 (fluidsimulation.h:1580).
        // Normal obstacles first, then the inversed ones: their union on a
        // fresh full-grid level set, negated (fluidsimulation.cpp:2955-2984).
        GridF inv(S.I + 1, S.J + 1, S.K + 1, (float)(((S.I + 1) + (S.J + 1) + (S.K + 1)) * S.dx));
        bool anyInversed = false;
        for (int pass = 0; pass < 2; pass++)
            for (size_t q = 0; q < S.obstacles.size(); q++) {
                const Obstacle &ob = S.obstacles[q];
                if (!ob.enabled || ob.inversed != (pass == 1)) continue;
                V3 jit = ob.hasJitter ? ob.jitter : ray_jitter(S.seed, 100 + (int)q, 0, S.dx);
                Idx go;
                GridF phi;
                int rc = obstacle_sdf(ob.mesh, S.I, S.J, S.K, S.dx, 3, jit, go, phi);
                if (rc < 0) throw std::runtime_error("obstacle mesh must lie inside its level set's grid");
                if (rc > 0) union_into(pass == 0 ? S.solid : inv, phi, go);
                anyInversed = anyInversed || pass == 1;
            }
        if (anyInversed) {
            for (float &v : inv.a) v = -v;
            union_into(S.solid, inv, Idx{0, 0, 0});
        }
        S.solidU = GridF(S.I + 1, S.J, S.K, 0.0f);
        S.solidV = GridF(S.I, S.J + 1, S.K, 0.0f);
        S.solidW = GridF(S.I, S.J, S.K + 1, 0.0f);
        S.solidUpToDate = true;
        S.weightsUpToDate = false;
    }
    update_near_solid(S);
    // _updateLiquidLevelSet + postProcess (fluidsimulation.cpp:3225-3247).
    liquid_sdf(S.I, S.J, S.K, S.dx, S.pos, S.radius, S.liquid);
    liquid_postprocess(S.liquid, S.solid, S.dx);
    // _advectVelocityField (fluidsimulation.cpp:3253-3277).
    mac_init(S.mac, S.I, S.J, S.K);
    valid_init(S.valid, S.I, S.J, S.K);
    int layers = (int)std::ceil(S.cfl) + 2;
    if (!S.pos.empty()) {
        p2g(S.I, S.J, S.K, S.dx, S.pos, S.vel, S.radius, S.mac, S.valid);
        extrapolate(S.mac.U, S.valid.U, layers);
        extrapolate(S.mac.V, S.valid.V, layers);
        extrapolate(S.mac.W, S.valid.W, layers);
    }
    S.saved = S.mac;
    // _applyConstantBodyForces (fluidsimulation.cpp:3450-3489), skipping the
    // faces the inflow sources constrain (:3372-3438, 3566-3586).
    if (!S.sources.empty()) ensure_sources(S);
    Valid ex;
    inflow_face_mask(S, ex);
    V3 bf = S.body_force();
    float beps = 1e-6f;
    auto force = [&](GridF &g, const GridB &skip, float f) {
        for (size_t q = 0; q < g.a.size(); q++)
            if (!skip.a[q]) g.a[q] += (float)(f * dt);
    };
    if (std::fabs(bf.x) > beps) force(S.mac.U, ex.U, bf.x);
    if (std::fabs(bf.y) > beps) force(S.mac.V, ex.V, bf.y);
    if (std::fabs(bf.z) > beps) force(S.mac.W, ex.W, bf.z);
    // _pressureSolve (fluidsimulation.cpp:3732-3769).
    if (!S.weightsUpToDate) {
        compute_weights(S.solid, S.I, S.J, S.K, S.w);
        S.weightsUpToDate = true;
    }
    PressureInputs in{S.I, S.J, S.K, S.dx, dt, S.pcgTol, S.pcgMaxIter, S.pcgMode, &S.liquid, &S.w,
                      &S.solidU, &S.solidV, &S.solidW, &S.mac, &S.valid};
    PressureResult pr = pressure_solve(in);
    S.lastIters = pr.iters; S.lastErr = pr.err; S.lastOk = pr.ok;
    extrapolate(S.mac.U, S.valid.U, layers);
    extrapolate(S.mac.V, S.valid.V, layers);
    extrapolate(S.mac.W, S.valid.W, layers);
    // _constrainVelocityFields (fluidsimulation.cpp:3937-3948).
    constrain(S.saved, S);
    constrain(S.mac, S);
    // _updatePICFLIPMarkerParticleVelocitiesThread (fluidsimulation.cpp:4082-4092).
    for (size_t i = 0; i < S.pos.size(); i++) {
        V3 p = S.pos[i];
        V3 vpic = mac_velocity(S.mac, p, S.dx, S.I, S.J, S.K);
        V3 vflip = S.vel[i] + vpic - mac_velocity(S.saved, p, S.dx, S.I, S.J, S.K);
        S.vel[i] = (float)S.picRatio * vpic + (float)(1 - S.picRatio) * vflip;
    }
    constrain_inflow_particles(S);  // fluidsimulation.cpp:4172-4185
    advance_particles(S, dt);
    // _updateFluidObjects (fluidsimulation.cpp:4794-4806).
    update_inflow_sources(S);
    update_outflow_sources(S);
}

// FluidSimulation::_calculateNextTimeStep (fluidsimulation.cpp:5593-5613).
static double next_time_step(const Sim &S, double dt) {
    double maxu = 0.0;
    if (S.currentFrame == 0 && S.stepNumber == 0) {
        // _predictMaximumMarkerParticleSpeed (:5529-5551): the fastest
        // enabled inflow source plus one step of body force.
        for (const InflowSource &src : S.sources)
            if (src.enabled && !src.outflow)
                maxu = std::fmax((double)std::fmax((double)length(src.velocity), 0.0), maxu);
        maxu += (double)length(S.body_force()) * dt;
    } else {
        double maxsq = 0.0;
        for (const V3 &v : S.vel) {
            double d = dot(v, v);
            if (d > maxsq) maxsq = d;
        }
        maxu = std::sqrt(maxsq);
    }
    maxu = std::fmax(0.0, maxu);
    return S.cfl * S.dx / (maxu + 1e-6);
}

// ---------------------------------------------------------------------------
// Surface reconstruction: FluidSimulation::_outputSurfaceMeshThread
// (fluidsimulation.cpp:5134-5219) with ParticleMesher::meshParticles
// (particlemesher.cpp:36-55, one polygonizer slice), Polygonizer3d
// (polygonizer3d.cpp:349-675) and the TriangleMesh post-processing
// (trianglemesh.cpp:536-583, 638-705).
// ---------------------------------------------------------------------------
static const uint64_t kMcTri[256] = {FLIPMI_MC_TRIANGLES};

// The single compute chunk of ParticleMesher::_initializeComputeChunkData*
// (particlemesher.cpp:124-275) for one slice.
struct MesherChunk {
    Idx bmin{0, 0, 0}, bmax{0, 0, 0};  // block range [bmin, bmax)
    Idx gmin{0, 0, 0}, gmax{0, 0, 0};  // node range [gmin, gmax)
    V3 offset;       // positionOffset
    int ni = 0, nj = 0, nk = 0;
};

static bool mesher_chunk(const std::vector<V3> &pts, int nsi, int nsj, int nsk, double subdx,
                         Grid<uint8_t> &active, MesherChunk &c) {
    const int bw = 10;  // particlemesher.h (_blockwidth)
    const int nbi = (nsi + bw - 1) / bw, nbj = (nsj + bw - 1) / bw, nbk = (nsk + bw - 1) / bw;
    Grid<uint8_t> seed(nbi, nbj, nbk, 0);
    const double blockdx = bw * subdx;
    for (const V3 &p : pts) {
        Idx g = pos_to_index(p, blockdx);
        if (seed.in(g.i, g.j, g.k)) seed(g.i, g.j, g.k) = 1;
    }
    // GridUtils::featherGrid26 (gridutils.cpp:265-298).
    active = seed;
    for (int k = 0; k < nbk; k++)
        for (int j = 0; j < nbj; j++)
            for (int i = 0; i < nbi; i++) {
                if (!seed(i, j, k)) continue;
                for (int dk = -1; dk <= 1; dk++)
                    for (int dj = -1; dj <= 1; dj++)
                        for (int di = -1; di <= 1; di++)
                            if (active.in(i + di, j + dj, k + dk)) active(i + di, j + dj, k + dk) = 1;
            }
    int split = 0, width = nbi;  // U
    if (nbj > width) { split = 1; width = nbj; }
    if (nbk > width) { split = 2; width = nbk; }
    Idx lo{nbi, nbj, nbk}, hi{0, 0, 0};
    bool any = false;
    for (int k = 0; k < nbk; k++)
        for (int j = 0; j < nbj; j++)
            for (int i = 0; i < nbi; i++) {
                if (!active(i, j, k)) continue;
                any = true;
                lo.i = std::min(lo.i, i); lo.j = std::min(lo.j, j); lo.k = std::min(lo.k, k);
                hi.i = std::max(hi.i, i + 1); hi.j = std::max(hi.j, j + 1); hi.k = std::max(hi.k, k + 1);
            }
    if (!any) return false;
    if (split != 0) { lo.i = std::max(lo.i - 1, 0); hi.i = std::min(hi.i + 1, nbi); }
    if (split != 1) { lo.j = std::max(lo.j - 1, 0); hi.j = std::min(hi.j + 1, nbj); }
    if (split != 2) { lo.k = std::max(lo.k - 1, 0); hi.k = std::min(hi.k + 1, nbk); }
    c.bmin = lo; c.bmax = hi;
    c.gmin = Idx{bw * lo.i, bw * lo.j, bw * lo.k};
    c.gmax = Idx{std::min(bw * hi.i, nsi), std::min(bw * hi.j, nsj), std::min(bw * hi.k, nsk)};
    c.offset = index_to_pos(c.gmin.i, c.gmin.j, c.gmin.k, subdx);
    c.ni = c.gmax.i - c.gmin.i; c.nj = c.gmax.j - c.gmin.j; c.nk = c.gmax.k - c.gmin.k;
    return true;
}

// ParticleMesher::_initializeScalarFieldData + _computeScalarField
// (particlemesher.cpp:301-450, 486-608): per node the negated minimum of
// |node - p| - r over the particles binned to the node's 10^3 block, in
// block-local float coordinates, capped at 3 r.
static void mesher_field(const std::vector<V3> &pts, const MesherChunk &c, const Grid<uint8_t> &active,
                         double subdx, double radius, GridF &field) {
    const int bw = 10;
    const float eps = 1e-6f;
    Box bb;  // AABB(pmin, pmax) (aabb.cpp:40-52), expanded by 2 (r + eps)
    {
        V3 pmin = c.offset, pmax = index_to_pos(c.gmax.i, c.gmax.j, c.gmax.k, subdx);
        double mnx = std::fmin(pmin.x, pmax.x), mny = std::fmin(pmin.y, pmax.y), mnz = std::fmin(pmin.z, pmax.z);
        double mxx = std::fmax(pmin.x, pmax.x), mxy = std::fmax(pmin.y, pmax.y), mxz = std::fmax(pmin.z, pmax.z);
        bb.pos = V3((float)mnx, (float)mny, (float)mnz);
        bb.w = mxx - mnx; bb.h = mxy - mny; bb.d = mxz - mnz;
        bb.expand(2.0f * (radius + eps));
    }
    const int cbi = (c.ni + bw - 1) / bw, cbj = (c.nj + bw - 1) / bw, cbk = (c.nk + bw - 1) / bw;
    auto block_active = [&](int i, int j, int k) {
        if (i < 0 || j < 0 || k < 0 || i >= cbi || j >= cbj || k >= cbk) return false;
        return active(c.bmin.i + i, c.bmin.j + j, c.bmin.k + k) != 0;
    };
    const float fill = (float)(3.0 * radius);
    std::vector<std::vector<float>> data((size_t)cbi * cbj * cbk);
    const float sr = 1.5f * (float)radius;  // particlemesher.h (_searchRadiusFactor)
    const float blockdx_f = (float)(bw * subdx);
    const double blockdx_d = bw * subdx;
    const float r = (float)radius;
    for (const V3 &p0 : pts) {
        if (!bb.inside(p0)) continue;
        const V3 p = p0 - c.offset;
        // _computeGridCountDataThread (particlemesher.cpp:486-541): the
        // blocks the particle is binned to.
        std::vector<Idx> blocks;
        Idx b = pos_to_index(p, blockdx_f);
        V3 bp = index_to_pos(b.i, b.j, b.k, blockdx_f);
        if (p.x - sr > bp.x && p.y - sr > bp.y && p.z - sr > bp.z && p.x + sr < bp.x + blockdx_f &&
            p.y + sr < bp.y + blockdx_f && p.z + sr < bp.z + blockdx_f) {
            if (block_active(b.i, b.j, b.k)) blocks.push_back(b);
        } else {
            Idx g0 = pos_to_index(p.x - sr, p.y - sr, p.z - sr, blockdx_f);
            Idx g1 = pos_to_index(p.x + sr, p.y + sr, p.z + sr, blockdx_f);
            for (int k = g0.k; k <= g1.k; k++)
                for (int j = g0.j; j <= g1.j; j++)
                    for (int i = g0.i; i <= g1.i; i++)
                        if (block_active(i, j, k)) blocks.push_back(Idx{i, j, k});
        }
        // _scalarFieldProducerThread (particlemesher.cpp:574-608).
        for (const Idx &q : blocks) {
            std::vector<float> &d = data[(size_t)q.i + (size_t)cbi * (q.j + (size_t)cbj * q.k)];
            if (d.empty()) d.assign(bw * bw * bw, fill);
            V3 pb = p - index_to_pos(q.i, q.j, q.k, blockdx_d);
            Idx lo = pos_to_index(V3(pb.x - sr, pb.y - sr, pb.z - sr), subdx);
            Idx hi = pos_to_index(V3(pb.x + sr, pb.y + sr, pb.z + sr), subdx);
            hi.i++; hi.j++; hi.k++;
            for (int k = std::max(lo.k, 0); k <= std::min(hi.k, bw - 1); k++)
                for (int j = std::max(lo.j, 0); j <= std::min(hi.j, bw - 1); j++)
                    for (int i = std::max(lo.i, 0); i <= std::min(hi.i, bw - 1); i++) {
                        float dist = length(index_to_pos(i, j, k, subdx) - pb) - r;
                        float &v = d[i + bw * (j + bw * k)];
                        if (dist < v) v = dist;
                    }
        }
    }
    field = GridF(c.ni, c.nj, c.nk, fill);
    for (int bk = 0; bk < cbk; bk++)
        for (int bj = 0; bj < cbj; bj++)
            for (int bi = 0; bi < cbi; bi++) {
                const std::vector<float> &d = data[(size_t)bi + (size_t)cbi * (bj + (size_t)cbj * bk)];
                if (d.empty()) continue;
                for (int k = 0; k < bw; k++)
                    for (int j = 0; j < bw; j++)
                        for (int i = 0; i < bw; i++) {
                            int a = bw * bi + i, e = bw * bj + j, f = bw * bk + k;
                            if (field.in(a, e, f)) field(a, e, f) = d[i + bw * (j + bw * k)];
                        }
            }
    for (float &v : field.a) v = -v;
}

// ScalarField::setSolidSDF (scalarfield.cpp:372-395) and
// MeshLevelSet::_trilinearInterpolateSolidGridPointsThread
// (meshlevelset.cpp:309-340): which field nodes lie in the solid.
static void mesher_solid_nodes(const GridF &solid, double dx, const MesherChunk &c, double subdx,
                               int I, int J, int K, GridB &vs) {
    vs = GridB(c.ni, c.nj, c.nk, 0);
    const bool matching = c.ni == I + 1 && c.nj == J + 1 && c.nk == K + 1 && std::fabs(subdx - dx) < 1e-12;
    if (matching) {
        for (int k = 0; k < c.nk; k++)
            for (int j = 0; j < c.nj; j++)
                for (int i = 0; i < c.ni; i++) vs(i, j, k) = solid(i, j, k) < 0;
        return;
    }
    const bool aligned2 = std::fabs(2 * subdx - dx) < 1e-6 && length(c.offset) < 1e-6;
    for (int k = 0; k < c.nk; k++)
        for (int j = 0; j < c.nj; j++)
            for (int i = 0; i < c.ni; i++) {
                float d;
                if (aligned2 && i % 2 == 0 && j % 2 == 0 && k % 2 == 0) {
                    d = solid(i >> 1, j >> 1, k >> 1);
                } else {
                    d = grid_trilinear(solid, index_to_pos(i, j, k, subdx) + c.offset, dx);
                }
                vs(i, j, k) = d < 0;
            }
}

// Polygonizer3d::polygonizeSurface with a solid SDF (polygonizer3d.cpp:
// 349-675): marching cubes over the cells with nodes on both sides of 0,
// in k-j-i order, one vertex per cut edge made by the first cell that
// meets it, clamped out of the solid along the edge.
static void polygonize(const GridF &field, const GridB &vsolid, const GridF &solid, double dx, V3 offset,
                       double subdx, SurfaceMesh &m) {
    const int ni = field.ni - 1, nj = field.nj - 1, nk = field.nk - 1;
    auto value = [&](int i, int j, int k) -> double {  // scalarfield.cpp:439-448
        double v = field(i, j, k);
        if (vsolid(i, j, k) && v > 0.0) v = 0.0;
        return v;
    };
    bool any = false;
    for (size_t q = 0; q < field.a.size() && !any; q++) {
        int i = (int)(q % field.ni), j = (int)((q / field.ni) % field.nj), k = (int)(q / ((size_t)field.ni * field.nj));
        any = value(i, j, k) > 0.0;
    }
    if (!any || ni <= 0 || nj <= 0 || nk <= 0) return;
    Grid<int> eU(ni, nj + 1, nk + 1, -1), eV(ni + 1, nj, nk + 1, -1), eW(ni + 1, nj + 1, nk, -1);
    static const int co[8] = {FLIPMI_MC_CORNER_OFFSETS};
    const float fdx = (float)subdx;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                int ci[8], cj[8], ck[8];
                double val[8];
                V3 pos[8];
                int cube = 0;
                for (int c = 0; c < 8; c++) {
                    ci[c] = i + (co[c] & 1); cj[c] = j + ((co[c] >> 1) & 1); ck[c] = k + ((co[c] >> 2) & 1);
                    val[c] = value(ci[c], cj[c], ck[c]);
                    pos[c] = fdx * V3((float)ci[c], (float)cj[c], (float)ck[c]);
                    if (val[c] > 0.0) cube |= 1 << c;
                }
                const int cut = mc_cut_edges(cube);
                if (cut == 0) continue;
                int vlist[12];
                for (int e = 0; e < 12; e++) {
                    if (!(cut >> e & 1)) continue;
                    // Edge keys (polygonizer3d.cpp:469-556): U edges by their
                    // lower corner, V and W likewise.
                    const int a = kMcEdgeA[e], b = kMcEdgeB[e];
                    const int lo = (e == 2 || e == 6) ? b : (e == 3 || e == 7) ? b : a;
                    int &slot = (e == 0 || e == 2 || e == 4 || e == 6) ? eU(ci[lo], cj[lo], ck[lo])
                               : (e >= 8) ? eV(ci[lo], cj[lo], ck[lo])
                                          : eW(ci[lo], cj[lo], ck[lo]);
                    if (slot == -1) {
                        // _vertexInterp (polygonizer3d.cpp:409-444).
                        const double teps = 1e-10;
                        double minmu = 0.0, maxmu = 1.0;
                        double s1 = grid_trilinear(solid, pos[a] + offset, dx);
                        double s2 = grid_trilinear(solid, pos[b] + offset, dx);
                        if ((s1 < 0.0 && s2 >= 0.0) || (s2 < 0.0 && s1 >= 0.0)) {
                            double diff = s2 - s1;
                            if (std::fabs(diff) > teps) {
                                double su = -s1 / diff;
                                if (s1 < 0.0) minmu = su; else maxmu = su;
                            } else {
                                maxmu = minmu;
                            }
                        }
                        minmu = std::fmax(minmu, teps);
                        maxmu = std::fmin(maxmu, 1.0 - teps);
                        double mu = (0.0 - val[a]) / (val[b] - val[a]);
                        if (mu < minmu) mu = minmu;
                        if (mu > maxmu) mu = maxmu;
                        m.verts.push_back(pos[a] + (float)mu * (pos[b] - pos[a]));
                        slot = (int)m.verts.size() - 1;
                    }
                    vlist[e] = slot;
                }
                const uint64_t w = kMcTri[cube];
                const int nt = (int)(w >> 60);
                for (int q = 0; q < 3 * nt; q++) m.tris.push_back(vlist[(w >> (4 * q)) & 15]);
            }
}

// TriangleMesh::removeTriangles + removeExtraneousVertices
// (trianglemesh.cpp:638-705): order-preserving compaction.
static void mesh_remove(SurfaceMesh &m, const std::vector<uint8_t> &drop) {
    std::vector<int> tris;
    for (size_t t = 0; t < drop.size(); t++)
        if (!drop[t]) tris.insert(tris.end(), m.tris.begin() + 3 * t, m.tris.begin() + 3 * t + 3);
    m.tris = tris;
    std::vector<int> remap(m.verts.size(), -1);
    for (int v : m.tris) remap[v] = 0;
    std::vector<V3> verts;
    for (size_t v = 0; v < m.verts.size(); v++)
        if (remap[v] == 0) { remap[v] = (int)verts.size(); verts.push_back(m.verts[v]); }
    m.verts = verts;
    for (int &v : m.tris) v = remap[v];
}

// TriangleMesh::smooth (trianglemesh.cpp:536-583).
static void mesh_smooth(SurfaceMesh &m, double value, int iterations) {
    if (iterations == 0) return;
    const size_t nv = m.verts.size(), nt = m.tris.size() / 3;
    std::vector<std::vector<int>> vt(nv);
    for (size_t t = 0; t < nt; t++)
        for (int c = 0; c < 3; c++) vt[m.tris[3 * t + c]].push_back((int)t);
    for (int it = 0; it < iterations; it++) {
        std::vector<V3> nvx(nv);
        for (size_t i = 0; i < nv; i++) {
            int count = 0;
            V3 avg;
            for (int t : vt[i])
                for (int c = 0; c < 3; c++) {
                    int v = m.tris[3 * (size_t)t + c];
                    if (v != (int)i) { avg = avg + m.verts[v]; count++; }
                }
            avg = avg / (float)count;
            V3 v = m.verts[i];
            nvx[i] = v + (float)value * (avg - v);
        }
        m.verts = nvx;
    }
}

// FluidSimulation::_outputSurfaceMeshThread (fluidsimulation.cpp:5134-5219)
// with the default meshing volume (none), preview and motion blur off.
static void surface_mesh(const std::vector<V3> &pts, const GridF &simSolid, int I, int J, int K, double dx,
                         double particleRadius, const SurfaceParams &sp, V3 domainOffset, double domainScale,
                         SurfaceMesh &out) {
    out = SurfaceMesh();
    if (pts.empty()) return;
    GridF solid = simSolid;  // constructMinimalSignedDistanceField (fluidsimulation.cpp:5232-5233)
    {   // _polygonizeOutputSurface (fluidsimulation.cpp:5050-5068), offset enabled
        float offset = (float)(sp.obstacleOffset * dx);
        if (std::fabs(offset) > 1e-9f)
            for (int k = 2; k < K - 1; k++)
                for (int j = 2; j < J - 1; j++)
                    for (int i = 2; i < I - 1; i++) solid(i, j, k) = solid(i, j, k) + offset;
    }
    const int sub = sp.subdivisions;
    const double subdx = dx / (double)sub;
    const double radius = particleRadius * sp.particleScale;
    Grid<uint8_t> active;
    MesherChunk c;
    SurfaceMesh m;
    if (mesher_chunk(pts, I * sub + 1, J * sub + 1, K * sub + 1, subdx, active, c)) {
        GridF field;
        mesher_field(pts, c, active, subdx, radius, field);
        GridB vs;
        mesher_solid_nodes(solid, dx, c, subdx, I, J, K, vs);
        polygonize(field, vs, solid, dx, c.offset, subdx, m);
        // m.translate(positionOffset); scale by localdx / subdx and back
        // (particlemesher.cpp:46-53).
        const double scale = 0.1 / subdx;  // particlemesher.h (_localdx)
        const float s = (float)scale, inv = (float)(1.0 / scale);
        for (V3 &v : m.verts) {
            v = v + c.offset;
            v = V3(v.x * s, v.y * s, v.z * s);
            v = V3(v.x * inv, v.y * inv, v.z * inv);
        }
    }
    if (sp.removeNearDomain) {  // _removeMeshNearDomain (fluidsimulation.cpp:4922-4956)
        const int width = 2 + sp.removeNearDomainDistance;
        std::vector<uint8_t> drop(m.tris.size() / 3, 0);
        for (size_t t = 0; t < drop.size(); t++) {
            V3 cen = (m.verts[m.tris[3 * t]] + m.verts[m.tris[3 * t + 1]] + m.verts[m.tris[3 * t + 2]]) / 3.0f;
            Idx g = pos_to_index(cen, dx);
            bool ok = g.i >= width && g.i < I - width && g.j >= width && g.j < J - width && g.k >= width &&
                      g.k < K - width;
            drop[t] = !ok;
        }
        mesh_remove(m, drop);
    }
    mesh_smooth(m, sp.smoothValue, sp.smoothIterations);
    if (sp.invertContactNormals) {  // _invertContactNormals (fluidsimulation.cpp:4898-4920)
        float eps = (float)(sp.contactThreshold * dx);
        std::vector<uint8_t> contact(m.verts.size());
        for (size_t v = 0; v < m.verts.size(); v++) contact[v] = grid_trilinear(simSolid, m.verts[v], dx) < eps;
        for (size_t t = 0; t < m.tris.size() / 3; t++)
            if (contact[m.tris[3 * t]] || contact[m.tris[3 * t + 1]] || contact[m.tris[3 * t + 2]])
                std::swap(m.tris[3 * t + 1], m.tris[3 * t + 2]);
    }
    const float ds = (float)domainScale;
    for (V3 &v : m.verts) {
        v = V3(v.x * ds, v.y * ds, v.z * ds);
        v = v + domainOffset;
    }
    out = m;
}

// FluidSimulation::_initializeParticleRadii (fluidsimulation.cpp:2644-2648).
static double marker_particle_radius(double dx) {
    double volume = dx * dx * dx / 8.0;
    double pi = 3.141592653;
    return std::pow(3 * volume / (4 * pi), 1.0 / 3.0);
}

static void surface_mesh_of(Sim &S) {
    surface_mesh(S.pos, S.solid, S.I, S.J, S.K, S.dx, marker_particle_radius(S.dx), S.surface, S.domainOffset,
                 S.domainScale, S.isomesh);
}

// FluidSimulation::update (fluidsimulation.cpp:5755-5846).
static void update(Sim &S, double dt) {
    if (!S.initialized) throw std::runtime_error("FluidSimulation must be initialized before update");
    if (dt < 0.0) throw std::domain_error("delta time must be greater than or equal to 0");
    double epsdt = 1e-6;
    bool zero = dt < epsdt;
    dt = std::max(dt, epsdt);
    S.frameDt = dt; S.remaining = dt; S.stepNumber = 0;
    bool skipped = zero && S.outputInitialized;
    double substepTime = S.frameDt / (double)S.minSteps;
    double eps = 1e-9;
    do {
        S.curStep = std::fmin(next_time_step(S, dt), S.remaining);
        double done = S.frameDt - S.remaining;
        double limit = (S.stepNumber + 1) * substepTime;
        if (done + S.curStep > limit) S.curStep = std::fmin(substepTime, S.remaining);
        if (S.stepNumber == S.maxSteps - 1) S.curStep = S.remaining;
        S.remaining -= S.curStep;
        if (!skipped) {
            step_fluid(S, S.curStep);
            // _outputSimulationData (fluidsimulation.cpp:5445-5463), at the
            // end of the frame's first substep.
            if (S.stepNumber == 0 && S.surfaceEnabled) surface_mesh_of(S);
        }
        S.stepNumber++;
    } while (S.remaining > eps);
    S.lastSubsteps = S.stepNumber;
    S.outputInitialized = true;
    S.currentFrame++;
}

static std::vector<V3> to_v3(const float *p, int n) {
    std::vector<V3> v((size_t)n);
    for (int i = 0; i < n; i++) v[i] = V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    return v;
}

template <class T>
static void copy_in(Grid<T> &g, const T *src) { std::memcpy(g.a.data(), src, sizeof(T) * g.a.size()); }
template <class T>
static void copy_out(const Grid<T> &g, T *dst) { std::memcpy(dst, g.a.data(), sizeof(T) * g.a.size()); }

}  // namespace orc

using namespace orc;

extern "C" {

void orc_domain_solid(int isize, int jsize, int ksize, double dx, float *phi_nodal, float *wU,
                      float *wV, float *wW, float *wC) {
    GridF phi;
    domain_solid_phi(isize, jsize, ksize, dx, phi);
    copy_out(phi, phi_nodal);
    Weights w;
    compute_weights(phi, isize, jsize, ksize, w);
    if (wU) copy_out(w.U, wU);
    if (wV) copy_out(w.V, wV);
    if (wW) copy_out(w.W, wW);
    if (wC) copy_out(w.C, wC);
}

void orc_liquid_sdf(int isize, int jsize, int ksize, double dx, const float *pos, int n,
                    double radius, const float *solid_phi_nodal, float *phi_out) {
    GridF phi;
    liquid_sdf(isize, jsize, ksize, dx, to_v3(pos, n), radius, phi);
    if (solid_phi_nodal) {
        GridF solid(isize + 1, jsize + 1, ksize + 1, 0.0f);
        copy_in(solid, solid_phi_nodal);
        liquid_postprocess(phi, solid, dx);
    }
    copy_out(phi, phi_out);
}

void orc_p2g(int isize, int jsize, int ksize, double dx, const float *pos, const float *vel, int n,
             double radius, float *u, float *v, float *w, uint8_t *vu, uint8_t *vv, uint8_t *vw) {
    MAC m;
    Valid va;
    mac_init(m, isize, jsize, ksize);
    valid_init(va, isize, jsize, ksize);
    p2g(isize, jsize, ksize, dx, to_v3(pos, n), to_v3(vel, n), radius, m, va);
    copy_out(m.U, u); copy_out(m.V, v); copy_out(m.W, w);
    copy_out(va.U, vu); copy_out(va.V, vv); copy_out(va.W, vw);
}

void orc_extrapolate(int ni, int nj, int nk, float *grid, uint8_t *valid, int layers) {
    GridF g(ni, nj, nk, 0.0f);
    GridB v(ni, nj, nk, 0);
    copy_in(g, grid);
    copy_in(v, valid);
    extrapolate(g, v, layers);
    copy_out(g, grid);
}

int orc_pressure_solve(int isize, int jsize, int ksize, double dx, double dt, double tol,
                       int maxiter, int mode, const float *liquid_phi,
                       const float *solid_phi_nodal, const float *wU, const float *wV,
                       const float *wW, const float *wC, float *u, float *v, float *w,
                       uint8_t *vu, uint8_t *vv, uint8_t *vw, float *su, float *sv, float *sw,
                       double *rhs_out, double *soln_out, int *n_rows, int *iters, double *err) {
    (void)solid_phi_nodal;
    GridF liquid(isize, jsize, ksize, 0.0f);
    copy_in(liquid, liquid_phi);
    Weights wg;
    wg.U = GridF(isize + 1, jsize, ksize, 0.0f); copy_in(wg.U, wU);
    wg.V = GridF(isize, jsize + 1, ksize, 0.0f); copy_in(wg.V, wV);
    wg.W = GridF(isize, jsize, ksize + 1, 0.0f); copy_in(wg.W, wW);
    wg.C = GridF(isize, jsize, ksize, 0.0f); copy_in(wg.C, wC);
    // Solid face velocities, in/out (null = zero).
    GridF sU(isize + 1, jsize, ksize, 0.0f), sV(isize, jsize + 1, ksize, 0.0f),
        sW(isize, jsize, ksize + 1, 0.0f);
    if (su) copy_in(sU, su);
    if (sv) copy_in(sV, sv);
    if (sw) copy_in(sW, sw);
    MAC m;
    mac_init(m, isize, jsize, ksize);
    copy_in(m.U, u); copy_in(m.V, v); copy_in(m.W, w);
    Valid va;
    valid_init(va, isize, jsize, ksize);
    copy_in(va.U, vu); copy_in(va.V, vv); copy_in(va.W, vw);
    PressureInputs in{isize, jsize, ksize, dx, dt, tol, maxiter, mode, &liquid, &wg, &sU, &sV, &sW,
                      &m, &va};
    PressureResult pr = pressure_solve(in);
    if (n_rows) *n_rows = pr.n;
    if (rhs_out) std::copy(pr.rhs.begin(), pr.rhs.end(), rhs_out);
    if (soln_out) std::copy(pr.soln.begin(), pr.soln.end(), soln_out);
    if (iters) *iters = pr.iters;
    if (err) *err = pr.err;
    copy_out(m.U, u); copy_out(m.V, v); copy_out(m.W, w);
    copy_out(va.U, vu); copy_out(va.V, vv); copy_out(va.W, vw);
    if (su) copy_out(sU, su);
    if (sv) copy_out(sV, sv);
    if (sw) copy_out(sW, sw);
    return pr.ok;
}

void orc_mic0_apply(int isize, int jsize, int ksize, double dx, double dt, const float *liquid_phi,
                    const float *wU, const float *wV, const float *wW, const float *wC,
                    const double *r_in, double *invdiag_out, double *z_out) {
    GridF liquid(isize, jsize, ksize, 0.0f);
    copy_in(liquid, liquid_phi);
    Weights wg;
    wg.U = GridF(isize + 1, jsize, ksize, 0.0f); copy_in(wg.U, wU);
    wg.V = GridF(isize, jsize + 1, ksize, 0.0f); copy_in(wg.V, wV);
    wg.W = GridF(isize, jsize, ksize + 1, 0.0f); copy_in(wg.W, wW);
    wg.C = GridF(isize, jsize, ksize, 0.0f); copy_in(wg.C, wC);
    PressureInputs in{isize, jsize, ksize, dx, dt, 1e-9, 1000, 0, &liquid, &wg, nullptr, nullptr,
                      nullptr, nullptr, nullptr};
    Rows R;
    build_rows(in, R);
    assemble_matrix(in, R);
    MIC M;
    mic0_factor(R, M);
    std::vector<double> r(r_in, r_in + R.cell.size()), z;
    mic0_apply(R, M, r, z);
    std::copy(M.invdiag.begin(), M.invdiag.end(), invdiag_out);
    std::copy(z.begin(), z.end(), z_out);
}

void orc_advect_particles(int isize, int jsize, int ksize, double dx, const float *u,
                          const float *v, const float *w, const float *pos, int n, double dt,
                          int order, float *out) {
    MAC m;
    mac_init(m, isize, jsize, ksize);
    copy_in(m.U, u); copy_in(m.V, v); copy_in(m.W, w);
    // ParticleAdvector::_RK1.._RK4 (particleadvector.cpp:412-445).
    for (int q = 0; q < n; q++) {
        V3 p0(pos[3 * q], pos[3 * q + 1], pos[3 * q + 2]), p1;
        auto vel = [&](V3 p) { return mac_velocity(m, p, dx, isize, jsize, ksize); };
        if (order == 4) {
            V3 k1 = vel(p0), k2 = vel(p0 + (float)(0.5 * dt) * k1),
               k3 = vel(p0 + (float)(0.5 * dt) * k2), k4 = vel(p0 + (float)dt * k3);
            p1 = p0 + (float)(dt / 6.0f) * (k1 + 2.0f * k2 + 2.0f * k3 + k4);
        } else if (order == 3) {
            p1 = rk3(m, p0, dt, dx, isize, jsize, ksize);
        } else if (order == 2) {
            V3 k1 = vel(p0), k2 = vel(p0 + (float)(0.5 * dt) * k1);
            p1 = p0 + (float)dt * k2;
        } else {
            V3 k1 = vel(p0);
            p1 = p0 + (float)dt * k1;
        }
        out[3 * q] = p1.x; out[3 * q + 1] = p1.y; out[3 * q + 2] = p1.z;
    }
}

void orc_trilinear_velocity(int isize, int jsize, int ksize, double dx, const float *u,
                            const float *v, const float *w, const float *pos, int n, float *out) {
    MAC m;
    mac_init(m, isize, jsize, ksize);
    copy_in(m.U, u); copy_in(m.V, v); copy_in(m.W, w);
    for (int q = 0; q < n; q++) {
        V3 r = mac_velocity(m, V3(pos[3 * q], pos[3 * q + 1], pos[3 * q + 2]), dx, isize, jsize, ksize);
        // ParticleAdvector::_validateOutput (particleadvector.cpp:506-516).
        if (std::isinf(r.x) || std::isnan(r.x) || std::isinf(r.y) || std::isnan(r.y) ||
            std::isinf(r.z) || std::isnan(r.z))
            r = V3();
        out[3 * q] = r.x; out[3 * q + 1] = r.y; out[3 * q + 2] = r.z;
    }
}

static TriMesh make_mesh(const float *verts, int nv, const int *tris, int nt) {
    TriMesh m;
    for (int i = 0; i < nv; i++) m.v.push_back(V3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]));
    m.t.assign(tris, tris + 3 * (size_t)nt);
    return m;
}

// MeshUtils::getGridNodesInsideTriangleMesh on an nni x nnj x nnk node grid
// with the given ray jitter; -1 if the mesh leaves the grid box.
int orc_mesh_nodes_inside(const float *verts, int nv, const int *tris, int nt, double dx, int nni,
                          int nnj, int nnk, const float *jitter, uint8_t *out) {
    GridB nodes;
    if (!nodes_inside(make_mesh(verts, nv, tris, nt), dx, nni, nnj, nnk,
                      V3(jitter[0], jitter[1], jitter[2]), nodes))
        return -1;
    copy_out(nodes, out);
    return 0;
}

// MeshLevelSet(isize, jsize, ksize, dx).fastCalculateSignedDistanceField(mesh,
// band) with the given ray jitter; phi is (isize+1)(jsize+1)(ksize+1).
int orc_mesh_sdf(const float *verts, int nv, const int *tris, int nt, int isize, int jsize,
                 int ksize, double dx, int band, const float *jitter, float *phi) {
    GridF g(isize + 1, jsize + 1, ksize + 1, 0.0f);
    if (!mesh_sdf(make_mesh(verts, nv, tris, nt), dx, band, V3(jitter[0], jitter[1], jitter[2]), g))
        return -1;
    copy_out(g, phi);
    return 0;
}

// MeshFluidSource::update + getCells of source `id` on an I x J x K grid,
// stand-alone: dims = sub-grid offset and node dims; phi / cells optional.
int orc_inflow_build(const float *verts, int nv, const int *tris, int nt, int I, int J, int K,
                     double dx, uint64_t seed, int id, int *dims, float *phi, uint8_t *cells) {
    InflowSource src;
    src.mesh = make_mesh(verts, nv, tris, nt);
    if (!source_build(src, I, J, K, dx, seed, id)) return -1;
    dims[0] = src.gridOffset.i; dims[1] = src.gridOffset.j; dims[2] = src.gridOffset.k;
    dims[3] = src.sdf.ni; dims[4] = src.sdf.nj; dims[5] = src.sdf.nk;
    if (phi) copy_out(src.sdf, phi);
    if (cells) copy_out(src.isCell, cells);
    return 0;
}

// MeshFluidSource(isize, jsize, ksize, dx) + updateMeshStatic + setInflow +
// setVelocity, added with FluidSimulation::addMeshFluidSource; returns its
// index.
// MeshObject(isize, jsize, ksize, dx) + updateMeshStatic(mesh), added with
// FluidSimulation::addMeshObstacle (meshobject.cpp:50-59, fluidsimulation.cpp:
// 1994-2008).  jitter (3 floats) fixes the ray jitter of its inside test;
// NULL draws it from the keyed generator.  Returns the obstacle's index, or
// -1 for a mesh of 25 or more islands.
int orc_sim_add_obstacle(void *h, const float *verts, int nv, const int *tris, int nt,
                         const float *jitter) {
    Sim *S = (Sim *)h;
    Obstacle ob;
    ob.mesh = referenced_only(make_mesh(verts, nv, tris, nt));
    if (island_count(ob.mesh) >= 25) return -1;
    if (jitter) { ob.hasJitter = true; ob.jitter = V3(jitter[0], jitter[1], jitter[2]); }
    S->obstacles.push_back(ob);
    S->solidUpToDate = false;
    return (int)S->obstacles.size() - 1;
}
// MeshObject::inverse() state (no rebuild, as in the reference).
void orc_sim_set_obstacle_inversed(void *h, int id, int inversed) {
    ((Sim *)h)->obstacles[id].inversed = inversed != 0;
}
// MeshObject::enable()/disable(): the solid is rebuilt at the next substep.
void orc_sim_set_obstacle(void *h, int id, int enabled) {
    Sim *S = (Sim *)h;
    S->obstacles[id].enabled = enabled != 0;
    S->solidUpToDate = false;
}

int orc_sim_add_inflow(void *h, const float *verts, int nv, const int *tris, int nt, double vx,
                       double vy, double vz) {
    Sim *S = (Sim *)h;
    InflowSource src;
    src.mesh = make_mesh(verts, nv, tris, nt);
    src.velocity = V3((float)vx, (float)vy, (float)vz);
    S->sources.push_back(src);
    return (int)S->sources.size() - 1;
}
// setOutflow() and outflowInverse() of one source (meshfluidsource.cpp:
// 82-84, 158-160).
void orc_sim_set_outflow(void *h, int id, int outflow, int inversed) {
    InflowSource &src = ((Sim *)h)->sources[id];
    src.outflow = outflow != 0;
    src.inversed = inversed != 0;
}
// enable()/disable(), enable/disableConstrainedFluidVelocity() and
// setSubstepEmissions() of one source (meshfluidsource.cpp:51-72, 146-156).
void orc_sim_set_inflow(void *h, int id, int enabled, int constrain, int emissions) {
    InflowSource &src = ((Sim *)h)->sources[id];
    src.enabled = enabled != 0;
    src.constrain = constrain != 0;
    src.emissions = emissions;
}
// setMarkerParticleJitterFactor / enable/disableJitterSurfaceMarkerParticles
// (fluidsimulation.cpp:185-210) and the seed of the jitter draws.
void orc_sim_set_jitter(void *h, double factor, int surface, uint64_t seed) {
    Sim *S = (Sim *)h;
    S->jitterFactor = factor;
    S->jitterSurface = surface != 0;
    S->seed = seed;
}
// The source's SDF (nodal, sub-grid) and its placement after the first
// update; dims/offset as ints (gi, gj, gk, ni, nj, nk).  -1 before it exists.
int orc_sim_inflow_sdf(void *h, int id, int *dims, float *phi) {
    InflowSource &src = ((Sim *)h)->sources[id];
    if (!src.ready) return -1;
    dims[0] = src.gridOffset.i; dims[1] = src.gridOffset.j; dims[2] = src.gridOffset.k;
    dims[3] = src.sdf.ni; dims[4] = src.sdf.nj; dims[5] = src.sdf.nk;
    if (phi) copy_out(src.sdf, phi);
    return 0;
}
// The source's cells as a flat I*J*K byte grid.
int orc_sim_inflow_cells(void *h, int id, uint8_t *cells) {
    InflowSource &src = ((Sim *)h)->sources[id];
    if (!src.ready) return -1;
    copy_out(src.isCell, cells);
    return 0;
}
// The jitter of one inside test of source id (which = 0: SDF signs, 1: cells).
void orc_ray_jitter(uint64_t seed, int id, int which, double dx, float *out) {
    V3 j = ray_jitter(seed, id, which, dx);
    out[0] = j.x; out[1] = j.y; out[2] = j.z;
}

void *orc_sim_create(int isize, int jsize, int ksize, double dx) {
    Sim *S = new Sim();
    S->I = isize; S->J = jsize; S->K = ksize; S->dx = dx;
    return S;
}
void orc_sim_destroy(void *h) { delete (Sim *)h; }
void orc_sim_add_body_force(void *h, double fx, double fy, double fz) {
    ((Sim *)h)->bodyForces.push_back(V3((float)fx, (float)fy, (float)fz));
}
void orc_sim_set_cfl(void *h, double cfl) { ((Sim *)h)->cfl = cfl; }
void orc_sim_set_substeps(void *h, int mn, int mx) { ((Sim *)h)->minSteps = mn; ((Sim *)h)->maxSteps = mx; }
void orc_sim_set_pic_ratio(void *h, double r) { ((Sim *)h)->picRatio = r; }
void orc_sim_set_pcg(void *h, int mode, double tol, int maxiter) {
    Sim *S = (Sim *)h;
    S->pcgMode = mode; S->pcgTol = tol; S->pcgMaxIter = maxiter;
}
// FluidSimulation::loadMarkerParticleData (fluidsimulation.cpp:2488-2509):
// the particles queue up; initialize() loads them.
void orc_sim_load_particles(void *h, const float *pos, const float *vel, int n) {
    Sim *S = (Sim *)h;
    for (int i = 0; i < n; i++) {
        S->loadPos.push_back(V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
        S->loadVel.push_back(V3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]));
    }
}
// setDomainOffset(double, double, double) / setDomainScale(double)
// (fluidsimulation.cpp:307-330).
void orc_sim_set_domain(void *h, double ox, double oy, double oz, double scale) {
    Sim *S = (Sim *)h;
    S->domainOffset = V3((float)ox, (float)oy, (float)oz);
    S->domainScale = scale;
}
// getMarkerParticle{Position,Velocity}DataRange (fluidsimulation.cpp:
// 2314-2338): positions leave in domain coordinates, p * scale + offset;
// an invalid range is a domain_error (-1 here).
int orc_sim_get_particle_range(void *h, int start, int end, float *pos, float *vel) {
    Sim *S = (Sim *)h;
    if (start < 0 || end > (int)S->pos.size() || start > end) return -1;
    for (int i = start; i < end; i++) {
        V3 p = S->pos[i] * (float)S->domainScale + S->domainOffset;
        float *o = pos + 3 * (size_t)(i - start), *u = vel + 3 * (size_t)(i - start);
        o[0] = p.x; o[1] = p.y; o[2] = p.z;
        u[0] = S->vel[i].x; u[1] = S->vel[i].y; u[2] = S->vel[i].z;
    }
    return 0;
}
int orc_sim_initialize(void *h) {
    Sim *S = (Sim *)h;
    if (S->initialized) return 0;
    // _loadParticles / _loadMarkerParticles (fluidsimulation.cpp:2669-2673,
    // 2773-2785): domain coordinates back to grid coordinates, then keep
    // the points inside the grid box.
    Box b;
    b.pos = V3(0, 0, 0);
    b.w = S->I * S->dx; b.h = S->J * S->dx; b.d = S->K * S->dx;
    for (size_t i = 0; i < S->loadPos.size(); i++) {
        V3 p = (S->loadPos[i] - S->domainOffset) / (float)S->domainScale;
        if (!b.inside(p)) continue;
        S->pos.push_back(p);
        S->vel.push_back(S->loadVel[i]);
    }
    S->loadPos.clear();
    S->loadVel.clear();
    // _initializeParticleRadii (fluidsimulation.cpp:2644-2650).
    S->radius = 0.5 * 1.0 * S->dx * std::sqrt(3.0);
    // Until the first build the solid SDF holds its distance upper bound
    // (fluidsimulation.cpp:2574, meshlevelset.cpp:35-42, 769-771).
    if (S->solid.a.empty())
        S->solid = GridF(S->I + 1, S->J + 1, S->K + 1,
                         (float)(((S->I + 1) + (S->J + 1) + (S->K + 1)) * S->dx));
    S->initialized = true;
    return 0;
}
int orc_sim_update(void *h, double dt) {
    try {
        update(*(Sim *)h, dt);
    } catch (std::exception &) {
        return -1;
    }
    return 0;
}
int orc_sim_num_particles(void *h) { return (int)((Sim *)h)->pos.size(); }
void orc_sim_get_particles(void *h, float *pos, float *vel) {
    Sim *S = (Sim *)h;
    for (size_t i = 0; i < S->pos.size(); i++) {
        pos[3 * i] = S->pos[i].x; pos[3 * i + 1] = S->pos[i].y; pos[3 * i + 2] = S->pos[i].z;
        vel[3 * i] = S->vel[i].x; vel[3 * i + 1] = S->vel[i].y; vel[3 * i + 2] = S->vel[i].z;
    }
}
void orc_sim_get_float_grid(void *h, int which, float *out) {
    Sim *S = (Sim *)h;
    switch (which) {
        case 0: copy_out(S->liquid, out); break;
        case 1: copy_out(S->solid, out); break;
        case 2: copy_out(S->mac.U, out); break;
        case 3: copy_out(S->mac.V, out); break;
        case 4: copy_out(S->mac.W, out); break;
        case 5: copy_out(S->w.U, out); break;
        case 6: copy_out(S->w.V, out); break;
        case 7: copy_out(S->w.W, out); break;
        case 8: copy_out(S->w.C, out); break;
        default: break;
    }
}
void orc_sim_get_bool_grid(void *h, int which, uint8_t *out) {
    Sim *S = (Sim *)h;
    switch (which) {
        case 0: copy_out(S->valid.U, out); break;
        case 1: copy_out(S->valid.V, out); break;
        case 2: copy_out(S->valid.W, out); break;
        case 3: copy_out(S->nearSolid, out); break;
        default: break;
    }
}
// Count grid of the last removal pass (zeros before the first substep), the
// particles that pass removed for their speed, and all speed removals since
// the simulation was created.
void orc_sim_get_count_grid(void *h, int32_t *out, int *speed_last, int *speed_total) {
    Sim *S = (Sim *)h;
    const size_t nc = (size_t)S->I * S->J * S->K;
    if (S->removalCount.a.size() == nc) copy_out(S->removalCount, out);
    else std::memset(out, 0, sizeof(int32_t) * nc);
    if (speed_last) *speed_last = S->removedBySpeed;
    if (speed_total) *speed_total = S->removedBySpeedTotal;
}

static SurfaceParams surface_params(int subdivisions, double smooth_value, int smooth_iterations,
                                    double particle_scale, double obstacle_offset, int remove_near_domain,
                                    int remove_distance, int invert_contact, double contact_threshold) {
    SurfaceParams p;
    p.subdivisions = subdivisions;
    p.smoothValue = smooth_value;
    p.smoothIterations = smooth_iterations;
    p.particleScale = particle_scale;
    p.obstacleOffset = obstacle_offset;
    p.removeNearDomain = remove_near_domain != 0;
    p.removeNearDomainDistance = remove_distance;
    p.invertContactNormals = invert_contact != 0;
    p.contactThreshold = contact_threshold;
    return p;
}

void orc_sim_set_surface(void *h, int enabled, int subdivisions, double smooth_value, int smooth_iterations,
                         double particle_scale, double obstacle_offset, int remove_near_domain,
                         int remove_distance, int invert_contact, double contact_threshold) {
    Sim *S = (Sim *)h;
    S->surfaceEnabled = enabled != 0;
    S->surface = surface_params(subdivisions, smooth_value, smooth_iterations, particle_scale, obstacle_offset,
                                remove_near_domain, remove_distance, invert_contact, contact_threshold);
}

// FluidSimulation::getIsomesh (fluidsimulation.cpp:2543-2546): counts, and
// the arrays when the pointers are not null.
void orc_sim_isomesh(void *h, int *nverts, int *ntris, float *verts, int *tris) {
    const SurfaceMesh &m = ((Sim *)h)->isomesh;
    *nverts = (int)m.verts.size();
    *ntris = (int)(m.tris.size() / 3);
    if (verts)
        for (size_t v = 0; v < m.verts.size(); v++) {
            verts[3 * v] = m.verts[v].x; verts[3 * v + 1] = m.verts[v].y; verts[3 * v + 2] = m.verts[v].z;
        }
    if (tris) std::memcpy(tris, m.tris.data(), sizeof(int) * m.tris.size());
}

// The surface of a particle set over a solid SDF, as the simulation makes
// it (surface_mesh above); the mesh is kept in a handle until freed.
void *orc_surface_mesh(int I, int J, int K, double dx, const float *pos, int n, const float *solid_phi,
                       int subdivisions, double smooth_value, int smooth_iterations, double particle_scale,
                       double obstacle_offset, int remove_near_domain, int remove_distance, int invert_contact,
                       double contact_threshold, const float *domain_offset, double domain_scale) {
    GridF solid(I + 1, J + 1, K + 1, 0.0f);
    copy_in(solid, solid_phi);
    SurfaceParams sp = surface_params(subdivisions, smooth_value, smooth_iterations, particle_scale,
                                      obstacle_offset, remove_near_domain, remove_distance, invert_contact,
                                      contact_threshold);
    SurfaceMesh *m = new SurfaceMesh();
    V3 off(domain_offset[0], domain_offset[1], domain_offset[2]);
    surface_mesh(to_v3(pos, n), solid, I, J, K, dx, marker_particle_radius(dx), sp, off, domain_scale, *m);
    return m;
}

void orc_mesh_get(void *h, int *nverts, int *ntris, float *verts, int *tris) {
    const SurfaceMesh &m = *(SurfaceMesh *)h;
    *nverts = (int)m.verts.size();
    *ntris = (int)(m.tris.size() / 3);
    if (verts)
        for (size_t v = 0; v < m.verts.size(); v++) {
            verts[3 * v] = m.verts[v].x; verts[3 * v + 1] = m.verts[v].y; verts[3 * v + 2] = m.verts[v].z;
        }
    if (tris) std::memcpy(tris, m.tris.data(), sizeof(int) * m.tris.size());
}

void orc_mesh_free(void *h) { delete (SurfaceMesh *)h; }

// The counting loop of FluidSimulation::_removeMarkerParticles
// (fluidsimulation.cpp:4325-4343) alone: particles in order, skipping the
// removed ones, each cell counting up to 250.
void orc_removal_counts(int I, int J, int K, double dx, const float *pos, int n, const uint8_t *removed,
                        int32_t *out) {
    Grid<int> count(I, J, K, 0);
    for (int i = 0; i < n; i++) {
        if (removed && removed[i]) continue;
        Idx g = pos_to_index(V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]), dx);
        if (!count.in(g.i, g.j, g.k)) continue;  // outside the grid: no cell
        if (count(g.i, g.j, g.k) >= 250) continue;
        count(g.i, g.j, g.k) += 1;
    }
    copy_out(count, out);
}

int orc_sim_frame_substeps(void *h) { return ((Sim *)h)->lastSubsteps; }
void orc_sim_pressure_info(void *h, int *iters, double *err, int *ok) {
    Sim *S = (Sim *)h;
    if (iters) *iters = S->lastIters;
    if (err) *err = S->lastErr;
    if (ok) *ok = S->lastOk;
}

}  // extern "C"
