// Surface reconstruction on the device: FluidSimulation's isomesh
// (fluidsimulation.cpp:5134-5243) -- the particle scalar field of
// ParticleMesher, marching cubes with the solid clamp of Polygonizer3d, and
// the TriangleMesh post-processing -- for one polygonizer slice.
#pragma once

#include "scene.h"

namespace flipmi {

// Device buffers of one surface build; grown as needed and reused.
struct SurfaceStore {
    DBuf<uint8_t> seed, active;   // 10^3 node blocks: particles / feathered
    DBuf<int> small;              // block bounds and totals read back
    DBuf<float> solid_copy;       // solid SDF with the obstacle meshing offset
    DBuf<float> field;            // chunk nodes: field, then solid-clamped values
    DBuf<uint8_t> cube;           // per chunk cell: marching-cubes case
    DBuf<int> vcnt, tcnt, vbase, tbase, scan_tmp;
    DBuf<int> edges;              // vertex id per cut edge, 3 node grids (U, V, W)
    DBuf<float> verts, verts_alt; // xyz
    DBuf<int> tris, tris_alt;     // three vertex ids per triangle
    DBuf<int> adj_cnt, adj_off, adj_list, keep, kpos;
    int64_t nverts = 0, ntris = 0;
    bool valid = false;
    size_t device_bytes() const;
};

struct SurfaceInputs {
    GridGeom g;
    const float4 *spos = nullptr;    // positions in half-cell order (x, y, z, id)
    const int *sub_offset = nullptr; // half-cell bins (8 I J K + 1)
    int n = 0;
    const float *solid = nullptr;    // nodal solid SDF of the simulation
    flipmi_surface_desc desc{};
    double kernel_radius = 0.0;      // marker particle radius x particle scale
    float domain_offset[3] = {0.0f, 0.0f, 0.0f};
    double domain_scale = 1.0;
};

// FluidSimulation::_initializeParticleRadii (fluidsimulation.cpp:2644-2648).
double marker_particle_radius(double dx);
// Argument checks (the reference's setter errors plus this build's limits);
// throws StatusError.
void validate_surface_desc(const flipmi_surface_desc &d, const GridGeom &g);
// Builds the mesh into st on stream s; synchronizes s to read the sizes.
void build_surface(const SurfaceInputs &in, SurfaceStore &st, hipStream_t s);
// Copies the mesh to host memory (either pointer may be null).
void copy_surface(const SurfaceStore &st, float *verts_xyz, int32_t *tris, hipStream_t s);

}  // namespace flipmi
