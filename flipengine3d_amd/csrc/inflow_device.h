// Device tests of the inflow velocity constraints, shared by the particle
// step and the body-force kernels.
#pragma once

#include "device_math.h"
#include "flipmi_kernels.h"

namespace flipmi {

// FluidSimulation::_constrainMarkerParticleVelocities (fluidsimulation.cpp:
// 4109-4170): a particle whose cell is a source cell and where the source
// SDF is not positive takes the source velocity.  The reference samples the
// sub-grid SDF at the particle's grid position without the sub-grid offset
// (fluidsimulation.cpp:4140); that sampling is kept as is.
__device__ inline bool inflow_holds_particle(const InflowView &v, const GridGeom &g, F3 p) {
    I3 c = pos_to_index(p, g.invdx);
    if (c.i < v.cmin[0] || c.j < v.cmin[1] || c.k < v.cmin[2] || c.i > v.cmax[0] ||
        c.j > v.cmax[1] || c.k > v.cmax[2])
        return false;
    if (!v.cells[(size_t)c.i + (size_t)g.I * ((size_t)c.j + (size_t)g.J * c.k)]) return false;
    float d = (float)nodal_trilinear(v.sdf, Dims{v.ni, v.nj, v.nk}, g.dx, g.invdx, p);
    return !(d > 0.0f);
}

// FluidSimulation::_getInflowConstrainedVelocityComponents (fluidsimulation.cpp:
// 3372-3438): a valid face where the source SDF, sampled at the face's grid
// position (again without the sub-grid offset), is negative.
__device__ inline bool inflow_holds_face(const InflowView &v, const GridGeom &g, F3 p) {
    float d = (float)nodal_trilinear(v.sdf, Dims{v.ni, v.nj, v.nk}, g.dx, g.invdx, p);
    return d < 0.0f;
}

}  // namespace flipmi
