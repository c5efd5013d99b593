// Memory-bound grid passes: fills, velocity extrapolation, saving the MAC
// grid with the body force, and the solid velocity constraint.
#include "device_math.h"
#include "flipmi_kernels.h"
#include "inflow_device.h"

namespace flipmi {

template <class T>
__global__ void k_fill(T *p, size_t n, T v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <class T>
static hipError_t fill(T *p, size_t n, T v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    k_fill<T><<<(unsigned)blocks, 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t *p, size_t n, uint8_t v, hipStream_t s) { return fill(p, n, v, s); }
hipError_t launch_fill_f32(float *p, size_t n, float v, hipStream_t s) { return fill(p, n, v, s); }
hipError_t launch_fill_f64(double *p, size_t n, double v, hipStream_t s) { return fill(p, n, v, s); }
hipError_t launch_fill_i32(int *p, size_t n, int v, hipStream_t s) { return fill(p, n, v, s); }

// Extrapolation status codes: 0 = valid on entry, L = filled in layer L,
// kBorder = face on the grid border (always finished, never a source),
// kUnknown = not reached yet.
constexpr uint8_t kBorder = 254;
constexpr uint8_t kUnknown = 255;

__global__ void k_extrap_init(int ni, int nj, int nk, const uint8_t *__restrict__ valid,
                              uint8_t *status) {
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (size_t)ni * nj * nk) return;
    int i, j, k;
    unflatten3(tid, ni, nj, i, j, k);
    uint8_t st;
    if (i == 0 || j == 0 || k == 0 || i == ni - 1 || j == nj - 1 || k == nk - 1) st = kBorder;
    else st = valid[tid] ? 0 : kUnknown;
    status[tid] = st;
}

// One layer of GridUtils::extrapolateGrid (gridutils.cpp:32-228): every
// unreached face next to a face finished in the previous layer takes the
// mean of its finished 6-neighbours, summed in the order +i,-i,+j,-j,+k,-k.
// Faces written in this pass are neither read as sources nor averaged, so
// the in-place update is race free.
__global__ void k_extrap_layer(int ni, int nj, int nk, float *grid, uint8_t *status, int layer) {
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (size_t)ni * nj * nk) return;
    if (status[tid] != kUnknown) return;
    // Unknown faces are never on the border, so all 6 neighbours exist.
    size_t sj = (size_t)ni, sk = (size_t)ni * nj;
    const size_t nb[6] = {tid + 1, tid - 1, tid + sj, tid - sj, tid + sk, tid - sk};
    uint8_t known = (uint8_t)(layer - 1);
    bool touch = false;
#pragma unroll
    for (int q = 0; q < 6; q++) touch |= status[nb[q]] == known;
    if (!touch) return;
    float sum = 0.0f;
    int count = 0;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        uint8_t st = status[nb[q]];
        if (st <= known || st == kBorder) {
            sum += grid[nb[q]];
            count++;
        }
    }
    grid[tid] = sum / (float)count;
    status[tid] = (uint8_t)layer;
}

hipError_t launch_extrapolate(int ni, int nj, int nk, float *grid, const uint8_t *valid,
                              uint8_t *status, int layers, hipStream_t s) {
    size_t n = (size_t)ni * nj * nk;
    unsigned nb = (unsigned)((n + 255) / 256);
    k_extrap_init<<<nb, 256, 0, s>>>(ni, nj, nk, valid, status);
    for (int l = 1; l <= layers; l++) k_extrap_layer<<<nb, 256, 0, s>>>(ni, nj, nk, grid, status, l);
    return hipGetLastError();
}

// _saveVelocityField + _applyConstantBodyForces (fluidsimulation.cpp:
// 3287-3297, 3450-3489): saved = mac, then mac += (float)(g * dt) on every
// face (no inflow faces are constrained in these scenes).
__global__ void k_save_force(size_t n, float *mac, float *saved, float add, int do_add) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = mac[i];
    saved[i] = v;
    if (do_add) mac[i] = v + add;
}

hipError_t launch_save_and_body_force(size_t n, const float *mac, float *saved, float add,
                                      int do_add, hipStream_t s) {
    k_save_force<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, const_cast<float *>(mac), saved,
                                                             add, do_add);
    return hipGetLastError();
}

// The same with the inflow constraint: valid faces held by an inflow source
// take no body force (fluidsimulation.cpp:3372-3438, 3450-3489).  Face
// positions are Grid3d::FaceIndexToPositionU/V/W (grid3d.h:111-133).
__global__ void k_save_force_inflow(GridGeom g, int comp, size_t n, float *mac, float *saved,
                                    float add, int do_add, const uint8_t *__restrict__ valid,
                                    InflowViews views) {
    size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    float v = mac[f];
    saved[f] = v;
    if (!do_add) return;
    if (valid[f]) {
        const int ni = g.I + (comp == 0), nj = g.J + (comp == 1);
        int i, j, k;
        unflatten3(f, ni, nj, i, j, k);
        F3 p = f3(comp == 0 ? (float)((double)(float)i * g.dx) : (float)(((double)(float)i + 0.5) * g.dx),
                  comp == 1 ? (float)((double)(float)j * g.dx) : (float)(((double)(float)j + 0.5) * g.dx),
                  comp == 2 ? (float)((double)(float)k * g.dx) : (float)(((double)(float)k + 0.5) * g.dx));
        for (int s = 0; s < views.n; s++)
            if (views.v[s].constrain && inflow_holds_face(views.v[s], g, p)) return;
    }
    mac[f] = v + add;
}

hipError_t launch_save_and_body_force_inflow(const GridGeom &g, int comp, const float *mac,
                                             float *saved, float add, int do_add,
                                             const uint8_t *valid, const InflowViews &views,
                                             hipStream_t s) {
    size_t n = (size_t)(g.I + (comp == 0)) * (g.J + (comp == 1)) * (g.K + (comp == 2));
    k_save_force_inflow<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, comp, n, const_cast<float *>(mac),
                                                                    saved, add, do_add, valid, views);
    return hipGetLastError();
}

// FluidSimulation::_constrainVelocityFieldThread (fluidsimulation.cpp:3873-3935)
// with the domain boundary's friction f = 0 (fluidsimulation.h:1524).
__global__ void k_constrain(size_t n, float *mac, const float *__restrict__ weight,
                            const float *__restrict__ solid_vel) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float w = weight[i];
    if (w == 0) {
        mac[i] = solid_vel[i];
    } else if (w < 1.0f) {
        const float f = 0.0f;
        mac[i] = f * solid_vel[i] + (1.0f - f) * mac[i];
    }
}

hipError_t launch_constrain(size_t n, float *mac, const float *weight, const float *solid_vel,
                            hipStream_t s) {
    k_constrain<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, mac, weight, solid_vel);
    return hipGetLastError();
}

}  // namespace flipmi
