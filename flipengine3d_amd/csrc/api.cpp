// extern "C" entry points of libflipmi (declared in include/flipmi.h).
//
// Scene entry points forward to flipmi::Scene; stage drop-ins copy the
// caller's host buffers to the device, run the same kernels the scene uses
// and copy results back.  No exception crosses the ABI: every failure is a
// flipmi_status plus a message in flipmi_last_error().
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "scene.h"
#include "surface.h"

using namespace flipmi;

struct flipmi_scene {
    Scene s;
};

namespace {

int fail(Ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

template <class F>
int guarded(Ctx *ctx, F &&fn) {
    try {
        fn();
        return FLIPMI_OK;
    } catch (const StatusError &e) {
        return fail(ctx, e.code, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(ctx, FLIPMI_ERR_OOM, "host allocation failed");
    } catch (...) {
        return fail(ctx, FLIPMI_ERR_HIP, "unexpected error");
    }
}

void hipck(hipError_t e, const char *what) {
    if (e != hipSuccess) throw StatusError{FLIPMI_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

template <class T>
void upload(DBuf<T> &d, const T *h, size_t n, hipStream_t s) {
    if (d.alloc(n) != hipSuccess) throw StatusError{FLIPMI_ERR_OOM, "out of device memory"};
    if (n) hipck(hipMemcpyAsync(d.p, h, sizeof(T) * n, hipMemcpyHostToDevice, s), "upload");
}
template <class T>
void alloc_dev(DBuf<T> &d, size_t n) {
    if (d.alloc(n) != hipSuccess) throw StatusError{FLIPMI_ERR_OOM, "out of device memory"};
}
template <class T>
void download(T *h, const DBuf<T> &d, size_t n, hipStream_t s) {
    if (n) hipck(hipMemcpyAsync(h, d.p, sizeof(T) * n, hipMemcpyDeviceToHost, s), "download");
}

// AoS float3 on the host -> three SoA device arrays.
void upload_soa(DBuf<float> out[3], const float *xyz, int64_t n, hipStream_t s) {
    std::vector<float> t((size_t)n);
    for (int a = 0; a < 3; a++) {
        for (int64_t q = 0; q < n; q++) t[q] = xyz[3 * q + a];
        upload(out[a], t.data(), (size_t)n, s);
        hipck(hipStreamSynchronize(s), "sync");
    }
}

}  // namespace

extern "C" {

void flipmi_scene_desc_default(flipmi_scene_desc *d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->cfl = 5.0;
    d->min_substeps = 1;
    d->max_substeps = 6;
    d->pic_ratio = 0.05;
    d->pcg_tolerance = 1e-9;
    d->pcg_max_iterations = 1000;
    d->pcg_precision = FLIPMI_PCG_FP64;
}

int flipmi_ctx_create(int device, flipmi_ctx **out) {
    if (!out) return FLIPMI_ERR_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return FLIPMI_ERR_HIP;
    if (device < 0 || device >= count) return FLIPMI_ERR_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return FLIPMI_ERR_HIP;
    flipmi_ctx *c = new (std::nothrow) flipmi_ctx();
    if (!c) return FLIPMI_ERR_OOM;
    c->c.device = device;
    if (hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return FLIPMI_ERR_HIP;
    }
    // Device address of this device's flipmi_extrapolate_noalloc scratch.
    if (noalloc_scratch_init(device) != hipSuccess) {
        (void)hipStreamDestroy(c->c.stream);
        delete c;
        return FLIPMI_ERR_HIP;
    }
    *out = c;
    return FLIPMI_OK;
}

int flipmi_ctx_destroy(flipmi_ctx *ctx) {
    if (!ctx) return FLIPMI_ERR_INVALID_ARG;
    (void)hipStreamSynchronize(ctx->c.stream);
    (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
    return FLIPMI_OK;
}

int flipmi_ctx_synchronize(flipmi_ctx *ctx) {
    if (!ctx) return FLIPMI_ERR_INVALID_ARG;
    return guarded(&ctx->c, [&] {
        hipck(hipSetDevice(ctx->c.device), "set device");
        hipck(hipDeviceSynchronize(), "device synchronize");
    });
}

const char *flipmi_last_error(const flipmi_ctx *ctx) { return ctx ? ctx->c.last_error.c_str() : ""; }

int flipmi_scene_create(flipmi_ctx *ctx, const flipmi_scene_desc *desc, flipmi_scene **out) {
    if (!ctx || !desc || !out) return FLIPMI_ERR_INVALID_ARG;
    *out = nullptr;
    return guarded(&ctx->c, [&] {
        validate_scene_desc(*desc);
        GridGeom g = make_geom(desc->isize, desc->jsize, desc->ksize, desc->dx);
        std::unique_ptr<flipmi_scene> sc(new flipmi_scene());
        Scene &s = sc->s;
        s.ctx = &ctx->c;
        s.desc = *desc;
        s.g = g;
        hipck(hipSetDevice(ctx->c.device), "set device");
        hipck(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "stream create");
        s.own_stream = true;
        *out = sc.release();
    });
}

int flipmi_scene_destroy(flipmi_scene *scene) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    delete scene;
    return FLIPMI_OK;
}

int flipmi_scene_load_particles(flipmi_scene *scene, const float *pos, const float *vel, int64_t n) {
    if (!scene || (n > 0 && !pos) || n < 0) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    if (s.initialized) return fail(s.ctx, FLIPMI_ERR_INVALID_ARG, "particles must be loaded before initialize");
    return guarded(s.ctx, [&] { s.load(pos, vel, n); });
}

int flipmi_scene_initialize(flipmi_scene *scene) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    return guarded(s.ctx, [&] {
        hipck(hipSetDevice(s.ctx->device), "set device");
        s.initialize();
    });
}

int flipmi_scene_update(flipmi_scene *scene, double dt, flipmi_frame_stats *stats) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    return guarded(s.ctx, [&] {
        hipck(hipSetDevice(s.ctx->device), "set device");
        flipmi_frame_stats st = s.update(dt);
        if (stats) *stats = st;
    });
}

int flipmi_scene_num_particles(flipmi_scene *scene, int64_t *n) {
    if (!scene || !n) return FLIPMI_ERR_INVALID_ARG;
    // Loaded particles join the scene at initialize (the reference counts
    // none before).
    *n = scene->s.initialized ? scene->s.n : 0;
    return FLIPMI_OK;
}

int flipmi_scene_set_domain(flipmi_scene *scene, double ox, double oy, double oz, double scale) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    s.domain_offset[0] = (float)ox;
    s.domain_offset[1] = (float)oy;
    s.domain_offset[2] = (float)oz;
    s.domain_scale = scale;
    return FLIPMI_OK;
}

int flipmi_scene_add_obstacle(flipmi_scene *scene, const float *verts, int nv, const int *tris, int nt,
                              int *obstacle_id) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    return guarded(s.ctx, [&] {
        int id = s.add_obstacle(verts, nv, tris, nt);
        if (obstacle_id) *obstacle_id = id;
    });
}

int flipmi_scene_set_obstacle(flipmi_scene *scene, int id, int enabled) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    if (id < 0 || id >= (int)s.obstacles.size()) return fail(s.ctx, FLIPMI_ERR_INVALID_ARG, "unknown obstacle");
    bool on = enabled != 0;
    if (s.obstacles[(size_t)id]->enabled != on) {
        s.obstacles[(size_t)id]->enabled = on;
        s.solid_ready = false;
    }
    return FLIPMI_OK;
}

int flipmi_scene_set_obstacle_inversed(flipmi_scene *scene, int id, int inversed) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    if (id < 0 || id >= (int)s.obstacles.size()) return fail(s.ctx, FLIPMI_ERR_INVALID_ARG, "unknown obstacle");
    s.obstacles[(size_t)id]->inversed = inversed != 0;
    return FLIPMI_OK;
}

int flipmi_scene_add_inflow_source(flipmi_scene *scene, const float *verts, int nv, const int *tris, int nt,
                                   const double velocity[3], int *source_id) {
    if (!scene || !velocity) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    return guarded(s.ctx, [&] {
        int id = s.add_inflow(verts, nv, tris, nt, velocity);
        if (source_id) *source_id = id;
    });
}

int flipmi_scene_get_inflow_source(flipmi_scene *scene, int id, int dims[6], float *sdf, uint8_t *cells) {
    if (!scene || !dims) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    if (!s.initialized) return fail(s.ctx, FLIPMI_ERR_NOT_INITIALIZED, "scene not initialized");
    if (id < 0 || id >= (int)s.sources.size()) return fail(s.ctx, FLIPMI_ERR_INVALID_ARG, "unknown inflow source");
    return guarded(s.ctx, [&] {
        s.ensure_inflow();
        const InflowSource &src = *s.sources[(size_t)id];
        dims[0] = src.go[0]; dims[1] = src.go[1]; dims[2] = src.go[2];
        dims[3] = src.ni; dims[4] = src.nj; dims[5] = src.nk;
        if (sdf) s.check(hipMemcpyAsync(sdf, src.sdf.p, src.sdf.bytes(), hipMemcpyDeviceToHost, s.stream), "copy");
        if (cells) s.check(hipMemcpyAsync(cells, src.cells.p, src.cells.bytes(), hipMemcpyDeviceToHost, s.stream), "copy");
        s.sync();
    });
}

int flipmi_scene_set_inflow_source(flipmi_scene *scene, int id, int enabled, int constrain, int emissions) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    if (id < 0 || id >= (int)s.sources.size()) return fail(s.ctx, FLIPMI_ERR_INVALID_ARG, "unknown inflow source");
    InflowSource &src = *s.sources[(size_t)id];
    src.enabled = enabled != 0;
    src.constrain = constrain != 0;
    src.emissions = emissions < 0 ? 0 : emissions;  // meshfluidsource.cpp:63-68
    return FLIPMI_OK;
}

int flipmi_scene_set_outflow_source(flipmi_scene *scene, int id, int outflow, int inversed) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    if (id < 0 || id >= (int)s.sources.size()) return fail(s.ctx, FLIPMI_ERR_INVALID_ARG, "unknown inflow source");
    s.sources[(size_t)id]->outflow = outflow != 0;
    s.sources[(size_t)id]->inversed = inversed != 0;
    return FLIPMI_OK;
}

int flipmi_scene_set_marker_jitter(flipmi_scene *scene, double factor, int jitter_surface, uint64_t seed) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    // The reference accepts any factor >= 0 (fluidsimulation.cpp:185-196);
    // factors above 1 can move a particle out of its subcell, which the
    // parallel emission does not order the way the reference's serial loop
    // does, so they are rejected.
    if (!(factor >= 0.0 && factor <= 1.0))
        return fail(s.ctx, FLIPMI_ERR_INVALID_ARG, "marker particle jitter factor must be in [0, 1]");
    s.jitter_factor = factor;
    s.jitter_surface = jitter_surface != 0;
    s.jitter_seed = seed;
    return FLIPMI_OK;
}

int flipmi_scene_get_particles_range(flipmi_scene *scene, int64_t start, int64_t end, float *pos, float *vel) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    if (!s.initialized) return fail(s.ctx, FLIPMI_ERR_NOT_INITIALIZED, "scene not initialized");
    if (start < 0 || end > (int64_t)s.n || start > end)
        return fail(s.ctx, FLIPMI_ERR_INVALID_ARG, "invalid particle range");
    const int64_t m = end - start;
    if (m == 0) return FLIPMI_OK;
    return guarded(s.ctx, [&] {
        std::vector<float> t((size_t)m);
        for (int a = 0; a < 6; a++) {
            float *dst = a < 3 ? pos : vel;
            if (!dst) continue;
            hipck(hipMemcpyAsync(t.data(), s.part[a].p + start, sizeof(float) * m, hipMemcpyDeviceToHost, s.stream),
                  "download");
            hipck(hipStreamSynchronize(s.stream), "sync");
            for (int64_t q = 0; q < m; q++) dst[3 * (size_t)q + (a % 3)] = t[q];
        }
        if (pos) s.to_domain(pos, m);
    });
}

int flipmi_scene_get_particles(flipmi_scene *scene, float *pos, float *vel) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    if (!s.initialized) return fail(s.ctx, FLIPMI_ERR_NOT_INITIALIZED, "scene not initialized");
    return guarded(s.ctx, [&] {
        std::vector<float> t((size_t)s.n);
        for (int a = 0; a < 6; a++) {
            float *dst = a < 3 ? pos : vel;
            if (!dst || s.n == 0) continue;
            hipck(hipMemcpyAsync(t.data(), s.part[a].p, sizeof(float) * s.n, hipMemcpyDeviceToHost, s.stream),
                  "download");
            hipck(hipStreamSynchronize(s.stream), "sync");
            for (int q = 0; q < s.n; q++) dst[3 * (size_t)q + (a % 3)] = t[q];
        }
        if (pos) s.to_domain(pos, s.n);
    });
}

void flipmi_surface_desc_default(flipmi_surface_desc *d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->enabled = 1;                    // fluidsimulation.h:1586
    d->subdivisions = 1;               // :1605
    d->polygonizer_slices = 1;         // :1606
    d->smoothing_value = 0.5;          // :1607
    d->smoothing_iterations = 2;       // :1608
    d->min_polyhedron_triangles = 0;   // :1609
    d->particle_scale = 3.0;           // :1611
    d->obstacle_meshing_offset = 0.0;  // :1593
    d->remove_near_domain = 0;         // :1594
    d->remove_near_domain_distance = 0;
    d->invert_contact_normals = 0;     // :1588
}

int flipmi_scene_set_surface(flipmi_scene *scene, const flipmi_surface_desc *desc) {
    if (!scene || !desc) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    return guarded(s.ctx, [&] {
        validate_surface_desc(*desc, s.g);
        s.surface = *desc;
    });
}

int flipmi_scene_get_isomesh(flipmi_scene *scene, int64_t *num_vertices, int64_t *num_triangles,
                             float *vertices_xyz, int32_t *triangles) {
    if (!scene || !num_vertices || !num_triangles) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    return guarded(s.ctx, [&] {
        const bool have = s.mesh && s.mesh->valid;
        *num_vertices = have ? s.mesh->nverts : 0;
        *num_triangles = have ? s.mesh->ntris : 0;
        if (have) copy_surface(*s.mesh, vertices_xyz, triangles, s.stream);
    });
}

int flipmi_scene_set_launch_graphs(flipmi_scene *scene, int enabled) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    scene->s.launch_graphs = enabled != 0;
    return FLIPMI_OK;
}

int flipmi_scene_get_grid(flipmi_scene *scene, int which, void *out) {
    if (!scene || !out) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    if (!s.initialized) return fail(s.ctx, FLIPMI_ERR_NOT_INITIALIZED, "scene not initialized");
    return guarded(s.ctx, [&] {
        const void *src = nullptr;
        size_t bytes = 0;
        switch (which) {
            case FLIPMI_GRID_LIQUID_PHI: src = s.liquid.p; bytes = s.liquid.bytes(); break;
            case FLIPMI_GRID_SOLID_PHI: src = s.solid.p; bytes = s.solid.bytes(); break;
            case FLIPMI_GRID_U: src = s.u.p; bytes = s.u.bytes(); break;
            case FLIPMI_GRID_V: src = s.v.p; bytes = s.v.bytes(); break;
            case FLIPMI_GRID_W: src = s.w.p; bytes = s.w.bytes(); break;
            case FLIPMI_GRID_WEIGHT_U: src = s.wU.p; bytes = s.wU.bytes(); break;
            case FLIPMI_GRID_WEIGHT_V: src = s.wV.p; bytes = s.wV.bytes(); break;
            case FLIPMI_GRID_WEIGHT_W: src = s.wW.p; bytes = s.wW.bytes(); break;
            case FLIPMI_GRID_WEIGHT_C: src = s.wC.p; bytes = s.wC.bytes(); break;
            case FLIPMI_GRID_VALID_U: src = s.vu.p; bytes = s.vu.bytes(); break;
            case FLIPMI_GRID_VALID_V: src = s.vv.p; bytes = s.vv.bytes(); break;
            case FLIPMI_GRID_VALID_W: src = s.vw.p; bytes = s.vw.bytes(); break;
            case FLIPMI_GRID_CELL_COUNT: {
                const size_t nc = (size_t)s.g.I * s.g.J * s.g.K;
                if (!s.removal_counts_ready) {
                    std::memset(out, 0, sizeof(int) * nc);
                    return;
                }
                DBuf<int> counts;
                alloc_dev(counts, nc);
                hipck(launch_removal_counts(s.g, s.count.p, s.removed_in_cell.p, 250, counts.p, s.stream),
                      "removal counts");
                download((int *)out, counts, nc, s.stream);
                hipck(hipStreamSynchronize(s.stream), "sync");
                return;
            }
            default: throw StatusError{FLIPMI_ERR_INVALID_ARG, "unknown grid id"};
        }
        hipck(hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, s.stream), "download");
        hipck(hipStreamSynchronize(s.stream), "sync");
    });
}

int flipmi_scene_stage_times(flipmi_scene *scene, flipmi_stage_times *out, int reset) {
    if (!scene || !out) return FLIPMI_ERR_INVALID_ARG;
    Scene &s = scene->s;
    for (int q = 0; q < FLIPMI_NUM_STAGES; q++) {
        out->ms[q] = s.stage_ms[q];
        out->calls[q] = s.stage_calls[q];
        if (reset) {
            s.stage_ms[q] = 0;
            s.stage_calls[q] = 0;
        }
    }
    return FLIPMI_OK;
}

int flipmi_scene_set_profiling(flipmi_scene *scene, int enable) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    scene->s.profiling = enable < 0 ? 0 : enable;
    return FLIPMI_OK;
}

int flipmi_scene_set_kernel_timing(flipmi_scene *scene, unsigned mask) {
    if (!scene) return FLIPMI_ERR_INVALID_ARG;
    scene->s.kernel_mask = mask;
    return FLIPMI_OK;
}

int flipmi_scene_set_kernel_sampling(flipmi_scene *scene, int every) {
    if (!scene || every < 1) return FLIPMI_ERR_INVALID_ARG;
    scene->s.kernel_every = every;
    return FLIPMI_OK;
}

// ---------------------------------------------------------------------------
// Stage drop-ins
// ---------------------------------------------------------------------------

int flipmi_domain_solid(flipmi_ctx *ctx, int I, int J, int K, double dx, float *phi_nodal, float *wU,
                        float *wV, float *wW, float *wC) {
    if (!ctx || !phi_nodal) return FLIPMI_ERR_INVALID_ARG;
    return guarded(&ctx->c, [&] {
        GridGeom g = make_geom(I, J, K, dx);
        hipStream_t s = ctx->c.stream;
        size_t nn = (size_t)(I + 1) * (J + 1) * (K + 1);
        DBuf<float> phi, dU, dV, dW, dC;
        alloc_dev(phi, nn);
        alloc_dev(dU, (size_t)(I + 1) * J * K); alloc_dev(dV, (size_t)I * (J + 1) * K);
        alloc_dev(dW, (size_t)I * J * (K + 1)); alloc_dev(dC, (size_t)I * J * K);
        BoxMesh mesh = make_domain_box(I, J, K, dx);
        hipck(launch_domain_solid(g, mesh, phi.p, s), "domain solid");
        hipck(launch_weights(g, phi.p, dU.p, dV.p, dW.p, dC.p, s), "weights");
        download(phi_nodal, phi, nn, s);
        if (wU) download(wU, dU, dU.n, s);
        if (wV) download(wV, dV, dV.n, s);
        if (wW) download(wW, dW, dW.n, s);
        if (wC) download(wC, dC, dC.n, s);
        hipck(hipStreamSynchronize(s), "sync");
    });
}

namespace {
// Sorts host particles on the device -- by cell, or into half-cell bins as
// the scene's step does -- and returns sorted SoA arrays.
struct SortedParticles {
    DBuf<float> raw[6], svel[3];
    DBuf<float4> spos;
    DBuf<int> cell, slot, count, offset, order, scan_tmp;
    void build(const GridGeom &g, const float *pos, const float *vel, int64_t n, hipStream_t s,
               bool half_cells = false) {
        if (n > INT32_MAX / 2) throw StatusError{FLIPMI_ERR_INVALID_ARG, "too many particles"};
        upload_soa(raw, pos, n, s);
        if (vel) upload_soa(raw + 3, vel, n, s);
        else for (int a = 3; a < 6; a++) { alloc_dev(raw[a], (size_t)std::max<int64_t>(n, 1)); hipck(launch_fill_f32(raw[a].p, raw[a].n, 0.0f, s), "fill"); }
        const size_t nbins = half_cells ? subcell_bins(g) : (size_t)g.I * g.J * g.K;
        if (nbins >= (size_t)INT32_MAX) throw StatusError{FLIPMI_ERR_INVALID_ARG, "grid too large"};
        size_t pn = (size_t)std::max<int64_t>(n, 1);
        alloc_dev(spos, pn);
        for (int a = 0; a < 3; a++) alloc_dev(svel[a], pn);
        alloc_dev(cell, pn); alloc_dev(slot, pn); alloc_dev(order, pn);
        alloc_dev(count, nbins); alloc_dev(offset, nbins + 1);
        alloc_dev(scan_tmp, std::max(scan_scratch_elems((int)nbins), scan_scratch_elems((int)pn)));
        SortBuffers sb{cell.p, slot.p, count.p, offset.p, order.p, scan_tmp.p, scan_tmp.n};
        if (half_cells) hipck(launch_sort_by_subcell(g, raw[0].p, raw[1].p, raw[2].p, (int)n, sb, s), "sort");
        else hipck(launch_sort_by_cell(g, raw[0].p, raw[1].p, raw[2].p, (int)n, sb, s), "sort");
        const float *src[6];
        for (int a = 0; a < 6; a++) src[a] = raw[a].p;
        float *sv[3] = {svel[0].p, svel[1].p, svel[2].p};
        hipck(launch_gather_sorted(order.p, offset.p + nbins, (int)n, src, spos.p, sv, s), "gather");
    }
};
}  // namespace

int flipmi_liquid_sdf(flipmi_ctx *ctx, int I, int J, int K, double dx, const float *pos, int64_t n,
                      double radius, const float *solid_phi_nodal, float *phi_out) {
    if (!ctx || !phi_out || (n > 0 && !pos) || n < 0) return FLIPMI_ERR_INVALID_ARG;
    return guarded(&ctx->c, [&] {
        GridGeom g = make_geom(I, J, K, dx);
        hipStream_t s = ctx->c.stream;
        SortedParticles sp;
        sp.build(g, pos, nullptr, n, s, true);
        size_t nn = (size_t)(I + 1) * (J + 1) * (K + 1);
        DBuf<float> solid, phi;
        if (solid_phi_nodal) upload(solid, solid_phi_nodal, nn, s);
        alloc_dev(phi, (size_t)I * J * K);
        hipck(launch_liquid_sdf(g, sp.spos.p, sp.offset.p, radius, solid.p, phi.p, s), "liquid sdf");
        download(phi_out, phi, phi.n, s);
        hipck(hipStreamSynchronize(s), "sync");
    });
}

int flipmi_removal_counts(flipmi_ctx *ctx, int I, int J, int K, double dx, const float *pos, int64_t n,
                          const uint8_t *removed, int32_t *counts_out) {
    if (!ctx || !counts_out || (n > 0 && !pos) || n < 0) return FLIPMI_ERR_INVALID_ARG;
    return guarded(&ctx->c, [&] {
        GridGeom g = make_geom(I, J, K, dx);
        hipStream_t s = ctx->c.stream;
        SortedParticles sp;
        sp.build(g, pos, nullptr, n, s);
        const size_t nc = (size_t)I * J * K;
        DBuf<uint8_t> rm;
        DBuf<int> ric, out;
        if (removed) {
            upload(rm, removed, (size_t)n, s);
        } else {
            alloc_dev(rm, (size_t)std::max<int64_t>(n, 1));
            hipck(launch_fill_u8(rm.p, rm.n, 0, s), "fill");
        }
        alloc_dev(ric, nc);
        alloc_dev(out, nc);
        hipck(launch_removed_in_cell(g, sp.raw[0].p, sp.raw[1].p, sp.raw[2].p, rm.p, (int)n, ric.p, s),
              "removed in cell");
        hipck(launch_removal_counts(g, sp.count.p, ric.p, 250, out.p, s), "removal counts");
        download(counts_out, out, nc, s);
        hipck(hipStreamSynchronize(s), "sync");
    });
}

struct flipmi_mesh {
    flipmi_ctx *ctx = nullptr;
    SurfaceStore st;
};

int flipmi_mesh_particles(flipmi_ctx *ctx, int I, int J, int K, double dx, const float *pos, int64_t n,
                          const float *solid_phi_nodal, const flipmi_surface_desc *desc, double kernel_radius,
                          const double *domain_offset, double domain_scale, flipmi_mesh **out) {
    if (!ctx || !out || !desc || !solid_phi_nodal || (n > 0 && !pos) || n < 0 || kernel_radius < 0.0)
        return FLIPMI_ERR_INVALID_ARG;
    *out = nullptr;
    flipmi_mesh *m = new flipmi_mesh();
    m->ctx = ctx;
    int rc = guarded(&ctx->c, [&] {
        GridGeom g = make_geom(I, J, K, dx);
        validate_surface_desc(*desc, g);
        // The mesher bins every particle into the grid's blocks
        // (particlemesher.cpp:113-118): positions must lie in the grid.
        for (int64_t q = 0; q < n; q++)
            for (int a = 0; a < 3; a++) {
                const double x = (double)pos[3 * q + a] * g.invdx;
                const int hi = a == 0 ? I : (a == 1 ? J : K);
                if (!(x >= 0.0 && x < (double)hi))
                    throw StatusError{FLIPMI_ERR_INVALID_ARG, "particle positions must lie inside the grid"};
            }
        hipStream_t s = ctx->c.stream;
        SortedParticles sp;
        sp.build(g, pos, nullptr, n, s, true);
        DBuf<float> solid;
        upload(solid, solid_phi_nodal, (size_t)(I + 1) * (J + 1) * (K + 1), s);
        SurfaceInputs in;
        in.g = g;
        in.spos = sp.spos.p;
        in.sub_offset = sp.offset.p;
        in.n = (int)n;
        in.solid = solid.p;
        in.desc = *desc;
        in.kernel_radius = kernel_radius > 0.0 ? kernel_radius : marker_particle_radius(dx) * desc->particle_scale;
        for (int a = 0; a < 3; a++) in.domain_offset[a] = domain_offset ? (float)domain_offset[a] : 0.0f;
        in.domain_scale = domain_scale;
        build_surface(in, m->st, s);
        hipck(hipStreamSynchronize(s), "sync");
    });
    if (rc != FLIPMI_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return FLIPMI_OK;
}

int flipmi_mesh_counts(flipmi_mesh *mesh, int64_t *num_vertices, int64_t *num_triangles) {
    if (!mesh || !num_vertices || !num_triangles) return FLIPMI_ERR_INVALID_ARG;
    *num_vertices = mesh->st.nverts;
    *num_triangles = mesh->st.ntris;
    return FLIPMI_OK;
}

int flipmi_mesh_copy(flipmi_mesh *mesh, float *vertices_xyz, int32_t *triangles) {
    if (!mesh) return FLIPMI_ERR_INVALID_ARG;
    return guarded(&mesh->ctx->c, [&] { copy_surface(mesh->st, vertices_xyz, triangles, mesh->ctx->c.stream); });
}

int flipmi_mesh_destroy(flipmi_mesh *mesh) {
    delete mesh;
    return FLIPMI_OK;
}

int flipmi_liquid_sdf_postprocess(flipmi_ctx *ctx, int I, int J, int K, double dx,
                                  const float *solid_phi_nodal, float *phi_inout) {
    if (!ctx || !solid_phi_nodal || !phi_inout) return FLIPMI_ERR_INVALID_ARG;
    return guarded(&ctx->c, [&] {
        GridGeom g = make_geom(I, J, K, dx);
        hipStream_t s = ctx->c.stream;
        size_t nn = (size_t)(I + 1) * (J + 1) * (K + 1);
        DBuf<float> solid, phi;
        upload(solid, solid_phi_nodal, nn, s);
        upload(phi, phi_inout, (size_t)I * J * K, s);
        hipck(launch_liquid_postprocess(g, solid.p, phi.p, s), "liquid post-process");
        download(phi_inout, phi, phi.n, s);
        hipck(hipStreamSynchronize(s), "sync");
    });
}

int flipmi_p2g(flipmi_ctx *ctx, int I, int J, int K, double dx, const float *pos, const float *vel,
               int64_t n, double radius, float *u, float *v, float *w, uint8_t *vu, uint8_t *vv,
               uint8_t *vw) {
    if (!ctx || !u || !v || !w || !vu || !vv || !vw || (n > 0 && (!pos || !vel)) || n < 0)
        return FLIPMI_ERR_INVALID_ARG;
    return guarded(&ctx->c, [&] {
        GridGeom g = make_geom(I, J, K, dx);
        hipStream_t s = ctx->c.stream;
        SortedParticles sp;
        sp.build(g, pos, vel, n, s, true);
        DBuf<float> du, dv, dw;
        DBuf<uint8_t> dvu, dvv, dvw;
        DBuf<int> ovf, ovf_count;
        alloc_dev(du, (size_t)(I + 1) * J * K); alloc_dev(dv, (size_t)I * (J + 1) * K);
        alloc_dev(dw, (size_t)I * J * (K + 1));
        alloc_dev(dvu, du.n); alloc_dev(dvv, dv.n); alloc_dev(dvw, dw.n);
        alloc_dev(ovf, p2g_face_count(g)); alloc_dev(ovf_count, 1);
        const float *svel[3] = {sp.svel[0].p, sp.svel[1].p, sp.svel[2].p};
        float *outs[3] = {du.p, dv.p, dw.p};
        uint8_t *valids[3] = {dvu.p, dvv.p, dvw.p};
        hipck(launch_p2g(g, sp.spos.p, svel, sp.offset.p, (int)n, radius, outs, valids, ovf_count.p, ovf.p, s),
              "p2g");
        download(u, du, du.n, s); download(v, dv, dv.n, s); download(w, dw, dw.n, s);
        download(vu, dvu, dvu.n, s); download(vv, dvv, dvv.n, s); download(vw, dvw, dvw.n, s);
        hipck(hipStreamSynchronize(s), "sync");
    });
}

int flipmi_extrapolate(flipmi_ctx *ctx, int ni, int nj, int nk, float *grid, const uint8_t *valid,
                       int layers) {
    if (!ctx || !grid || !valid || ni < 1 || nj < 1 || nk < 1 || layers < 0 || layers > 250)
        return FLIPMI_ERR_INVALID_ARG;
    if ((double)ni * nj * nk >= 4294967296.0) return FLIPMI_ERR_INVALID_ARG;  // 32-bit kernel indexing
    return guarded(&ctx->c, [&] {
        hipStream_t s = ctx->c.stream;
        size_t n = (size_t)ni * nj * nk;
        DBuf<float> dg;
        DBuf<uint8_t> dv, st;
        upload(dg, grid, n, s);
        upload(dv, valid, n, s);
        alloc_dev(st, n);
        hipck(launch_extrapolate(ni, nj, nk, dg.p, dv.p, st.p, layers, s), "extrapolate");
        download(grid, dg, n, s);
        hipck(hipStreamSynchronize(s), "sync");
    });
}

int flipmi_pressure_solve(flipmi_ctx *ctx, const flipmi_pressure_args *a, flipmi_pressure_report *rep) {
    if (!ctx || !a || !a->liquid_phi || !a->weight_u || !a->weight_v || !a->weight_w || !a->weight_c ||
        !a->u || !a->v || !a->w || !a->valid_u || !a->valid_v || !a->valid_w)
        return FLIPMI_ERR_INVALID_ARG;
    bool ok = true;
    int rc = guarded(&ctx->c, [&] {
        GridGeom g = make_geom(a->isize, a->jsize, a->ksize, a->dx);
        if (a->max_iterations < 1) throw StatusError{FLIPMI_ERR_INVALID_ARG, "max_iterations must be positive"};
        hipStream_t s = ctx->c.stream;
        int I = a->isize, J = a->jsize, K = a->ksize;
        size_t nc = (size_t)I * J * K, nU = (size_t)(I + 1) * J * K, nV = (size_t)I * (J + 1) * K,
               nW = (size_t)I * J * (K + 1);
        DBuf<float> liq, wu, wv, ww, wc, su, sv, sw, u, v, w;
        DBuf<uint8_t> vu, vv, vw;
        upload(liq, a->liquid_phi, nc, s);
        upload(wu, a->weight_u, nU, s); upload(wv, a->weight_v, nV, s);
        upload(ww, a->weight_w, nW, s); upload(wc, a->weight_c, nc, s);
        alloc_dev(su, nU); alloc_dev(sv, nV); alloc_dev(sw, nW);
        if (a->solid_u) hipck(hipMemcpyAsync(su.p, a->solid_u, su.bytes(), hipMemcpyHostToDevice, s), "upload");
        else hipck(launch_fill_f32(su.p, nU, 0.0f, s), "fill");
        if (a->solid_v) hipck(hipMemcpyAsync(sv.p, a->solid_v, sv.bytes(), hipMemcpyHostToDevice, s), "upload");
        else hipck(launch_fill_f32(sv.p, nV, 0.0f, s), "fill");
        if (a->solid_w) hipck(hipMemcpyAsync(sw.p, a->solid_w, sw.bytes(), hipMemcpyHostToDevice, s), "upload");
        else hipck(launch_fill_f32(sw.p, nW, 0.0f, s), "fill");
        upload(u, a->u, nU, s); upload(v, a->v, nV, s); upload(w, a->w, nW, s);
        upload(vu, a->valid_u, nU, s); upload(vv, a->valid_v, nV, s); upload(vw, a->valid_w, nW, s);
        // _conditionSolidVelocityField (pressuresolver.cpp:124-268) precedes the
        // assembly; with all solid velocities zero it changes nothing.
        const bool solid_given = a->solid_u || a->solid_v || a->solid_w;
        if (solid_given) {
            DBuf<int> parent;
            DBuf<unsigned> flags;
            alloc_dev(parent, nc);
            alloc_dev(flags, nc);
            hipck(launch_condition_solid(g, liq.p, wu.p, wv.p, ww.p, su.p, sv.p, sw.p, parent.p, flags.p, s),
                  "solid velocity conditioning");
            hipck(hipStreamSynchronize(s), "sync");
        }
        if (a->precision != FLIPMI_PCG_FP64 && a->precision != FLIPMI_PCG_MIXED)
            throw StatusError{FLIPMI_ERR_INVALID_ARG, "unknown pcg precision"};
        PressureStore ps;
        ps.alloc(g, s, a->precision == FLIPMI_PCG_MIXED);
        PressureWork pw;
        pw.g = g; pw.dt = a->dt; pw.tol = a->tolerance; pw.acceptable = a->acceptable_tolerance;
        pw.maxiter = a->max_iterations; pw.precision = a->precision;
        pw.liquid = liq.p; pw.wU = wu.p; pw.wV = wv.p; pw.wW = ww.p; pw.wC = wc.p;
        pw.sU = su.p; pw.sV = sv.p; pw.sW = sw.p;
        pw.u = u.p; pw.v = v.p; pw.w = w.p; pw.vu = vu.p; pw.vv = vv.p; pw.vw = vw.p;
        PcgScalars *hs = nullptr;
        hipck(hipHostMalloc((void **)&hs, sizeof(PcgScalars), hipHostMallocDefault), "pinned");
        PressureOutcome o;
        try {
            o = run_pressure(pw, ps.pb, hs, s, nullptr);
        } catch (...) {
            (void)hipHostFree(hs);
            throw;
        }
        (void)hipHostFree(hs);
        ok = o.success;
        download(a->u, u, nU, s); download(a->v, v, nV, s); download(a->w, w, nW, s);
        download(a->valid_u, vu, nU, s); download(a->valid_v, vv, nV, s); download(a->valid_w, vw, nW, s);
        if (a->solid_u) download(a->solid_u, su, nU, s);
        if (a->solid_v) download(a->solid_v, sv, nV, s);
        if (a->solid_w) download(a->solid_w, sw, nW, s);
        std::vector<float> hl(nc);
        int rows = 0;
        std::vector<double> x;
        const SlotMap &sm = ps.pb.sm;
        if (a->pressure_rows) {
            x.resize(sm.count());
            hipck(hipMemcpyAsync(x.data(), ps.pb.x, sm.count() * sizeof(double), hipMemcpyDeviceToHost, s),
                  "download");
        }
        download(hl.data(), liq, nc, s);
        hipck(hipStreamSynchronize(s), "sync");
        // Rows in reference order: interior cells, k outer, j, i inner, phi < 0.
        for (int k = 1; k < K - 1; k++)
            for (int j = 1; j < J - 1; j++)
                for (int i = 1; i < I - 1; i++) {
                    size_t c = (size_t)i + (size_t)I * ((size_t)j + (size_t)J * k);
                    if (hl[c] < 0) {
                        if (a->pressure_rows) a->pressure_rows[rows] = o.solved ? x[sm.slot(i, j, k)] : 0.0;
                        rows++;
                    }
                }
        if (rep) {
            rep->iterations = o.iterations;
            rep->residual = o.residual;
            rep->rows = rows;
            rep->success = o.success ? 1 : 0;
            rep->refinements = o.refinements;
        }
    });
    if (rc != FLIPMI_OK) return rc;
    return ok ? FLIPMI_OK : fail(&ctx->c, FLIPMI_ERR_NOT_CONVERGED, "***Pressure Solver FAILED");
}

int flipmi_mic0_apply(flipmi_ctx *ctx, int I, int J, int K, double dx, double dt, const float *liquid_phi,
                      const float *wU, const float *wV, const float *wW, const float *wC,
                      const double *r_rows, double *z_rows, double *invdiag_rows) {
    if (!ctx || !liquid_phi || !wU || !wV || !wW || !wC || !r_rows || !z_rows) return FLIPMI_ERR_INVALID_ARG;
    return guarded(&ctx->c, [&] {
        GridGeom g = make_geom(I, J, K, dx);
        hipStream_t s = ctx->c.stream;
        size_t nc = (size_t)I * J * K, nU = (size_t)(I + 1) * J * K, nV = (size_t)I * (J + 1) * K,
               nW = (size_t)I * J * (K + 1);
        DBuf<float> liq, wu, wv, ww, wc, zero_u, zero_v, zero_w;
        upload(liq, liquid_phi, nc, s);
        upload(wu, wU, nU, s); upload(wv, wV, nV, s); upload(ww, wW, nW, s); upload(wc, wC, nc, s);
        alloc_dev(zero_u, nU); alloc_dev(zero_v, nV); alloc_dev(zero_w, nW);
        hipck(launch_fill_f32(zero_u.p, nU, 0.0f, s), "fill");
        hipck(launch_fill_f32(zero_v.p, nV, 0.0f, s), "fill");
        hipck(launch_fill_f32(zero_w.p, nW, 0.0f, s), "fill");
        PressureStore ps;
        ps.alloc(g, s, false);
        hipck(launch_pressure_reset(ps.pb, s), "reset");
        hipck(launch_pressure_assemble(g, dt, liq.p, wu.p, wv.p, ww.p, wc.p, zero_u.p, zero_v.p, zero_w.p,
                                       zero_u.p, zero_v.p, zero_w.p, ps.pb, s),
              "assemble");
        hipck(launch_group_mask(ps.pb, s), "group mask");
        hipck(launch_mic0_factor(g, liq.p, ps.pb, s), "factor");
        // Scatter r from reference row order into slot order.
        std::vector<float> hl(nc);
        download(hl.data(), liq, nc, s);
        hipck(hipStreamSynchronize(s), "sync");
        const SlotMap &sm = ps.pb.sm;
        size_t ns = sm.count();
        std::vector<double> rd(ns, 0.0);
        std::vector<size_t> rows;
        for (int k = 1; k < K - 1; k++)
            for (int j = 1; j < J - 1; j++)
                for (int i = 1; i < I - 1; i++) {
                    size_t c = (size_t)i + (size_t)I * ((size_t)j + (size_t)J * k);
                    if (hl[c] < 0) {
                        size_t p = sm.slot(i, j, k);
                        rd[p] = r_rows[rows.size()];
                        rows.push_back(p);
                    }
                }
        // r and z live in the store's own slot arrays (with their zero and
        // sink groups); z starts at zero from the allocation.
        hipck(hipMemcpyAsync(ps.pb.r, rd.data(), ns * sizeof(double), hipMemcpyHostToDevice, s), "upload");
        hipck(launch_precond_apply(g, ps.pb, ps.pb.r, ps.pb.z, s), "precond");
        std::vector<double> zd(ns), inv(ns);
        hipck(hipMemcpyAsync(zd.data(), ps.pb.z, ns * sizeof(double), hipMemcpyDeviceToHost, s), "download");
        hipck(hipMemcpyAsync(inv.data(), ps.pb.invdiag, ns * sizeof(double), hipMemcpyDeviceToHost, s), "download");
        hipck(hipStreamSynchronize(s), "sync");
        PcgScalars hs;
        hipck(hipMemcpy(&hs, ps.pb.sc, sizeof(hs), hipMemcpyDeviceToHost), "scalars");
        if (hs.abort) throw StatusError{FLIPMI_ERR_TIMEOUT, "wavefront sweep wait timed out"};
        for (size_t q = 0; q < rows.size(); q++) {
            z_rows[q] = zd[rows[q]];
            if (invdiag_rows) invdiag_rows[q] = inv[rows[q]];
        }
    });
}

int flipmi_advect_particles(flipmi_ctx *ctx, int I, int J, int K, double dx, const float *u, const float *v,
                            const float *w, const float *pos_in, int64_t n, double dt, int order,
                            float *pos_out) {
    if (!ctx || !u || !v || !w || (n > 0 && (!pos_in || !pos_out)) || n < 0 || order < 1 || order > 4)
        return FLIPMI_ERR_INVALID_ARG;
    return guarded(&ctx->c, [&] {
        GridGeom g = make_geom(I, J, K, dx);
        if (n > INT32_MAX / 3) throw StatusError{FLIPMI_ERR_INVALID_ARG, "too many particles"};
        hipStream_t s = ctx->c.stream;
        DBuf<float> du, dv, dw, pin, pout;
        upload(du, u, (size_t)(I + 1) * J * K, s); upload(dv, v, (size_t)I * (J + 1) * K, s);
        upload(dw, w, (size_t)I * J * (K + 1), s);
        upload(pin, pos_in, (size_t)n * 3, s);
        alloc_dev(pout, (size_t)n * 3);
        hipck(launch_advect_rk(g, du.p, dv.p, dw.p, pin.p, (int)n, dt, order, pout.p, s), "advect");
        download(pos_out, pout, (size_t)n * 3, s);
        hipck(hipStreamSynchronize(s), "sync");
    });
}

int flipmi_trilinear_velocity(flipmi_ctx *ctx, int I, int J, int K, double dx, const float *u, const float *v,
                              const float *w, const float *pos, int64_t n, float *vel_out) {
    if (!ctx || !u || !v || !w || (n > 0 && (!pos || !vel_out)) || n < 0) return FLIPMI_ERR_INVALID_ARG;
    return guarded(&ctx->c, [&] {
        GridGeom g = make_geom(I, J, K, dx);
        if (n > INT32_MAX / 3) throw StatusError{FLIPMI_ERR_INVALID_ARG, "too many particles"};
        hipStream_t s = ctx->c.stream;
        DBuf<float> du, dv, dw, pin, pout;
        upload(du, u, (size_t)(I + 1) * J * K, s); upload(dv, v, (size_t)I * (J + 1) * K, s);
        upload(dw, w, (size_t)I * J * (K + 1), s);
        upload(pin, pos, (size_t)n * 3, s);
        alloc_dev(pout, (size_t)n * 3);
        hipck(launch_trilinear_velocity(g, du.p, dv.p, dw.p, pin.p, (int)n, pout.p, s), "trilinear");
        download(vel_out, pout, (size_t)n * 3, s);
        hipck(hipStreamSynchronize(s), "sync");
    });
}

}  // extern "C"
