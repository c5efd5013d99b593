// Host orchestration of the FLIP step on one device stream.
//
// Scene::update mirrors FluidSimulation::update (fluidsimulation.cpp:
// 5755-5846) and Scene::step_fluid mirrors FluidSimulation::_stepFluid
// (fluidsimulation.cpp:5471-5508) for scenes with domain walls and a constant
// body force; every stage runs as HIP kernels on device-resident data.  The
// host reads back a handful of scalars per substep (maximum speed, |rhs|,
// PCG progress, the speed histogram and the surviving particle count).
#include "scene.h"
#include "surface.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace flipmi {

// ---------------------------------------------------------------------------
// Domain geometry
// ---------------------------------------------------------------------------

namespace {
struct HBox {
    float pos[3];
    double ext[3];
};

// AABB(0,0,0, I dx, J dx, K dx).expand(-3 dx - 1e-4) (fluidsimulation.cpp:
// 2834-2839, aabb.cpp:122-128): float min corner, double extents.
HBox boundary_box(int isize, int jsize, int ksize, double dx) {
    HBox b;
    double v = -3 * dx - 1e-4;
    double h = 0.5 * v;
    for (int a = 0; a < 3; a++) b.pos[a] = 0.0f - (float)h;
    b.ext[0] = isize * dx + v;
    b.ext[1] = jsize * dx + v;
    b.ext[2] = ksize * dx + v;
    return b;
}

inline int floor_index(double x, double dx) { return (int)std::floor(x * (1.0 / dx)); }
}  // namespace

BoxMesh make_domain_box(int isize, int jsize, int ksize, double dx) {
    const int band = 3;  // fluidsimulation.h:1580 (_solidLevelSetExactBand)
    HBox bb = boundary_box(isize, jsize, ksize, dx);
    float px = bb.pos[0], py = bb.pos[1], pz = bb.pos[2];
    float xw = (float)(px + bb.ext[0]), yh = (float)(py + bb.ext[1]), zd = (float)(pz + bb.ext[2]);
    // Vertices and triangles of FluidSimulation::_getTriangleMeshFromAABB.
    const float vert[8][3] = {{px, py, pz}, {xw, py, pz}, {xw, py, zd}, {px, py, zd},
                              {px, yh, pz}, {xw, yh, pz}, {xw, yh, zd}, {px, yh, zd}};
    const int tri[12][3] = {{0, 1, 2}, {0, 2, 3}, {4, 7, 6}, {4, 6, 5}, {0, 3, 7}, {0, 7, 4},
                            {1, 5, 6}, {1, 6, 2}, {0, 4, 5}, {0, 5, 1}, {3, 2, 6}, {3, 6, 7}};
    BoxMesh m{};
    m.ntri = 0;
    for (int t = 0; t < 12; t++) {
        BoxTri T{};
        double mn[3], mx[3];
        for (int a = 0; a < 3; a++) {
            mn[a] = mx[a] = vert[tri[t][0]][a];
        }
        for (int c = 0; c < 3; c++) {
            for (int a = 0; a < 3; a++) {
                T.v[c][a] = vert[tri[t][c]][a];
                mn[a] = std::fmin((double)vert[tri[t][c]][a], mn[a]);
                mx[a] = std::fmax((double)vert[tri[t][c]][a], mx[a]);
            }
        }
        // AABB(Triangle) min corner in float, extents + 1e-9, max point in float.
        float bmin[3], bmax[3];
        for (int a = 0; a < 3; a++) {
            bmin[a] = (float)mn[a];
            double ext = mx[a] - mn[a] + 1e-9;
            bmax[a] = bmin[a] + (float)ext;
        }
        int gmin[3], gmax[3], lim[3] = {isize, jsize, ksize};
        for (int a = 0; a < 3; a++) {
            gmin[a] = floor_index((double)bmin[a], dx) - band;
            gmax[a] = floor_index((double)bmax[a], dx) + band + 1;
        }
        bool outside = false;
        for (int a = 0; a < 3; a++)
            if (gmax[a] < 0 || gmin[a] >= lim[a] + 1) outside = true;
        if (outside) continue;
        for (int a = 0; a < 3; a++) {
            T.gmin[a] = std::max(gmin[a], 0);
            T.gmax[a] = std::min(gmax[a], lim[a]);
        }
        m.tri[m.ntri++] = T;
    }
    for (int a = 0; a < 3; a++) {
        m.bpos[a] = bb.pos[a];
        m.bext[a] = bb.ext[a];
    }
    return m;
}

void collision_box(int isize, int jsize, int ksize, double dx, float pos[3], double ext[3]) {
    HBox b = boundary_box(isize, jsize, ksize, dx);
    double v = -0.1f * dx;  // fluidsimulation.h:1685 (_solidBufferWidth)
    double h = 0.5 * v;
    for (int a = 0; a < 3; a++) {
        pos[a] = b.pos[a] - (float)h;
        ext[a] = b.ext[a] + v;
    }
}

// ---------------------------------------------------------------------------
// Pressure solve driver
// ---------------------------------------------------------------------------

size_t PressureStore::device_bytes() const {
    size_t b = edge.bytes() + mask.bytes() + progress.bytes() + ticket.bytes() + irange.bytes() + scal.bytes() +
               alist.bytes();
    for (int q = 0; q <= kSlotArrays; q++) b += arr[q].bytes();
    for (int q = 0; q < kWeightArrays; q++) b += wgt[q].bytes();
    for (int q = 0; q < kSlotArrays32; q++) b += arr32[q].bytes();
    return b;
}

void PressureStore::alloc(const GridGeom &g, hipStream_t s, bool mixed) {
    graphs.clear();  // captured launches point into the buffers replaced here
    SlotMap sm = make_slot_map(g.I, g.J, g.K);
    // Slot arrays carry the zero and sink groups after the last tile.
    size_t ns = sm.alloc_count();
    // The kernels index slots in 32-bit arithmetic.
    if (ns > 0xffffffffull)
        throw StatusError{FLIPMI_ERR_INVALID_ARG, "pressure grid too large: more than 2^32 slots"};
    int ntj = sm.ntj, ntk = sm.ntk;
    int npart = 2 * 2048 + ntj * ntk;  // the assembly's 2 x 2048 block shares; the PCG's per-block and per-tile partials
    for (int q = 0; q < kSlotArrays; q++)
        if (arr[q].alloc(ns) != hipSuccess)
            throw StatusError{FLIPMI_ERR_OOM, "pressure arrays: out of device memory"};
    for (int q = 0; q < kWeightArrays; q++)
        if (wgt[q].alloc(ns) != hipSuccess)
            throw StatusError{FLIPMI_ERR_OOM, "pressure arrays: out of device memory"};
    size_t nrec = edge_record_count(sm);
    if (arr[kSlotArrays].alloc((size_t)npart) != hipSuccess || edge.alloc(nrec) != hipSuccess ||
        progress.alloc((size_t)ntj * ntk) != hipSuccess ||
        ticket.alloc(1) != hipSuccess || scal.alloc(1) != hipSuccess ||
        irange.alloc((size_t)2 * ntj * ntk) != hipSuccess)
        throw StatusError{FLIPMI_ERR_OOM, "pressure scratch: out of device memory"};
    (void)launch_fill_i32((int *)progress.p, progress.n, 0, s);
    (void)launch_fill_i32((int *)ticket.p, 1, 0, s);
    (void)launch_fill_u8((uint8_t *)scal.p, sizeof(PcgScalars), 0, s);
    (void)launch_fill_i32((int *)&scal.p->sweep_epoch, 1, 1, s);
    // Slots outside the grid are never written by the kernels: start at zero.
    for (int q = 0; q < kSlotArrays; q++) (void)launch_fill_f64(arr[q].p, ns, 0.0, s);
    for (int q = 0; q < kWeightArrays; q++) (void)launch_fill_f32(wgt[q].p, ns, -0.0f, s);  // no coupling
    (void)launch_fill_u8((uint8_t *)edge.p, nrec * sizeof(unsigned long long), 0, s);
    pb.sm = sm;
    pb.diag = arr[0].p;
    pb.rhs = arr[1].p; pb.x = arr[2].p; pb.r = arr[3].p; pb.z = arr[4].p; pb.s = arr[5].p;
    pb.invdiag = arr[6].p; pb.Lx = arr[7].p; pb.Ly = arr[8].p; pb.Lz = arr[9].p;
    pb.Lxm = arr[10].p; pb.Lym = arr[11].p; pb.Lzm = arr[12].p;
    pb.wx = wgt[0].p; pb.wy = wgt[1].p; pb.wz = wgt[2].p;
    pb.partials = arr[kSlotArrays].p;
    if (mixed) {
        for (int q = 0; q < kSlotArrays32; q++) {
            if (arr32[q].alloc(ns) != hipSuccess)
                throw StatusError{FLIPMI_ERR_OOM, "mixed-precision arrays: out of device memory"};
            (void)launch_fill_f32(arr32[q].p, ns, 0.0f, s);
        }
        float **dst[kSlotArrays32] = {&pb.invdiag32, &pb.Lx32, &pb.Ly32, &pb.Lz32, &pb.Lxm32, &pb.Lym32,
                                      &pb.Lzm32, &pb.r32, &pb.z32, &pb.s32, &pb.s32b, &pb.q32};
        for (int q = 0; q < kSlotArrays32; q++) *dst[q] = arr32[q].p;
    }
    pb.npartials = npart;
    pb.edge = edge.p;
    const int groups_per_tile = sm.S / 4;
    pb.mask_words = (groups_per_tile + 63) / 64;
    if (mask.alloc((size_t)ntj * ntk * pb.mask_words) != hipSuccess)
        throw StatusError{FLIPMI_ERR_OOM, "pressure scratch: out of device memory"};
    // Every group counts as active until the first assembly marks them.
    (void)launch_fill_u8((uint8_t *)mask.p, mask.n * sizeof(unsigned long long), 0xff, s);
    pb.group_mask = mask.p;
    const size_t ngroups = sm.count() >> 8;
    if (alist.alloc(ngroups + 1) != hipSuccess)
        throw StatusError{FLIPMI_ERR_OOM, "pressure scratch: out of device memory"};
    (void)launch_fill_i32(alist.p, ngroups + 1, 0, s);
    pb.active_list = alist.p;
    pb.active_count = alist.p + ngroups;
    pb.epoch = 0;
    pb.progress = progress.p;
    pb.ntiles_j = ntj;
    pb.ntiles_k = ntk;
    pb.ticket = ticket.p;
    pb.tile_irange = irange.p;
    pb.sc = scal.p;
}

static void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw StatusError{FLIPMI_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

// One batch of PCG iterations as a graph: captured the first time this
// (batch size, starting s buffer) occurs, replayed afterwards.  The kernels
// read their scalars (alpha, beta, the sweep epoch, done) from device memory,
// so a replay computes exactly what launching the kernels would; only the
// host's bookkeeping of the launches (the s buffer swap of each SpMV, the
// count of sweep launches) is advanced by hand on a replay.
static void launch_pcg_batch_graph(PcgGraphs &G, const GridGeom &g, bool mixed, PressureBuffers &pb, int batch,
                                   hipStream_t s) {
    const auto key = std::make_pair(batch, (const void *)pb.s32);
    auto it = G.exec.find(key);
    if (it == G.exec.end()) {
        hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "pcg graph capture");
        hipError_t le = hipSuccess;
        for (int q = 0; q < batch && le == hipSuccess; q++) le = launch_pcg_iteration(g, mixed, pb, s);
        hipGraph_t graph = nullptr;
        hipError_t ce = hipStreamEndCapture(s, &graph);
        hip_check(le, "pcg iteration (capture)");
        hip_check(ce, "pcg graph capture");
        hipGraphExec_t exec = nullptr;
        hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        hip_check(ie, "pcg graph instantiate");
        it = G.exec.emplace(key, exec).first;
    } else {
        if (mixed && (batch & 1)) std::swap(pb.s32, pb.s32b);
        pb.epoch += pcg_sweeps_per_iteration(mixed) * (unsigned)batch;
    }
    hip_check(hipGraphLaunch(it->second, s), "pcg graph launch");
}

PressureOutcome run_pressure(const PressureWork &w, PressureBuffers &pb, PcgScalars *hs,
                             hipStream_t s, Scene *prof) {
    PressureOutcome out;
    hipEvent_t ev = nullptr;
    if (prof) prof->begin_stage(FLIPMI_STAGE_ASSEMBLE, &ev);
    hip_check(launch_pressure_reset(pb, s), "pressure reset");
    hip_check(launch_pressure_assemble(w.g, w.dt, w.liquid, w.wU, w.wV, w.wW, w.wC, w.u, w.v, w.w,
                                       w.sU, w.sV, w.sW, pb, s),
              "pressure assemble");
    if (prof) prof->end_stage(FLIPMI_STAGE_ASSEMBLE, ev);
    hip_check(hipMemcpyAsync(hs, pb.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s), "scalars");
    hip_check(hipStreamSynchronize(s), "assemble sync");
    double bnorm;
    std::memcpy(&bnorm, &hs->absmax_bits, sizeof(double));
    out.rows = hs->rows;
    // Early exit of PressureSolver::solve when max|rhs| < tolerance
    // (pressuresolver.cpp:52-62): nothing is solved or applied.
    if (bnorm < w.tol) {
        out.iterations = 0;
        out.residual = 0.0;
        out.success = true;
        return out;
    }
    if (prof) prof->begin_stage(FLIPMI_STAGE_MIC0, &ev);
    const bool mixed = w.precision == FLIPMI_PCG_MIXED;
    if (mixed && !pb.r32) throw StatusError{FLIPMI_ERR_INVALID_ARG, "mixed precision arrays not allocated"};
    hip_check(launch_group_mask(pb, s), "group mask");
    hip_check(launch_mic0_factor(w.g, w.liquid, pb, s), "mic0 factor");
    hip_check(launch_pcg_init(w.g, w.tol, w.maxiter, mixed, pb, s), "pcg init");
    if (prof) prof->end_stage(FLIPMI_STAGE_MIC0, ev);
    if (prof) prof->begin_stage(FLIPMI_STAGE_PCG, &ev);
    // Iterations are enqueued in batches between host checks of the device
    // scalars.  After the first batch, the batch size is the number of
    // iterations the observed convergence rate still needs (plus a margin),
    // so few launches are wasted past convergence.
    int batch = 8;
    int launched = 0;
    int prev_iters = 0;
    double prev_resid = bnorm;
    const double target = w.tol * bnorm;
    for (;;) {
        // A batch runs from a captured graph once the first SpMV (which reads
        // s as initialised) is behind, unless kernels are being timed or the
        // sweep tag is near its wrap-around (which clears the records).
        const bool use_graph = w.graphs && pb.s_plain == 0 && !(prof && prof->profiling >= 2) &&
                               pb.epoch + pcg_sweeps_per_iteration(mixed) * (unsigned)batch + 8u < 0xfffffff0u;
        for (int q = 0; q < batch; q++) {
            const bool timed = prof && prof->profiling >= 2 && prof->kernel_mask != 0 &&
                               prof->pcg_iteration_count++ % prof->kernel_every == 0;
            if (timed) {
                // Time the selected kernels of the iteration on the launch stream.
                for (int q = 0; q < kPcgPieces; q++) {
                    int piece = pcg_piece_at(q);
                    if (!pcg_piece_used(piece, mixed)) continue;
                    int stage = piece == kPcgResidual ? FLIPMI_STAGE_K_RESIDUAL : FLIPMI_STAGE_K_SPMV + piece;
                    KernelTiming kt;
                    if ((prof->kernel_mask >> piece) & 1u) {
                        kt.start = prof->take_event();
                        kt.stop = prof->take_event();
                    }
                    hip_check(launch_pcg_piece(piece, w.g, mixed, pb, s, kt), "pcg kernel");
                    if (kt.start) prof->add_timed(stage, kt.start, kt.stop);
                }
            } else if (use_graph) {
                // (Never at profiling level 2: there every iteration runs on
                // its own, timed or not.)
                launch_pcg_batch_graph(*w.graphs, w.g, mixed, pb, batch, s);
                break;  // the graph holds the whole batch
            } else {
                hip_check(launch_pcg_iteration(w.g, mixed, pb, s), "pcg iteration");
            }
        }
        launched += batch;
        hip_check(hipMemcpyAsync(hs, pb.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s), "scalars");
        hip_check(hipStreamSynchronize(s), "pcg sync");
        if (hs->abort) throw StatusError{FLIPMI_ERR_TIMEOUT, "wavefront sweep wait timed out"};
        if (hs->done) break;
        if (launched > w.maxiter + 16)
            throw StatusError{FLIPMI_ERR_HIP, "pcg did not terminate within its iteration cap"};
        int next = std::min(batch * 2, 32);
        const int done_iters = hs->iters - prev_iters;
        if (done_iters > 0 && hs->resid > 0 && hs->resid < prev_resid && target > 0) {
            const double rate = std::log(hs->resid / prev_resid) / done_iters;  // < 0
            const double need = std::log(target / hs->resid) / rate;
            if (need > 0 && need < 1e6) next = (int)std::ceil(need) + 2;
        }
        prev_iters = hs->iters;
        prev_resid = hs->resid;
        batch = std::max(2, std::min(next, 64));
    }
    if (prof) prof->end_stage(FLIPMI_STAGE_PCG, ev);
    if (prof) prof->stage_calls[FLIPMI_STAGE_PCG_ITERS] += hs->iters;
    out.iterations = hs->failed ? 0 : hs->iters;
    out.residual = hs->resid;
    out.refinements = mixed ? hs->refinements : 0;
    bool converged = hs->converged != 0;
    // PressureSolver::_solveLinearSystem (pressuresolver.cpp:810-840).
    out.success = converged || (!hs->failed && out.iterations == w.maxiter && out.residual < w.acceptable);
    if (out.success) {
        if (prof) prof->begin_stage(FLIPMI_STAGE_APPLY, &ev);
        hip_check(launch_pressure_apply(w.g, pb.sm, w.dt, w.liquid, w.wU, w.wV, w.wW, pb.x, w.u, w.v, w.w,
                                        w.vu, w.vv, w.vw, s),
                  "pressure apply");
        if (prof) prof->end_stage(FLIPMI_STAGE_APPLY, ev);
        out.solved = true;
    }
    return out;
}

// ---------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------

Scene::~Scene() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto &pe : pending_events) {
        (void)hipEventDestroy(pe.a);
        (void)hipEventDestroy(pe.b);
    }
    for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
    if (host_stage) (void)hipHostFree(host_stage);
    delete press;
    delete mesh;
    if (own_stream && stream) (void)hipStreamDestroy(stream);
}

void Scene::check(hipError_t e, const char *what) { hip_check(e, what); }

void Scene::sync() { check(hipStreamSynchronize(stream), "stream sync"); }

hipEvent_t Scene::take_event() {
    if (!event_pool.empty()) {
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    check(hipEventCreate(&e), "event create");
    return e;
}

void Scene::begin_stage(int stage, hipEvent_t *a) {
    (void)stage;
    if (!profiling) return;
    *a = take_event();
    check(hipEventRecord(*a, stream), "event record");
}

void Scene::end_stage(int stage, hipEvent_t a) {
    if (!profiling) return;
    hipEvent_t b = take_event();
    check(hipEventRecord(b, stream), "event record");
    pending_events.push_back(PendingEvent{stage, a, b});
}

void Scene::add_timed(int stage, hipEvent_t a, hipEvent_t b) {
    pending_events.push_back(PendingEvent{stage, a, b});
}

void Scene::collect_events() {
    if (pending_events.empty()) return;
    sync();
    for (auto &pe : pending_events) {
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, pe.a, pe.b), "event elapsed");
        stage_ms[pe.stage] += ms;
        stage_calls[pe.stage] += 1;
        event_pool.push_back(pe.a);
        event_pool.push_back(pe.b);
    }
    pending_events.clear();
}

// Scratch of the prefix scans: the half-cell bins or the particles.
static size_t scan_elems(const GridGeom &g, size_t particles) {
    return std::max(scan_scratch_elems((int)subcell_bins(g)), scan_scratch_elems((int)particles));
}

void Scene::alloc_all() {
    int I = desc.isize, J = desc.jsize, K = desc.ksize;
    size_t nc = (size_t)I * J * K;
    size_t nn = (size_t)(I + 1) * (J + 1) * (K + 1);
    size_t nU = (size_t)(I + 1) * J * K, nV = (size_t)I * (J + 1) * K, nW = (size_t)I * J * (K + 1);
    size_t nf = std::max(nU, std::max(nV, nW));
    auto need = [&](hipError_t e) {
        if (e != hipSuccess) throw StatusError{FLIPMI_ERR_OOM, "scene allocation: out of device memory"};
    };
    size_t pc = (size_t)std::max(cap, 1);
    for (int a = 0; a < 6; a++) {
        need(part[a].alloc(pc));
        need(part_alt[a].alloc(pc));
    }
    need(spos.alloc(pc));
    for (int a = 0; a < 3; a++) need(svel[a].alloc(pc));
    need(cell.alloc(pc)); need(slot.alloc(pc)); need(order.alloc(pc));
    need(keep.alloc(pc)); need(keep_pos.alloc(pc)); need(rank.alloc(pc));
    need(solid_removed.alloc(pc));
    need(count.alloc(nc)); need(offset.alloc(nc + 1)); need(removed_in_cell.alloc(nc));
    const size_t nsub = subcell_bins(g);
    if (nsub >= (size_t)INT32_MAX) throw StatusError{FLIPMI_ERR_INVALID_ARG, "grid too large for the half-cell sort"};
    need(sub_count.alloc(nsub)); need(sub_offset.alloc(nsub + 1));
    need(p2g_ovf.alloc(p2g_face_count(g))); need(p2g_ovf_count.alloc(1));
    need(scan_tmp.alloc(scan_elems(g, pc)));
    need(new_count.alloc(1)); need(over_cap.alloc(1)); need(speed_hist.alloc(64)); need(max_speed_bits.alloc(1));
    need(solid.alloc(nn));
    need(wU.alloc(nU)); need(wV.alloc(nV)); need(wW.alloc(nW)); need(wC.alloc(nc));
    need(sU.alloc(nU)); need(sV.alloc(nV)); need(sW.alloc(nW));
    nci = (int)std::ceil((double)I * g.dx / (3 * g.dx));
    ncj = (int)std::ceil((double)J * g.dx / (3 * g.dx));
    nck = (int)std::ceil((double)K * g.dx / (3 * g.dx));
    need(near_a.alloc((size_t)nci * ncj * nck)); need(near_b.alloc((size_t)nci * ncj * nck));
    need(liquid.alloc(nc));
    need(u.alloc(nU)); need(v.alloc(nV)); need(w.alloc(nW));
    need(su.alloc(nU)); need(sv.alloc(nV)); need(sw.alloc(nW));
    need(vu.alloc(nU)); need(vv.alloc(nV)); need(vw.alloc(nW));
    need(status.alloc(nf));
    press = new PressureStore();
    press->alloc(g, stream, desc.pcg_precision == FLIPMI_PCG_MIXED);
    static_assert(sizeof(PcgScalars) <= 4096, "the PCG scalars are read back through the staging buffer");
    check(hipHostMalloc(&host_stage, 4096, hipHostMallocDefault), "pinned staging");
    // Fresh MeshLevelSet: phi = distance upper bound everywhere
    // (meshlevelset.cpp:35-42, 760-762); zero solid face velocities.
    float upper = (float)((double)((I + 1) + (J + 1) + (K + 1)) * g.dx);
    check(launch_fill_f32(solid.p, nn, upper, stream), "fill");
    check(launch_fill_f32(sU.p, nU, 0.0f, stream), "fill");
    check(launch_fill_f32(sV.p, nV, 0.0f, stream), "fill");
    check(launch_fill_f32(sW.p, nW, 0.0f, stream), "fill");
}

// FluidSimulation::loadMarkerParticleData (fluidsimulation.cpp:2488-2509):
// the particles queue up as given; initialize() admits them.
void Scene::load(const float *pos, const float *vel, int64_t count_in) {
    pending_pos.insert(pending_pos.end(), pos, pos + 3 * (size_t)count_in);
    if (vel) pending_vel.insert(pending_vel.end(), vel, vel + 3 * (size_t)count_in);
    else pending_vel.insert(pending_vel.end(), 3 * (size_t)count_in, 0.0f);
}

void Scene::admit_pending() {
    // FluidSimulation::_loadMarkerParticles (fluidsimulation.cpp:2773-2785):
    // p = (p - offset) / scale in vmath's float arithmetic (the division is
    // a multiplication by (float)(1.0 / (float)scale), vmath.cpp:100-103),
    // then keep the points inside [0, I dx) x [0, J dx) x [0, K dx)
    // (AABB::isPointInside, aabb.cpp:130-133).
    const float inv = (float)(1.0 / (double)(float)domain_scale);
    const float bx = 0.0f, by = 0.0f, bz = 0.0f;
    const double ex = desc.isize * desc.dx, ey = desc.jsize * desc.dx, ez = desc.ksize * desc.dx;
    const size_t m = pending_pos.size() / 3;
    size_t kept = 0;
    for (size_t q = 0; q < m; q++) {
        const float x = (pending_pos[3 * q] - domain_offset[0]) * inv;
        const float y = (pending_pos[3 * q + 1] - domain_offset[1]) * inv;
        const float z = (pending_pos[3 * q + 2] - domain_offset[2]) * inv;
        if (!(x >= bx && y >= by && z >= bz && (double)x < (double)bx + ex &&
              (double)y < (double)by + ey && (double)z < (double)bz + ez))
            continue;
        pending_pos[3 * kept] = x; pending_pos[3 * kept + 1] = y; pending_pos[3 * kept + 2] = z;
        for (int a = 0; a < 3; a++) pending_vel[3 * kept + a] = pending_vel[3 * q + a];
        kept++;
    }
    pending_pos.resize(3 * kept);
    pending_vel.resize(3 * kept);
}

// getMarkerParticlePositionData (fluidsimulation.cpp:2408-2413):
// position * scale + offset, in float.
void Scene::to_domain(float *pos_xyz, int64_t count) const {
    const float sc = (float)domain_scale;
    for (int64_t q = 0; q < 3 * count; q++) pos_xyz[q] = pos_xyz[q] * sc + domain_offset[q % 3];
}

void Scene::upload_pending() {
    n = (int)(pending_pos.size() / 3);
    if (n > 0) {
        // Host AoS -> device SoA.
        std::vector<float> tmp((size_t)n);
        for (int a = 0; a < 6; a++) {
            const std::vector<float> &src = a < 3 ? pending_pos : pending_vel;
            for (int q = 0; q < n; q++) tmp[q] = src[3 * (size_t)q + (a % 3)];
            check(hipMemcpyAsync(part[a].p, tmp.data(), sizeof(float) * n, hipMemcpyHostToDevice, stream),
                  "upload particles");
            sync();
        }
    }
    pending_pos.clear(); pending_vel.clear();
    pending_pos.shrink_to_fit(); pending_vel.shrink_to_fit();
}

void Scene::initialize() {
    if (initialized) return;
    admit_pending();
    int64_t np = (int64_t)(pending_pos.size() / 3);
    if (np > INT32_MAX / 2) throw StatusError{FLIPMI_ERR_INVALID_ARG, "too many particles"};
    cap = (int)std::max<int64_t>(np, desc.max_particles);
    alloc_all();
    upload_pending();
    // FluidSimulation::_initializeParticleRadii (fluidsimulation.cpp:2644-2650).
    radius = 0.5 * 1.0 * g.dx * std::sqrt(3.0);
    initialized = true;
}

bool Scene::rearm(const float *pos, const float *vel, int64_t count_in) {
    removal_counts_ready = false;
    if (mesh) {
        mesh->valid = false;
        mesh->nverts = mesh->ntris = 0;
    }
    if (!initialized || !pending_events.empty()) return false;
    pending_pos.clear();
    pending_vel.clear();
    load(pos, vel, count_in);
    admit_pending();
    if ((int64_t)(pending_pos.size() / 3) > cap) {
        pending_pos.clear();
        pending_vel.clear();
        return false;
    }
    upload_pending();
    // The state FluidSimulation has right after initialize(); the static
    // solid, weights and near-solid grid depend only on the grid and stay.
    current_frame = 0;
    step_number = 0;
    frame_dt = remaining = cur_step = 0.0;
    output_initialized = false;
    last_iters = 0;
    last_ok = 1;
    last_resid = 0.0;
    last_rows = 0;
    last_refinements = 0;
    for (int q = 0; q < FLIPMI_NUM_STAGES; q++) {
        stage_ms[q] = 0.0;
        stage_calls[q] = 0;
    }
    return true;
}

size_t Scene::device_bytes() const {
    size_t b = 0;
    for (int a = 0; a < 6; a++) b += part[a].bytes() + part_alt[a].bytes();
    b += spos.bytes() + svel[0].bytes() + svel[1].bytes() + svel[2].bytes();
    b += cell.bytes() + slot.bytes() + count.bytes() + offset.bytes() + order.bytes() + scan_tmp.bytes();
    b += removed_in_cell.bytes() + sub_count.bytes() + sub_offset.bytes() + p2g_ovf.bytes() + p2g_ovf_count.bytes();
    b += keep.bytes() + keep_pos.bytes() + rank.bytes() + new_count.bytes() + over_cap.bytes() + solid_removed.bytes();
    b += speed_hist.bytes() + max_speed_bits.bytes();
    b += solid.bytes() + wU.bytes() + wV.bytes() + wW.bytes() + wC.bytes() + sU.bytes() + sV.bytes() + sW.bytes();
    b += near_a.bytes() + near_b.bytes() + liquid.bytes() + u.bytes() + v.bytes() + w.bytes();
    b += su.bytes() + sv.bytes() + sw.bytes() + vu.bytes() + vv.bytes() + vw.bytes() + status.bytes();
    if (press) b += press->device_bytes();
    if (mesh) b += mesh->device_bytes();
    for (auto &src : sources) b += src->sdf.bytes() + src->cells.bytes() + src->list.bytes();
    b += inflow_mask.bytes() + emit_cand.bytes() + emit_flags.bytes() + emit_pos.bytes() + emit_scan.bytes() +
         emit_nbase.bytes();
    return b;
}

size_t estimate_scene_bytes(const flipmi_scene_desc &d, int64_t cap) {
    // The allocations of Scene::alloc_all and PressureStore::alloc.
    const size_t I = (size_t)d.isize, J = (size_t)d.jsize, K = (size_t)d.ksize;
    const size_t nc = I * J * K, nn = (I + 1) * (J + 1) * (K + 1);
    const size_t nU = (I + 1) * J * K, nV = I * (J + 1) * K, nW = I * J * (K + 1);
    const size_t nf = std::max(nU, std::max(nV, nW));
    const size_t pc = (size_t)std::max<int64_t>(cap, 1);
    size_t b = pc * (19 * sizeof(float) + 6 * sizeof(int) + 1);  // part, part_alt, spos, svel
    b += sizeof(int) * (nc + nc + 1 + nc);  // count, offset, removed_in_cell
    const GridGeom g = make_geom(d.isize, d.jsize, d.ksize, d.dx);
    b += sizeof(int) * (2 * subcell_bins(g) + 1 + p2g_face_count(g) + 1);  // half-cell sort, P2G overflow
    b += sizeof(int) * scan_elems(g, pc);
    b += sizeof(int) * (1 + 64) + sizeof(unsigned);
    b += sizeof(float) * nn;
    b += sizeof(float) * 2 * (nU + nV + nW) + sizeof(float) * nc;  // weights (faces, centre), solid face velocities
    const double dx = d.dx;
    const size_t nci = (size_t)std::ceil((double)I * dx / (3 * dx)), ncj = (size_t)std::ceil((double)J * dx / (3 * dx)),
                 nck = (size_t)std::ceil((double)K * dx / (3 * dx));
    b += 2 * nci * ncj * nck;
    b += sizeof(float) * nc + sizeof(float) * 2 * (nU + nV + nW) + (nU + nV + nW) + nf;
    SlotMap sm = make_slot_map(d.isize, d.jsize, d.ksize);
    const size_t ns = sm.alloc_count(), ntiles = (size_t)sm.ntj * sm.ntk;
    b += sizeof(double) * (PressureStore::kSlotArrays * ns + 2 * 2048 + ntiles);
    b += sizeof(float) * PressureStore::kWeightArrays * ns;
    if (d.pcg_precision == FLIPMI_PCG_MIXED) b += sizeof(float) * PressureStore::kSlotArrays32 * ns;
    b += sizeof(unsigned long long) * edge_record_count(sm);
    b += sizeof(unsigned long long) * ntiles * (size_t)((sm.S / 4 + 63) / 64);
    b += sizeof(unsigned) * (ntiles + 1) + sizeof(int) * 2 * ntiles + sizeof(PcgScalars);
    b += sizeof(int) * ((sm.count() >> 8) + 1);
    return b;
}

GridGeom make_geom(int I, int J, int K, double dx) {
    if (I < 1 || J < 1 || K < 1 || !(dx > 0.0))
        throw StatusError{FLIPMI_ERR_INVALID_ARG, "dimensions and cell size must be greater than 0"};
    if ((double)I * J * K > 1.5e9) throw StatusError{FLIPMI_ERR_INVALID_ARG, "grid too large"};
    // Kernels index grids (nodes included) in 32-bit arithmetic.
    if ((double)(I + 1) * (J + 1) * (K + 1) >= 4294967296.0)
        throw StatusError{FLIPMI_ERR_INVALID_ARG, "grid too large"};
    GridGeom g;
    g.I = I; g.J = J; g.K = K; g.dx = dx; g.invdx = 1.0 / dx;
    return g;
}

void validate_scene_desc(const flipmi_scene_desc &d) {
    (void)make_geom(d.isize, d.jsize, d.ksize, d.dx);
    if (d.min_substeps < 1 || d.max_substeps < d.min_substeps || d.max_substeps > 64)
        throw StatusError{FLIPMI_ERR_INVALID_ARG, "substep bounds must satisfy 1 <= min <= max <= 64"};
    if (!(d.cfl > 0)) throw StatusError{FLIPMI_ERR_INVALID_ARG, "CFL number must be positive"};
    if (d.pcg_max_iterations < 1) throw StatusError{FLIPMI_ERR_INVALID_ARG, "pcg_max_iterations must be positive"};
    if (d.pcg_precision != FLIPMI_PCG_FP64 && d.pcg_precision != FLIPMI_PCG_MIXED)
        throw StatusError{FLIPMI_ERR_INVALID_ARG, "unknown pcg precision"};
}

void Scene::ensure_solid() {
    // FluidSimulation::_updateSolidLevelSet (fluidsimulation.cpp:3028-3072):
    // computed once, when marker particles or mesh fluid sources exist.
    if (solid_ready || (n == 0 && sources.empty())) return;
    BoxMesh mesh = make_domain_box(desc.isize, desc.jsize, desc.ksize, desc.dx);
    check(launch_domain_solid(g, mesh, solid.p, stream), "domain solid");
    union_obstacles();
    weights_ready = false;
    // _updateNearSolidGrid (fluidsimulation.cpp:3094-3126) depends only on
    // the static solid, so it is built once as well.
    int layers = (int)std::ceil((float)desc.cfl / (float)3);
    check(launch_near_solid(g, solid.p, near_a.p, near_b.p, nci, ncj, nck, layers, &near, stream),
          "near solid");
    solid_ready = true;
}

void Scene::sort_particles(int *over) {
    SortBuffers sb{cell.p, slot.p, count.p, offset.p, order.p, scan_tmp.p, scan_tmp.n, over};
    check(launch_sort_by_cell(g, part[0].p, part[1].p, part[2].p, n, sb, stream), "sort");
}

void Scene::sort_subcells() {
    SortBuffers sb{cell.p, slot.p, sub_count.p, sub_offset.p, order.p, scan_tmp.p, scan_tmp.n};
    check(launch_sort_by_subcell(g, part[0].p, part[1].p, part[2].p, n, sb, stream), "half-cell sort");
}

void Scene::build_mesh() {
    hipEvent_t ev = nullptr;
    begin_stage(FLIPMI_STAGE_SURFACE, &ev);
    if (!mesh) mesh = new SurfaceStore();
    if (n > 0) {
        sort_subcells();
        const float *src[6];
        for (int a = 0; a < 6; a++) src[a] = part[a].p;
        float *sorted_vel[3] = {svel[0].p, svel[1].p, svel[2].p};
        check(launch_gather_sorted(order.p, sub_offset.p + subcell_bins(g), n, src, spos.p, sorted_vel, stream),
              "gather");
    }
    SurfaceInputs in;
    in.g = g;
    in.spos = spos.p;
    in.sub_offset = sub_offset.p;
    in.n = n;
    in.solid = solid.p;
    in.desc = surface;
    in.kernel_radius = marker_particle_radius(g.dx) * surface.particle_scale;
    for (int a = 0; a < 3; a++) in.domain_offset[a] = domain_offset[a];
    in.domain_scale = domain_scale;
    build_surface(in, *mesh, stream);
    end_stage(FLIPMI_STAGE_SURFACE, ev);
}

void Scene::extrapolate_all(int layers) {
    hipEvent_t ev = nullptr;
    begin_stage(FLIPMI_STAGE_EXTRAPOLATE, &ev);
    check(launch_extrapolate(g.I + 1, g.J, g.K, u.p, vu.p, status.p, layers, stream), "extrapolate");
    check(launch_extrapolate(g.I, g.J + 1, g.K, v.p, vv.p, status.p, layers, stream), "extrapolate");
    check(launch_extrapolate(g.I, g.J, g.K + 1, w.p, vw.p, status.p, layers, stream), "extrapolate");
    end_stage(FLIPMI_STAGE_EXTRAPOLATE, ev);
}

void Scene::pressure_solve(double dt) {
    if (!weights_ready) {
        check(launch_weights(g, solid.p, wU.p, wV.p, wW.p, wC.p, stream), "weights");
        weights_ready = true;
    }
    PressureWork pw;
    pw.g = g; pw.dt = dt; pw.tol = desc.pcg_tolerance; pw.acceptable = 1.0;
    pw.maxiter = desc.pcg_max_iterations; pw.precision = desc.pcg_precision;
    pw.liquid = liquid.p; pw.wU = wU.p; pw.wV = wV.p; pw.wW = wW.p; pw.wC = wC.p;
    pw.sU = sU.p; pw.sV = sV.p; pw.sW = sW.p;
    pw.u = u.p; pw.v = v.p; pw.w = w.p; pw.vu = vu.p; pw.vv = vv.p; pw.vw = vw.p;
    pw.graphs = launch_graphs ? &press->graphs : nullptr;
    PressureOutcome o =
        run_pressure(pw, press->pb, (PcgScalars *)host_stage, stream, profiling ? this : nullptr);
    last_iters = o.iterations;
    last_resid = o.residual;
    last_ok = o.success ? 1 : 0;
    last_rows = o.rows;
    last_refinements = o.refinements;
}

void Scene::advance_particles(double dt) {
    hipEvent_t ev = nullptr;
    begin_stage(FLIPMI_STAGE_G2P_ADVECT, &ev);
    ParticleStepArgs a{};
    a.px = part[0].p; a.py = part[1].p; a.pz = part[2].p;
    a.vx = part[3].p; a.vy = part[4].p; a.vz = part[5].p;
    a.n = n;
    a.u = u.p; a.v = v.p; a.w = w.p;
    a.su = su.p; a.sv = sv.p; a.sw = sw.p;
    a.solid = solid.p;
    a.near_solid = near;
    a.nci = nci; a.ncj = ncj; a.nck = nck;
    a.near_cell = 3 * g.dx;
    a.pic_ratio_f = (float)desc.pic_ratio;
    a.flip_ratio_f = (float)(1 - desc.pic_ratio);
    a.dt = dt;
    a.cfl = desc.cfl;
    collision_box(desc.isize, desc.jsize, desc.ksize, desc.dx, a.bpos, a.bext);
    a.solid_removed = solid_removed.p;
    a.speed_hist = speed_hist.p;
    // Speed limit uses the frame step (fluidsimulation.cpp:4385).
    a.speed_step = desc.cfl * g.dx / frame_dt;
    a.maxsteps = desc.max_substeps;
    a.inflow = inflow_views(false);
    check(launch_fill_i32(speed_hist.p, 64, 0, stream), "fill");
    check(launch_particle_step(g, a, stream), "particle step");
    end_stage(FLIPMI_STAGE_G2P_ADVECT, ev);

    // _removeMarkerParticles (fluidsimulation.cpp:4324-4353).
    begin_stage(FLIPMI_STAGE_REMOVE, &ev);
    int *hist = (int *)host_stage;
    check(hipMemcpyAsync(hist, speed_hist.p, sizeof(int) * desc.max_substeps, hipMemcpyDeviceToHost, stream),
          "hist");
    // The per-cell order is built only when some cell is over the cap.
    sort_particles(over_cap.p);
    check(launch_removed_in_cell(g, part[0].p, part[1].p, part[2].p, solid_removed.p, n, removed_in_cell.p,
                                 stream),
          "removed in cell");
    removal_counts_ready = true;
    check(launch_fill_i32(rank.p, (size_t)std::max(n, 1), -1, stream), "fill");
    check(launch_cell_rank_cells(offset.p, g.I * g.J * g.K, order.p, solid_removed.p, over_cap.p, rank.p,
                                 stream),
          "cell rank");
    sync();
    // _getMarkerParticleSpeedLimit (fluidsimulation.cpp:4298-4322).
    double step = desc.cfl * g.dx / frame_dt;
    int maxRemoval = (int)std::fmin((int)((double)n * 0.0005), 35);
    double maxspeed = desc.max_substeps * step;
    int cur = 0;
    for (int i = desc.max_substeps - 1; i > 0; i--) {
        if (cur + hist[i] > maxRemoval) break;
        cur += hist[i];
        maxspeed = i * step;
    }
    float maxspeed_f = (float)maxspeed;
    double maxsq = maxspeed_f * maxspeed_f;
    check(launch_keep_flags(part[3].p, part[4].p, part[5].p, n, solid_removed.p, rank.p, maxsq,
                            keep.p, stream),
          "keep flags");
    const float *src[6];
    float *dst[6];
    for (int a2 = 0; a2 < 6; a2++) {
        src[a2] = part[a2].p;
        dst[a2] = part_alt[a2].p;
    }
    check(launch_compact_keep(keep.p, n, keep_pos.p, scan_tmp.p, src, dst, new_count.p, stream), "compact");
    int *nc_host = (int *)host_stage + 64;
    check(hipMemcpyAsync(nc_host, new_count.p, sizeof(int), hipMemcpyDeviceToHost, stream), "count");
    sync();
    n = *nc_host;
    for (int a2 = 0; a2 < 6; a2++) std::swap(part[a2].p, part_alt[a2].p);
    end_stage(FLIPMI_STAGE_REMOVE, ev);
}

double Scene::next_time_step(double dt) {
    // FluidSimulation::_calculateNextTimeStep (fluidsimulation.cpp:5593-5613).
    double maxu;
    if (current_frame == 0 && step_number == 0) {
        float bf[3] = {(float)desc.body_force[0], (float)desc.body_force[1], (float)desc.body_force[2]};
        float len = std::sqrt(bf[0] * bf[0] + bf[1] * bf[1] + bf[2] * bf[2]);
        // _predictMaximumMarkerParticleSpeed (fluidsimulation.cpp:5529-5551):
        // the fastest enabled inflow source plus one step of body force.
        maxu = 0.0;
        for (auto &src : sources) {
            if (!src->enabled || src->outflow) continue;
            const float *u3 = src->vel;
            float sp = std::sqrt(u3[0] * u3[0] + u3[1] * u3[1] + u3[2] * u3[2]);
            maxu = std::fmax(std::fmax((double)sp, 0.0), maxu);
        }
        maxu += (double)len * dt;
    } else {
        check(launch_max_speed(part[3].p, part[4].p, part[5].p, n, max_speed_bits.p, stream), "max speed");
        unsigned *bits = (unsigned *)host_stage;
        check(hipMemcpyAsync(bits, max_speed_bits.p, sizeof(unsigned), hipMemcpyDeviceToHost, stream),
              "max speed copy");
        sync();
        float maxsq;
        std::memcpy(&maxsq, bits, sizeof(float));
        maxu = std::sqrt((double)maxsq);
    }
    maxu = std::fmax(0.0, maxu);
    return desc.cfl * g.dx / (maxu + 1e-6);
}

// ---------------------------------------------------------------------------
// Mesh inflow sources
// ---------------------------------------------------------------------------
namespace {

// AABB(std::vector<vec3>) (aabb.cpp:56-80): float corner, double extents
// widened by 1e-9.
struct HostBox {
    float pos[3];
    double ext[3];
};
HostBox points_box(const float *v, size_t nv) {
    double mn[3] = {v[0], v[1], v[2]}, mx[3] = {v[0], v[1], v[2]};
    for (size_t q = 0; q < nv; q++)
        for (int a = 0; a < 3; a++) {
            mn[a] = std::fmin((double)v[3 * q + a], mn[a]);
            mx[a] = std::fmax((double)v[3 * q + a], mx[a]);
        }
    HostBox b;
    for (int a = 0; a < 3; a++) {
        b.pos[a] = (float)mn[a];
        b.ext[a] = mx[a] - mn[a] + 1e-9;
    }
    return b;
}
// Grid3d::positionToGridIndex and GridIndexToPosition (grid3d.h:49-83).
int floor_index(float x, double dx) { return (int)std::floor((double)x * (1.0 / dx)); }
float index_position(int i, double dx) { return (float)((double)(float)i * dx); }
// AABB::isPointInside (aabb.cpp:130-133) for the box [0, ext).
bool inside_origin_box(const float *p, const double ext[3]) {
    for (int a = 0; a < 3; a++)
        if (!(p[a] >= 0.0f && (double)p[a] < 0.0 + ext[a])) return false;
    return true;
}

uint64_t splitmix64_host(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
double draw_unit_host(uint64_t seed, uint64_t stream, uint64_t item) {
    uint64_t h = splitmix64_host(splitmix64_host(seed ^ (stream * 0xD1B54A32D192ED03ull)) + item);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
// The ray jitter of one inside test (MeshUtils::_getCollisionGridZ,
// meshutils.cpp:107-110: _randomDouble(jit, -jit) per axis, jit = 0.001 dx),
// drawn from the keyed generator: `which` 0 keys the source SDF's signs, 1
// the source cells.
void ray_jitter(uint64_t seed, int source, int which, double dx, float out[3]) {
    const double jit = 0.001 * dx;
    const uint64_t stream = ((uint64_t)(source + 1) << 8) | (uint64_t)which;
    for (int a = 0; a < 3; a++) out[a] = (float)(jit + draw_unit_host(seed, stream, a) * (-jit - jit));
}

// MeshLevelSet::_initializeTriangleData for one triangle (meshlevelset.cpp:
// 878-915): the node range of its AABB widened by the band, clamped to the
// (isize+1)(jsize+1)(ksize+1) node grid; false when it misses the grid.
bool triangle_band(const float v[3][3], int isize, int jsize, int ksize, double dx, int band,
                   BoxTri &T) {
    HostBox tb = points_box(&v[0][0], 3);
    const int size[3] = {isize, jsize, ksize};
    for (int a = 0; a < 3; a++) {
        float mx = tb.pos[a] + (float)tb.ext[a];  // AABB::getMaxPoint (aabb.cpp:493-495)
        T.gmin[a] = floor_index(tb.pos[a], dx) - band;
        T.gmax[a] = floor_index(mx, dx) + band + 1;
    }
    for (int a = 0; a < 3; a++)
        if (T.gmax[a] < 0 || T.gmin[a] >= size[a] + 1) return false;
    for (int a = 0; a < 3; a++) {
        T.gmin[a] = std::max(T.gmin[a], 0);
        T.gmax[a] = std::min(T.gmax[a], size[a]);
    }
    std::memcpy(T.v, v, sizeof(T.v));
    return true;
}

// The ray-column inputs of MeshUtils::getGridNodesInsideTriangleMesh for a
// mesh inside the node grid's box (meshutils.cpp:347-372): the mesh shifted
// by half a cell and moved to the grid offset `go` of its bounding box, the
// local cell grid (ci, cj, ck) and each triangle's ray-column range.
struct RayInput {
    std::vector<RayTri> tris;
    int ci = 0, cj = 0, ck = 0;
    int go[3] = {0, 0, 0};
};
RayInput ray_input(const std::vector<float> &verts, const std::vector<int> &tris, double dx) {
    RayInput r;
    const size_t nv = verts.size() / 3;
    std::vector<float> m(verts);
    const float hw = (float)(0.5 * dx);
    for (float &x : m) x = x + hw;  // TriangleMesh::translate (trianglemesh.cpp:732-736)
    HostBox bb = points_box(m.data(), nv);
    float off[3];
    for (int a = 0; a < 3; a++) {
        r.go[a] = floor_index(bb.pos[a], dx);
        off[a] = index_position(r.go[a], dx);
    }
    for (size_t q = 0; q < nv; q++)
        for (int a = 0; a < 3; a++) m[3 * q + a] = m[3 * q + a] + -off[a];
    int gmax[3];
    for (int a = 0; a < 3; a++) {
        float lo = bb.pos[a] - off[a];
        gmax[a] = floor_index(lo + (float)bb.ext[a], dx);  // Grid3d::getGridIndexBounds
    }
    r.ci = gmax[0] + 1; r.cj = gmax[1] + 1; r.ck = gmax[2] + 1;
    const size_t nt = tris.size() / 3;
    r.tris.resize(nt);
    for (size_t t = 0; t < nt; t++) {
        RayTri &T = r.tris[t];
        double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
        for (int c = 0; c < 3; c++) {
            const float *p = &m[3 * (size_t)tris[3 * t + c]];
            for (int a = 0; a < 3; a++) T.v[c][a] = p[a];
            for (int a = 0; a < 2; a++) {
                lo[a] = std::min(lo[a], (double)p[a]);
                hi[a] = std::max(hi[a], (double)p[a]);
            }
        }
        // A column's ray runs inside its cell, so only the columns the
        // triangle's box covers (one column of slack for rounding) can hit it.
        for (int a = 0; a < 2; a++) {
            T.c0[a] = (int)std::floor(lo[a] / dx) - 1;
            T.c1[a] = (int)std::floor(hi[a] / dx) + 1;
        }
    }
    return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// Static mesh obstacles
// ---------------------------------------------------------------------------
namespace {

constexpr int kSolidBand = 3;  // _solidLevelSetExactBand (fluidsimulation.h:1580)

// Level-set sub-grid of an obstacle: MeshObject::_getMeshIslandLevelSet for
// the whole mesh (meshobject.cpp:519-553), the mesh box widened by the band
// and clamped to the domain.  Empty when a side is not positive.
struct ObstacleGrid {
    int go[3] = {0, 0, 0};
    int ni = 0, nj = 0, nk = 0;  // node dims
    float off[3] = {0, 0, 0};
    bool empty = true;
};
ObstacleGrid obstacle_grid(const std::vector<float> &verts, int I, int J, int K, double dx) {
    ObstacleGrid og;
    HostBox bb = points_box(verts.data(), verts.size() / 3);
    const int top[3] = {I - 1, J - 1, K - 1};
    int gmax[3];
    for (int a = 0; a < 3; a++) {
        og.go[a] = (int)std::fmax(floor_index(bb.pos[a], dx) - kSolidBand, 0);
        gmax[a] = (int)std::fmin(floor_index(bb.pos[a] + (float)bb.ext[a], dx) + kSolidBand + 1, top[a]);
        og.off[a] = index_position(og.go[a], dx);
    }
    const int w = gmax[0] - og.go[0], h = gmax[1] - og.go[1], d = gmax[2] - og.go[2];
    if (w <= 0 || h <= 0 || d <= 0) return og;
    og.empty = false;
    og.ni = w + 1; og.nj = h + 1; og.nk = d + 1;
    return og;
}

// Islands: components of vertices joined by triangle edges
// (MeshUtils::_splitIntoMeshIslands, meshutils.cpp:455-563).
int count_islands(size_t nv, const std::vector<int> &tris) {
    std::vector<int> parent(nv);
    for (size_t q = 0; q < nv; q++) parent[q] = (int)q;
    auto find = [&](int a) {
        while (parent[a] != a) a = parent[a] = parent[parent[a]];
        return a;
    };
    for (size_t t = 0; t + 2 < tris.size(); t += 3) {
        int a = find(tris[t]);
        parent[find(tris[t + 1])] = a;
        parent[find(tris[t + 2])] = find(a);
    }
    int n = 0;
    for (size_t q = 0; q < nv; q++) n += find((int)q) == (int)q ? 1 : 0;
    return n;
}

}  // namespace

int Scene::add_obstacle(const float *verts, int nv, const int *tris, int nt) {
    if (!verts || !tris || nv < 3 || nt < 1)
        throw StatusError{FLIPMI_ERR_INVALID_ARG, "obstacle mesh needs at least 3 vertices and 1 triangle"};
    for (int q = 0; q < 3 * nt; q++)
        if (tris[q] < 0 || tris[q] >= nv)
            throw StatusError{FLIPMI_ERR_INVALID_ARG, "obstacle mesh triangle index out of range"};
    auto ob = std::make_unique<ObstacleMesh>();
    // TriangleMesh::removeExtraneousVertices: keep used vertices, in order.
    std::vector<int> map((size_t)nv, -1);
    for (int q = 0; q < 3 * nt; q++) map[(size_t)tris[q]] = 0;
    for (int q = 0; q < nv; q++) {
        if (map[(size_t)q] != 0) continue;
        for (int a = 0; a < 3; a++) {
            if (!std::isfinite(verts[3 * q + a]))
                throw StatusError{FLIPMI_ERR_INVALID_ARG, "obstacle mesh vertex is not finite"};
            ob->verts.push_back(verts[3 * q + a]);
        }
        map[(size_t)q] = (int)(ob->verts.size() / 3) - 1;
    }
    for (int q = 0; q < 3 * nt; q++) ob->tris.push_back(map[(size_t)tris[q]]);
    // Meshes of 25 or more islands take the reference's per-island path
    // (meshobject.cpp:254-262), which is not built.
    if (count_islands(ob->verts.size() / 3, ob->tris) >= 25)
        throw StatusError{FLIPMI_ERR_INVALID_ARG, "obstacle meshes of 25 or more islands are not supported"};
    // The inside test of the obstacle's level set needs the mesh inside the
    // sub-grid's box (the reference splits other meshes, meshutils.cpp:392-428).
    const double dx = g.dx;
    ObstacleGrid og = obstacle_grid(ob->verts, g.I, g.J, g.K, dx);
    if (!og.empty) {
        const double box[3] = {(og.ni - 1) * dx, (og.nj - 1) * dx, (og.nk - 1) * dx};
        for (size_t q = 0; q < ob->verts.size() / 3; q++) {
            float p[3];
            for (int a = 0; a < 3; a++) p[a] = ob->verts[3 * q + a] + -og.off[a];
            if (!inside_origin_box(p, box))
                throw StatusError{FLIPMI_ERR_INVALID_ARG,
                                  "obstacle mesh must lie at least 4 cells inside the domain grid"};
        }
    }
    obstacles.push_back(std::move(ob));
    solid_ready = false;
    return (int)obstacles.size() - 1;
}

// FluidSimulation::_addStaticObjectsToSDF (fluidsimulation.cpp:2932-2985):
// each enabled obstacle's level set on its sub-grid (exact band 3, signs by
// the ray test), unioned into the solid SDF in order.
void Scene::union_obstacles() {
    const double dx = g.dx;
    auto need = [&](hipError_t e) {
        if (e != hipSuccess) throw StatusError{FLIPMI_ERR_OOM, "obstacle allocation: out of device memory"};
    };
    // One obstacle's level set, unioned into the nodal field dst.
    auto add_island = [&](const ObstacleMesh &ob, size_t q, float *dst) {
        ObstacleGrid og = obstacle_grid(ob.verts, g.I, g.J, g.K, dx);
        if (og.empty) return;
        const size_t nt = ob.tris.size() / 3;
        // Triangle bands (meshlevelset.cpp:878-915): box from the untranslated
        // vertices, then moved by the grid offset, like the vertices.
        std::vector<BoxTri> band;
        for (size_t t = 0; t < nt; t++) {
            float v[3][3];
            for (int c = 0; c < 3; c++)
                for (int a = 0; a < 3; a++) v[c][a] = ob.verts[3 * (size_t)ob.tris[3 * t + c] + a];
            HostBox tb = points_box(&v[0][0], 3);
            BoxTri T;
            bool keep = true;
            const int size[3] = {og.ni - 1, og.nj - 1, og.nk - 1};
            for (int a = 0; a < 3; a++) {
                float lo = tb.pos[a] - og.off[a];
                float hi = lo + (float)tb.ext[a];
                T.gmin[a] = floor_index(lo, dx) - kSolidBand;
                T.gmax[a] = floor_index(hi, dx) + kSolidBand + 1;
                if (T.gmax[a] < 0 || T.gmin[a] >= size[a] + 1) keep = false;
            }
            if (!keep) continue;
            for (int a = 0; a < 3; a++) {
                T.gmin[a] = std::max(T.gmin[a], 0);
                T.gmax[a] = std::min(T.gmax[a], size[a]);
            }
            for (int c = 0; c < 3; c++)
                for (int a = 0; a < 3; a++) T.v[c][a] = v[c][a] - og.off[a];
            band.push_back(T);
        }
        const size_t nn = (size_t)og.ni * og.nj * og.nk;
        DBuf<float> phi;
        need(phi.alloc(nn));
        const float upper = (float)((double)(og.ni + og.nj + og.nk) * dx);  // meshlevelset.cpp:769-771
        DBuf<BoxTri> dband;
        if (band.empty()) {
            check(launch_fill_f32(phi.p, nn, upper, stream), "fill");
        } else {
            need(dband.alloc(band.size()));
            check(hipMemcpyAsync(dband.p, band.data(), dband.bytes(), hipMemcpyHostToDevice, stream), "upload");
            check(launch_mesh_exact_band(dband.p, (int)band.size(), og.ni, og.nj, og.nk, dx, upper, phi.p, stream),
                  "obstacle band");
        }
        std::vector<float> local(ob.verts.size());
        for (size_t q2 = 0; q2 < local.size(); q2++) local[q2] = ob.verts[q2] + -og.off[q2 % 3];
        RayInput r = ray_input(local, ob.tris, dx);
        DBuf<RayTri> dray;
        DBuf<int> overflow;
        need(dray.alloc(nt));
        need(overflow.alloc(1));
        check(launch_fill_i32(overflow.p, 1, 0, stream), "fill");
        check(hipMemcpyAsync(dray.p, r.tris.data(), dray.bytes(), hipMemcpyHostToDevice, stream), "upload");
        float jit[3];
        ray_jitter(jitter_seed, 100 + (int)q, 0, dx, jit);
        check(launch_ray_inside(dray.p, (int)nt, r.ci, r.cj, r.ck, dx, jit, r.go, og.ni, og.nj, og.nk, 1, nullptr,
                                phi.p, overflow.p, stream),
              "obstacle signs");
        check(launch_union_min(phi.p, og.ni, og.nj, og.nk, og.go, g.I + 1, g.J + 1, g.K + 1, dst, stream),
              "obstacle union");
        int ov = 0;
        check(hipMemcpyAsync(&ov, overflow.p, sizeof(int), hipMemcpyDeviceToHost, stream), "overflow");
        sync();
        if (ov != 0)
            throw StatusError{FLIPMI_ERR_INVALID_ARG, "obstacle mesh: a ray column crosses more than 64 triangles"};
    };
    bool inversed = false;
    for (size_t q = 0; q < obstacles.size(); q++) {
        if (!obstacles[q]->enabled) continue;
        if (obstacles[q]->inversed) inversed = true;
        else add_island(*obstacles[q], q, solid.p);
    }
    if (!inversed) return;
    // Inversed obstacles: their union, on a fresh full-grid level set, is
    // negated and then unioned into the solid (fluidsimulation.cpp:2968-2984).
    const int Ni = g.I + 1, Nj = g.J + 1, Nk = g.K + 1;
    const size_t nn = (size_t)Ni * Nj * Nk;
    DBuf<float> temp;
    need(temp.alloc(nn));
    check(launch_fill_f32(temp.p, nn, (float)((double)(Ni + Nj + Nk) * dx), stream), "fill");
    for (size_t q = 0; q < obstacles.size(); q++)
        if (obstacles[q]->enabled && obstacles[q]->inversed) add_island(*obstacles[q], q, temp.p);
    check(launch_negate(temp.p, nn, stream), "negate");
    const int origin[3] = {0, 0, 0};
    check(launch_union_min(temp.p, Ni, Nj, Nk, origin, Ni, Nj, Nk, solid.p, stream), "inversed union");
    sync();
}

int Scene::add_inflow(const float *verts, int nv, const int *tris, int nt, const double vel[3]) {
    if (!verts || !tris || nv < 3 || nt < 1)
        throw StatusError{FLIPMI_ERR_INVALID_ARG, "inflow mesh needs at least 3 vertices and 1 triangle"};
    if (sources.size() >= (size_t)kMaxInflow)
        throw StatusError{FLIPMI_ERR_INVALID_ARG, "too many inflow sources"};
    for (int q = 0; q < 3 * nt; q++)
        if (tris[q] < 0 || tris[q] >= nv)
            throw StatusError{FLIPMI_ERR_INVALID_ARG, "inflow mesh triangle index out of range"};
    const double dx = g.dx;
    const double domain[3] = {g.I * dx, g.J * dx, g.K * dx};
    for (int q = 0; q < nv; q++) {
        if (!std::isfinite(verts[3 * q]) || !std::isfinite(verts[3 * q + 1]) || !std::isfinite(verts[3 * q + 2]))
            throw StatusError{FLIPMI_ERR_INVALID_ARG, "inflow mesh vertex is not finite"};
        // MeshUtils::getGridNodesInsideTriangleMesh splits meshes that leave
        // the node grid's box (meshutils.cpp:392-428); only contained meshes
        // are supported here.
        if (!inside_origin_box(verts + 3 * q, domain))
            throw StatusError{FLIPMI_ERR_INVALID_ARG, "inflow mesh must lie inside the domain grid"};
    }
    auto src = std::make_unique<InflowSource>();
    src->verts.assign(verts, verts + 3 * (size_t)nv);
    src->tris.assign(tris, tris + 3 * (size_t)nt);
    for (int a = 0; a < 3; a++) src->vel[a] = (float)vel[a];
    // Sub-grid of the source SDF: MeshFluidSource::_getGridBoundsFromTriangleMesh
    // (meshfluidsource.cpp:318-333) with grid pad 4 (meshfluidsource.h:133).
    HostBox bb = points_box(verts, (size_t)nv);
    const double pad = 4.0 * dx, h = 0.5 * pad;  // AABB::expand (aabb.cpp:122-128)
    const int top[3] = {g.I - 1, g.J - 1, g.K - 1};
    int gmin[3], gmax[3];
    for (int a = 0; a < 3; a++) {
        float lo = bb.pos[a] - (float)h;
        double ext = bb.ext[a] + pad;
        gmin[a] = (int)std::fmax(floor_index(lo, dx), 0);
        gmax[a] = (int)std::fmin(floor_index(lo + (float)ext, dx), top[a]);
        src->go[a] = gmin[a];
        src->off[a] = index_position(gmin[a], dx);
    }
    src->ni = std::max(gmax[0] - gmin[0] + 1, 1) + 1;
    src->nj = std::max(gmax[1] - gmin[1] + 1, 1) + 1;
    src->nk = std::max(gmax[2] - gmin[2] + 1, 1) + 1;
    const double sub[3] = {(src->ni - 1) * dx, (src->nj - 1) * dx, (src->nk - 1) * dx};
    for (int q = 0; q < nv; q++) {
        float p[3];
        for (int a = 0; a < 3; a++) p[a] = verts[3 * q + a] + -src->off[a];
        if (!inside_origin_box(p, sub))
            throw StatusError{FLIPMI_ERR_INVALID_ARG,
                              "inflow mesh must lie at least 4 cells inside the domain grid"};
    }
    sources.push_back(std::move(src));
    return (int)sources.size() - 1;
}

void Scene::build_inflow(InflowSource &src, int id) {
    const double dx = g.dx;
    const int I = g.I, J = g.J, K = g.K;
    const size_t nt = src.tris.size() / 3;
    auto need = [&](hipError_t e) {
        if (e != hipSuccess) throw StatusError{FLIPMI_ERR_OOM, "inflow source allocation: out of device memory"};
    };
    DBuf<int> overflow;
    need(overflow.alloc(1));
    check(launch_fill_i32(overflow.p, 1, 0, stream), "fill");
    // Source SDF: MeshLevelSet::fastCalculateSignedDistanceField with exact
    // band 2 (meshfluidsource.h:116) on the mesh moved to the sub-grid.
    std::vector<float> local(src.verts.size());
    for (size_t q = 0; q < local.size(); q++) local[q] = src.verts[q] + -src.off[q % 3];
    std::vector<BoxTri> band;
    for (size_t t = 0; t < nt; t++) {
        float v[3][3];
        for (int c = 0; c < 3; c++)
            for (int a = 0; a < 3; a++) v[c][a] = local[3 * (size_t)src.tris[3 * t + c] + a];
        BoxTri T;
        if (triangle_band(v, src.ni - 1, src.nj - 1, src.nk - 1, dx, 2, T)) band.push_back(T);
    }
    const size_t nn = (size_t)src.ni * src.nj * src.nk;
    need(src.sdf.alloc(nn));
    const float upper = (float)((double)(src.ni + src.nj + src.nk) * dx);  // meshlevelset.cpp:769-771
    DBuf<BoxTri> dband;
    if (band.empty()) {
        check(launch_fill_f32(src.sdf.p, nn, upper, stream), "fill");
    } else {
        need(dband.alloc(band.size()));
        check(hipMemcpyAsync(dband.p, band.data(), dband.bytes(), hipMemcpyHostToDevice, stream), "upload");
        check(launch_mesh_exact_band(dband.p, (int)band.size(), src.ni, src.nj, src.nk, dx, upper, src.sdf.p,
                                     stream),
              "exact band");
    }
    float jit[3];
    RayInput rs = ray_input(local, src.tris, dx);
    DBuf<RayTri> dray;
    need(dray.alloc(nt));
    check(hipMemcpyAsync(dray.p, rs.tris.data(), dray.bytes(), hipMemcpyHostToDevice, stream), "upload");
    ray_jitter(jitter_seed, id, 0, dx, jit);
    check(launch_ray_inside(dray.p, (int)nt, rs.ci, rs.cj, rs.ck, dx, jit, rs.go, src.ni, src.nj, src.nk, 1,
                            nullptr, src.sdf.p, overflow.p, stream),
          "sdf signs");
    // Source cells: MeshObject::getCells on the domain's node grid.
    RayInput rc = ray_input(src.verts, src.tris, dx);
    DBuf<RayTri> dray2;
    need(dray2.alloc(nt));
    check(hipMemcpyAsync(dray2.p, rc.tris.data(), dray2.bytes(), hipMemcpyHostToDevice, stream), "upload");
    const size_t nnodes = (size_t)(I + 1) * (J + 1) * (K + 1), nc = (size_t)I * J * K;
    DBuf<uint8_t> nodes;
    need(nodes.alloc(nnodes));
    check(launch_fill_u8(nodes.p, nnodes, 0, stream), "fill");
    ray_jitter(jitter_seed, id, 1, dx, jit);
    check(launch_ray_inside(dray2.p, (int)nt, rc.ci, rc.cj, rc.ck, dx, jit, rc.go, I + 1, J + 1, K + 1, 0,
                            nodes.p, nullptr, overflow.p, stream),
          "source nodes");
    need(src.cells.alloc(nc));
    DBuf<int> flags, pos, cnt;
    need(flags.alloc(nc)); need(pos.alloc(nc)); need(cnt.alloc(1));
    check(launch_source_cell_scan(nodes.p, I, J, K, src.cells.p, flags.p, pos.p, scan_tmp.p, cnt.p, stream),
          "source cells");
    int host[2];
    check(hipMemcpyAsync(&host[0], cnt.p, sizeof(int), hipMemcpyDeviceToHost, stream), "count");
    check(hipMemcpyAsync(&host[1], overflow.p, sizeof(int), hipMemcpyDeviceToHost, stream), "overflow");
    sync();
    if (host[1] != 0)
        throw StatusError{FLIPMI_ERR_INVALID_ARG, "inflow mesh: a ray column crosses more than 64 triangles"};
    src.ncells = host[0];
    if (src.ncells > 0) {
        need(src.list.alloc((size_t)src.ncells));
        check(launch_source_cell_list(flags.p, pos.p, nc, src.list.p, stream), "cell list");
        std::vector<int> cl((size_t)src.ncells);
        check(hipMemcpyAsync(cl.data(), src.list.p, sizeof(int) * cl.size(), hipMemcpyDeviceToHost, stream),
              "cell list");
        sync();
        for (int a = 0; a < 3; a++) { src.cmin[a] = INT32_MAX; src.cmax[a] = -1; }
        for (int c : cl) {
            int ijk[3] = {c % I, (c / I) % J, c / (I * J)};
            for (int a = 0; a < 3; a++) {
                src.cmin[a] = std::min(src.cmin[a], ijk[a]);
                src.cmax[a] = std::max(src.cmax[a], ijk[a]);
            }
        }
    }
    src.built = true;
}

void Scene::ensure_inflow() {
    bool fresh = false;
    for (size_t q = 0; q < sources.size(); q++) {
        if (sources[q]->built) continue;
        build_inflow(*sources[q], (int)q);
        fresh = true;
    }
    if (!fresh) return;
    auto need = [&](hipError_t e) {
        if (e != hipSuccess) throw StatusError{FLIPMI_ERR_OOM, "inflow allocation: out of device memory"};
    };
    const size_t nc = (size_t)g.I * g.J * g.K;
    if (!inflow_mask.p) {
        need(inflow_mask.alloc((nc + 3) / 4));
        check(launch_fill_i32((int *)inflow_mask.p, inflow_mask.n, 0, stream), "fill");
    }
    int most = 0;
    for (auto &s : sources) most = std::max(most, s->ncells);
    const size_t m = (size_t)std::max(8 * most, 1);
    if (emit_flags.n < m) {
        need(emit_cand.alloc(3 * m));
        need(emit_flags.alloc(m));
        need(emit_pos.alloc(m));
        need(emit_scan.alloc(scan_scratch_elems((int)m)));
    }
    if (!emit_nbase.p) need(emit_nbase.alloc(1));
}

InflowViews Scene::inflow_views(bool emitting) const {
    InflowViews vs{};
    for (size_t q = 0; q < sources.size() && q < (size_t)kMaxInflow; q++) {
        const InflowSource &s = *sources[q];
        InflowView &v = vs.v[vs.n++];
        v.sdf = s.sdf.p;
        v.ni = s.ni; v.nj = s.nj; v.nk = s.nk;
        v.cells = s.cells.p;
        v.list = s.list.p;
        v.ncells = s.ncells;
        for (int a = 0; a < 3; a++) {
            v.off[a] = s.off[a];
            v.cmin[a] = s.cmin[a];
            v.cmax[a] = s.cmax[a];
            v.vel[a] = s.vel[a];
        }
        // _updateInflowMeshFluidSource (fluidsimulation.cpp:4564-4606): a
        // source with 0 substep emissions emits at the first substep only.
        const bool on = s.built && s.enabled;
        v.emit = emitting && on && !s.outflow && s.ncells > 0 && (s.emissions > 0 || step_number == 0);
        v.constrain = on && !s.outflow && s.constrain && s.ncells > 0 ? 1 : 0;
        v.outflow = on && s.outflow ? 1 : 0;
        v.inversed = s.inversed ? 1 : 0;
    }
    return vs;
}

bool Scene::any_inflow_constraint() const {
    for (auto &s : sources)
        if (s->enabled && !s->outflow && s->constrain) return true;
    return false;
}

void Scene::remove_outflow() {
    bool any = false;
    for (auto &s : sources) any = any || (s->enabled && s->outflow);
    if (!any || n == 0) return;
    hipEvent_t ev = nullptr;
    begin_stage(FLIPMI_STAGE_INFLOW, &ev);
    InflowViews vs = inflow_views(false);
    check(launch_outflow_keep(g, part[0].p, part[1].p, part[2].p, n, vs, keep.p, stream), "outflow keep");
    const float *src[6];
    float *dst[6];
    for (int a = 0; a < 6; a++) {
        src[a] = part[a].p;
        dst[a] = part_alt[a].p;
    }
    check(launch_compact_keep(keep.p, n, keep_pos.p, scan_tmp.p, src, dst, new_count.p, stream), "compact");
    int *nc_host = (int *)host_stage + 192;
    check(hipMemcpyAsync(nc_host, new_count.p, sizeof(int), hipMemcpyDeviceToHost, stream), "count");
    sync();
    n = *nc_host;
    for (int a = 0; a < 6; a++) std::swap(part[a].p, part_alt[a].p);
    end_stage(FLIPMI_STAGE_INFLOW, ev);
}

void Scene::grow(int64_t need_n) {
    if (need_n <= cap) return;
    const int64_t limit = INT32_MAX / 2;
    if (need_n > limit) throw StatusError{FLIPMI_ERR_OOM, "particle count exceeds the supported maximum"};
    const int64_t target = std::min<int64_t>(std::max<int64_t>(need_n, (int64_t)cap + cap / 2), limit);
    const size_t pc = (size_t)target;
    auto need = [&](hipError_t e) {
        if (e != hipSuccess) throw StatusError{FLIPMI_ERR_OOM, "particle growth: out of device memory"};
    };
    for (int a = 0; a < 6; a++) {
        DBuf<float> fresh;
        need(fresh.alloc(pc));
        if (n > 0)
            check(hipMemcpyAsync(fresh.p, part[a].p, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, stream),
                  "grow");
        sync();
        std::swap(part[a].p, fresh.p);
        std::swap(part[a].n, fresh.n);
        part_alt[a].release();
        need(part_alt[a].alloc(pc));
    }
    spos.release();
    need(spos.alloc(pc));
    for (int a = 0; a < 3; a++) {
        svel[a].release();
        need(svel[a].alloc(pc));
    }
    cell.release(); slot.release(); order.release(); keep.release(); keep_pos.release(); rank.release();
    solid_removed.release(); scan_tmp.release();
    need(cell.alloc(pc)); need(slot.alloc(pc)); need(order.alloc(pc));
    need(keep.alloc(pc)); need(keep_pos.alloc(pc)); need(rank.alloc(pc));
    need(solid_removed.alloc(pc));
    need(scan_tmp.alloc(scan_elems(g, pc)));
    cap = (int)target;
}

// FluidSimulation::_updateInflowMeshFluidSources (fluidsimulation.cpp:
// 4675-4697) at the end of a substep.
void Scene::emit_inflow() {
    if (sources.empty()) return;
    InflowViews vs = inflow_views(true);
    int64_t extra = 0;
    for (int q = 0; q < vs.n; q++)
        if (vs.v[q].emit) extra += 8 * (int64_t)vs.v[q].ncells * std::max(sources[q]->emissions, 1);
    if (extra == 0) return;
    hipEvent_t ev = nullptr;
    begin_stage(FLIPMI_STAGE_INFLOW, &ev);
    grow((int64_t)n + extra);
    check(launch_inflow_mask(g, part[0].p, part[1].p, part[2].p, n, vs, inflow_mask.p, stream), "inflow mask");
    check(launch_fill_i32(emit_nbase.p, 1, n, stream), "fill");
    EmitOut out{part[0].p, part[1].p, part[2].p, part[3].p, part[4].p, part[5].p};
    EmitParams ep{};
    ep.jitter = 0.25 * (jitter_factor - 1e-3) * g.dx;  // fluidsimulation.cpp:2623-2626
    ep.jitter_surface = jitter_surface ? 1 : 0;
    ep.seed = jitter_seed;
    for (int q = 0; q < vs.n; q++) {
        if (!vs.v[q].emit) continue;
        const int passes = std::max(sources[q]->emissions, 1);
        for (int e = 0; e < passes; e++) {
            ep.stream = 0x1000000ull + emission_count++;
            check(launch_inflow_emit(g, vs.v[q], solid.p, inflow_mask.p, ep, emit_cand.p, emit_flags.p,
                                     emit_pos.p, emit_scan.p, emit_nbase.p, out, stream),
                  "inflow emit");
        }
    }
    int *nb = (int *)host_stage + 128;
    check(hipMemcpyAsync(nb, emit_nbase.p, sizeof(int), hipMemcpyDeviceToHost, stream), "count");
    sync();
    n = *nb;
    end_stage(FLIPMI_STAGE_INFLOW, ev);
}

void Scene::step_fluid(double dt) {
    ensure_solid();
    if (!sources.empty()) ensure_inflow();
    hipEvent_t ev = nullptr;
    // Liquid SDF + post-process (fluidsimulation.cpp:3225-3247).
    begin_stage(FLIPMI_STAGE_SORT, &ev);
    sort_subcells();
    const float *src[6];
    for (int a = 0; a < 6; a++) src[a] = part[a].p;
    float *sorted_vel[3] = {svel[0].p, svel[1].p, svel[2].p};
    check(launch_gather_sorted(order.p, sub_offset.p + subcell_bins(g), n, src, spos.p, sorted_vel, stream),
          "gather");
    end_stage(FLIPMI_STAGE_SORT, ev);
    begin_stage(FLIPMI_STAGE_CLASSIFY, &ev);
    check(launch_liquid_sdf(g, spos.p, sub_offset.p, radius, solid.p, liquid.p, stream), "liquid sdf");
    end_stage(FLIPMI_STAGE_CLASSIFY, ev);
    // _advectVelocityField (fluidsimulation.cpp:3253-3277).
    int layers = (int)std::ceil(desc.cfl) + 2;
    size_t nU = u.n, nV = v.n, nW = w.n;
    if (n > 0) {
        begin_stage(FLIPMI_STAGE_P2G, &ev);
        const float *sv3[3] = {svel[0].p, svel[1].p, svel[2].p};
        float *outs[3] = {u.p, v.p, w.p};
        uint8_t *valids[3] = {vu.p, vv.p, vw.p};
        check(launch_p2g(g, spos.p, sv3, sub_offset.p, n, radius, outs, valids, p2g_ovf_count.p, p2g_ovf.p,
                         stream),
              "p2g");
        end_stage(FLIPMI_STAGE_P2G, ev);
        extrapolate_all(layers);
    } else {
        check(launch_fill_f32(u.p, nU, 0.0f, stream), "fill");
        check(launch_fill_f32(v.p, nV, 0.0f, stream), "fill");
        check(launch_fill_f32(w.p, nW, 0.0f, stream), "fill");
        check(launch_fill_u8(vu.p, nU, 0, stream), "fill");
        check(launch_fill_u8(vv.p, nV, 0, stream), "fill");
        check(launch_fill_u8(vw.p, nW, 0, stream), "fill");
    }
    // _saveVelocityField + _applyConstantBodyForces (fluidsimulation.cpp:
    // 3287-3297, 3450-3489).
    begin_stage(FLIPMI_STAGE_GLUE, &ev);
    float bf[3] = {(float)desc.body_force[0], (float)desc.body_force[1], (float)desc.body_force[2]};
    float *macs[3] = {u.p, v.p, w.p};
    float *saved[3] = {su.p, sv.p, sw.p};
    size_t sizes[3] = {nU, nV, nW};
    const bool held = any_inflow_constraint();
    const InflowViews views = held ? inflow_views(false) : InflowViews{};
    const uint8_t *valid[3] = {vu.p, vv.p, vw.p};
    for (int c = 0; c < 3; c++) {
        int do_add = std::fabs(bf[c]) > 1e-6f ? 1 : 0;
        float add = (float)((double)bf[c] * dt);
        if (held)
            check(launch_save_and_body_force_inflow(g, c, macs[c], saved[c], add, do_add, valid[c], views, stream),
                  "save/force");
        else
            check(launch_save_and_body_force(sizes[c], macs[c], saved[c], add, do_add, stream), "save/force");
    }
    end_stage(FLIPMI_STAGE_GLUE, ev);
    // _pressureSolve (fluidsimulation.cpp:3732-3769).
    pressure_solve(dt);
    extrapolate_all(layers);
    // _constrainVelocityFields (fluidsimulation.cpp:3937-3948).
    begin_stage(FLIPMI_STAGE_GLUE, &ev);
    const float *weights[3] = {wU.p, wV.p, wW.p};
    const float *svel[3] = {sU.p, sV.p, sW.p};
    for (int c = 0; c < 3; c++) {
        check(launch_constrain(sizes[c], saved[c], weights[c], svel[c], stream), "constrain");
        check(launch_constrain(sizes[c], macs[c], weights[c], svel[c], stream), "constrain");
    }
    end_stage(FLIPMI_STAGE_GLUE, ev);
    // _updateMarkerParticleVelocities + _advanceMarkerParticles.
    if (n > 0) advance_particles(dt);
    // _updateFluidObjects (fluidsimulation.cpp:4794-4806).
    emit_inflow();
    remove_outflow();
}

flipmi_frame_stats Scene::update(double dt) {
    if (!initialized)
        throw StatusError{FLIPMI_ERR_NOT_INITIALIZED, "FluidSimulation must be initialized before update"};
    if (dt < 0.0) throw StatusError{FLIPMI_ERR_INVALID_ARG, "delta time must be greater than or equal to 0"};
    hipEvent_t f0 = take_event(), f1 = take_event();
    check(hipEventRecord(f0, stream), "event");
    double epsdt = 1e-6;
    bool zero = dt < epsdt;
    dt = std::max(dt, epsdt);
    frame_dt = dt;
    remaining = dt;
    step_number = 0;
    bool skipped = zero && output_initialized;
    double substepTime = frame_dt / (double)desc.min_substeps;
    const double eps = 1e-9;
    do {
        cur_step = std::fmin(next_time_step(dt), remaining);
        double done = frame_dt - remaining;
        double limit = (step_number + 1) * substepTime;
        if (done + cur_step > limit) cur_step = std::fmin(substepTime, remaining);
        if (step_number == desc.max_substeps - 1) cur_step = remaining;
        remaining -= cur_step;
        if (!skipped) {
            step_fluid(cur_step);
            // _outputSimulationData (fluidsimulation.cpp:5445-5463): the
            // surface mesh at the end of the frame's first substep.
            if (step_number == 0 && surface.enabled) build_mesh();
        }
        step_number++;
    } while (remaining > eps);
    check(hipEventRecord(f1, stream), "event");
    sync();
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, f0, f1), "event");
    event_pool.push_back(f0);
    event_pool.push_back(f1);
    collect_events();
    output_initialized = true;
    current_frame++;
    flipmi_frame_stats st{};
    st.substeps = step_number;
    st.num_particles = n;
    st.pcg_iterations = last_iters;
    st.pcg_residual = last_resid;
    st.pressure_ok = last_ok;
    st.gpu_ms = ms;
    st.pressure_rows = last_rows;
    st.pcg_refinements = last_refinements;
    return st;
}

}  // namespace flipmi
