// Serving: a device-resident request loop over independent scenes.
//
// Requests (a scene description, its particles, a dt and a number of
// FluidSimulation::update calls) are queued as they arrive and admitted in
// arrival order whenever a worker slot is free and the scene's device memory
// fits the budget.  Each worker slot is one host thread with its own HIP
// stream, so the kernels of the scenes in flight overlap on the device; a
// request that arrives while others run starts as soon as a slot and memory
// free up (continuous batching).  Finished scenes stay resident, idle, keyed
// by their grid and settings: the next request of the same kind re-arms one
// (new particles, state of a freshly initialised scene, static solid and
// weights kept) instead of allocating, and idle scenes are evicted, least
// recently used first, when a new kind needs the memory.  A request's result
// does not depend on what else runs: every scene owns all of its buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "scene.h"

using namespace flipmi;

namespace {

using Clock = std::chrono::steady_clock;

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw StatusError{FLIPMI_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    size_t k = (size_t)std::llround(q * (double)(v.size() - 1));
    return v[std::min(k, v.size() - 1)];
}

// Scenes are interchangeable when everything but the particles matches.
bool same_kind(const flipmi_scene_desc &a, const flipmi_scene_desc &b) {
    return a.isize == b.isize && a.jsize == b.jsize && a.ksize == b.ksize && a.dx == b.dx &&
           a.body_force[0] == b.body_force[0] && a.body_force[1] == b.body_force[1] &&
           a.body_force[2] == b.body_force[2] && a.cfl == b.cfl && a.min_substeps == b.min_substeps &&
           a.max_substeps == b.max_substeps && a.pic_ratio == b.pic_ratio && a.pcg_tolerance == b.pcg_tolerance &&
           a.pcg_max_iterations == b.pcg_max_iterations && a.pcg_precision == b.pcg_precision;
}

struct Request {
    int64_t id = 0;
    flipmi_scene_desc desc{};
    std::vector<float> pos, vel;
    double dt = 0.0;
    int steps = 0;
    Clock::time_point t_submit, t_start, t_done;
    bool done = false;
    flipmi_request_report rep{};
    std::string error;
    std::vector<double> step_ms;
    std::vector<float> out_pos, out_vel;
};

struct Pooled {
    std::unique_ptr<Scene> scene;
    size_t bytes = 0;
    bool busy = false;
    uint64_t last_use = 0;
};

}  // namespace

struct flipmi_server {
    flipmi_ctx *ctx = nullptr;
    flipmi_server_desc desc{};
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::map<int64_t, std::unique_ptr<Request>> requests;
    std::vector<std::unique_ptr<Pooled>> pool;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;
    bool stopping = false;
    int64_t next_id = 1;
    size_t resident = 0, peak = 0;
    uint64_t tick = 0;
    flipmi_server_stats st{};
    std::vector<double> all_step_ms;

    void worker(int slot);
    // Under the lock: a scene for the head request (reserved and marked
    // busy), or null when it has to wait.  Sets *fail for a request that can
    // never fit.
    Pooled *admit(Request *r, bool *reused, bool *fail, std::vector<std::unique_ptr<Pooled>> &evicted);
};

namespace {

int64_t capacity_for(const flipmi_server &s, const Request &r) {
    const double cells = (double)r.desc.isize * r.desc.jsize * r.desc.ksize;
    const int64_t n = (int64_t)(r.pos.size() / 3);
    return std::max<int64_t>({n, r.desc.max_particles, (int64_t)std::ceil(s.desc.particle_capacity_per_cell * cells)});
}

}  // namespace

Pooled *flipmi_server::admit(Request *r, bool *reused, bool *fail, std::vector<std::unique_ptr<Pooled>> &evicted) {
    *fail = false;
    *reused = false;
    const int64_t n = (int64_t)(r->pos.size() / 3);
    // An idle resident scene of the same kind that holds the particles.
    for (auto &p : pool) {
        if (!p->busy && same_kind(p->scene->desc, r->desc) && p->scene->cap >= n) {
            p->busy = true;
            *reused = true;
            return p.get();
        }
    }
    const size_t need = estimate_scene_bytes(r->desc, capacity_for(*this, *r));
    if ((int64_t)need > desc.memory_budget) {
        *fail = true;
        return nullptr;
    }
    // Evict idle scenes, least recently used first, until the new one fits.
    while (resident + need > (size_t)desc.memory_budget) {
        auto victim = pool.end();
        for (auto it = pool.begin(); it != pool.end(); ++it)
            if (!(*it)->busy && (victim == pool.end() || (*it)->last_use < (*victim)->last_use)) victim = it;
        if (victim == pool.end()) return nullptr;  // wait for a running scene to finish
        resident -= (*victim)->bytes;
        evicted.push_back(std::move(*victim));
        pool.erase(victim);
        st.scenes_evicted++;
    }
    std::unique_ptr<Pooled> p(new Pooled());
    p->bytes = need;
    p->busy = true;
    resident += need;
    peak = std::max(peak, resident);
    pool.push_back(std::move(p));
    return pool.back().get();
}

void flipmi_server::worker(int slot) {
    (void)hipSetDevice(ctx->c.device);
    hipStream_t stream = streams[slot];
    for (;;) {
        Request *r = nullptr;
        Pooled *p = nullptr;
        bool reused = false;
        std::vector<std::unique_ptr<Pooled>> evicted;
        {
            std::unique_lock<std::mutex> lk(m);
            for (;;) {
                if (!queue.empty()) {
                    bool fail = false;
                    p = admit(queue.front(), &reused, &fail, evicted);
                    if (fail) {
                        Request *bad = queue.front();
                        queue.pop_front();
                        bad->rep.status = FLIPMI_ERR_BUDGET;
                        bad->error = "the scene does not fit the serving memory budget";
                        bad->done = true;
                        bad->t_start = bad->t_done = Clock::now();
                        st.failed++;
                        st.queued--;
                        cv_done.notify_all();
                        continue;
                    }
                    if (p) {
                        r = queue.front();
                        queue.pop_front();
                        st.queued--;
                        st.active++;
                        break;
                    }
                }
                if (stopping && queue.empty()) return;
                cv_work.wait(lk);
            }
        }
        evicted.clear();  // frees the evicted scenes' memory outside the lock
        r->t_start = Clock::now();
        int status = FLIPMI_OK;
        std::string error;
        try {
            const int64_t n = (int64_t)(r->pos.size() / 3);
            const float *vel = r->vel.empty() ? nullptr : r->vel.data();
            bool armed = false;
            if (reused) {
                p->scene->stream = stream;
                armed = p->scene->rearm(r->pos.data(), vel, n);
            }
            if (!armed) {
                // A new scene on this slot's stream.
                std::unique_ptr<Scene> s(new Scene());
                s->ctx = &ctx->c;
                s->desc = r->desc;
                s->desc.max_particles = capacity_for(*this, *r);
                s->g = make_geom(r->desc.isize, r->desc.jsize, r->desc.ksize, r->desc.dx);
                s->stream = stream;
                s->own_stream = false;
                s->launch_graphs = this->desc.launch_graphs != 0;
                s->load(r->pos.data(), vel, n);
                p->scene.reset();
                p->scene = std::move(s);
                p->scene->initialize();
                reused = false;
            }
            Scene &s = *p->scene;
            for (int q = 0; q < r->steps; q++) {
                const int64_t np = s.n;
                Clock::time_point t0 = Clock::now();
                flipmi_frame_stats fs = s.update(r->dt);
                r->step_ms.push_back(ms_between(t0, Clock::now()));
                r->rep.particle_substeps += np * fs.substeps;
                if (!fs.pressure_ok) r->rep.pressure_failures++;
                r->rep.steps_done++;
            }
            // Final particles, AoS.
            r->out_pos.assign(3 * (size_t)s.n, 0.0f);
            r->out_vel.assign(3 * (size_t)s.n, 0.0f);
            std::vector<float> t((size_t)s.n);
            for (int a = 0; a < 6 && s.n > 0; a++) {
                hip_check(hipMemcpyAsync(t.data(), s.part[a].p, sizeof(float) * s.n, hipMemcpyDeviceToHost, stream),
                          "download");
                hip_check(hipStreamSynchronize(stream), "sync");
                std::vector<float> &dst = a < 3 ? r->out_pos : r->out_vel;
                for (int q = 0; q < s.n; q++) dst[3 * (size_t)q + (a % 3)] = t[q];
            }
            r->rep.num_particles = s.n;
        } catch (const StatusError &e) {
            status = e.code;
            error = e.msg;
        } catch (const std::bad_alloc &) {
            status = FLIPMI_ERR_OOM;
            error = "host allocation failed";
        } catch (...) {
            status = FLIPMI_ERR_HIP;
            error = "unexpected error";
        }
        (void)hipStreamSynchronize(stream);
        r->t_done = Clock::now();
        {
            std::lock_guard<std::mutex> lk(m);
            r->rep.status = status;
            r->error = error;
            r->rep.reused_scene = reused ? 1 : 0;
            r->rep.queue_ms = ms_between(r->t_submit, r->t_start);
            r->rep.run_ms = ms_between(r->t_start, r->t_done);
            r->rep.step_ms_p50 = percentile(r->step_ms, 0.5);
            r->rep.step_ms_p99 = percentile(r->step_ms, 0.99);
            r->done = true;
            st.active--;
            if (status == FLIPMI_OK) {
                st.completed++;
                st.particle_substeps += r->rep.particle_substeps;
                st.steps += r->rep.steps_done;
                all_step_ms.insert(all_step_ms.end(), r->step_ms.begin(), r->step_ms.end());
                if (reused) st.scenes_reused++;
                else st.scenes_created++;
            } else {
                st.failed++;
            }
            if (status == FLIPMI_OK && p->scene) {
                // Account the scene's actual allocation from now on.
                size_t actual = p->scene->device_bytes();
                resident = resident - p->bytes + actual;
                p->bytes = actual;
                peak = std::max(peak, resident);
                p->busy = false;
                p->last_use = ++tick;
            } else {
                // A failed scene is not kept.
                resident -= p->bytes;
                for (auto it = pool.begin(); it != pool.end(); ++it)
                    if (it->get() == p) {
                        evicted.push_back(std::move(*it));
                        pool.erase(it);
                        break;
                    }
            }
        }
        evicted.clear();
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

extern "C" {

void flipmi_server_desc_default(flipmi_server_desc *d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->max_active = 8;
    d->memory_budget = (int64_t)200 << 30;
    d->particle_capacity_per_cell = 3.0;
    d->launch_graphs = 1;
}

int flipmi_server_create(flipmi_ctx *ctx, const flipmi_server_desc *desc, flipmi_server **out) {
    if (!ctx || !out) return FLIPMI_ERR_INVALID_ARG;
    *out = nullptr;
    flipmi_server_desc d;
    flipmi_server_desc_default(&d);
    if (desc) d = *desc;
    if (d.max_active < 1 || d.max_active > 64 || d.memory_budget <= 0 || !(d.particle_capacity_per_cell >= 0.0)) {
        ctx->c.last_error = "server: max_active must be in [1, 64], the budget positive";
        return FLIPMI_ERR_INVALID_ARG;
    }
    if (hipSetDevice(ctx->c.device) != hipSuccess) return FLIPMI_ERR_HIP;
    std::unique_ptr<flipmi_server> s(new (std::nothrow) flipmi_server());
    if (!s) return FLIPMI_ERR_OOM;
    s->ctx = ctx;
    s->desc = d;
    for (int q = 0; q < d.max_active; q++) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (hipStream_t x : s->streams) (void)hipStreamDestroy(x);
            ctx->c.last_error = "server: stream creation failed";
            return FLIPMI_ERR_HIP;
        }
        s->streams.push_back(st);
    }
    flipmi_server *raw = s.get();
    for (int q = 0; q < d.max_active; q++) s->workers.emplace_back([raw, q] { raw->worker(q); });
    *out = s.release();
    return FLIPMI_OK;
}

int flipmi_server_submit(flipmi_server *s, const flipmi_request *req, int64_t *id) {
    if (!s || !req || (req->n > 0 && !req->pos_xyz) || req->n < 0 || req->steps < 0 || !(req->dt >= 0.0))
        return FLIPMI_ERR_INVALID_ARG;
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r) return FLIPMI_ERR_OOM;
    try {
        validate_scene_desc(req->scene);
        r->desc = req->scene;
        r->pos.assign(req->pos_xyz, req->pos_xyz + 3 * (size_t)req->n);
        if (req->vel_xyz) r->vel.assign(req->vel_xyz, req->vel_xyz + 3 * (size_t)req->n);
    } catch (const StatusError &e) {
        std::lock_guard<std::mutex> lk(s->m);
        s->ctx->c.last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        return FLIPMI_ERR_OOM;
    }
    r->dt = req->dt;
    r->steps = req->steps;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->m);
        if (s->stopping) return FLIPMI_ERR_INVALID_ARG;
        r->id = s->next_id++;
        if (id) *id = r->id;
        s->queue.push_back(r.get());
        s->requests[r->id] = std::move(r);
        s->st.submitted++;
        s->st.queued++;
    }
    s->cv_work.notify_one();
    return FLIPMI_OK;
}

int flipmi_server_wait(flipmi_server *s, int64_t id, double timeout_s, flipmi_request_report *rep) {
    if (!s) return FLIPMI_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) {
        s->ctx->c.last_error = "server: unknown request id";
        return FLIPMI_ERR_INVALID_ARG;
    }
    Request *r = it->second.get();
    auto ready = [&] { return r->done; };
    if (timeout_s < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return FLIPMI_ERR_TIMEOUT;
    if (rep) *rep = r->rep;
    if (r->rep.status != FLIPMI_OK) s->ctx->c.last_error = r->error;
    return r->rep.status;
}

int flipmi_server_take(flipmi_server *s, int64_t id, float *pos_xyz, float *vel_xyz) {
    if (!s) return FLIPMI_ERR_INVALID_ARG;
    std::unique_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->m);
        auto it = s->requests.find(id);
        if (it == s->requests.end() || !it->second->done) {
            s->ctx->c.last_error = "server: unknown or unfinished request";
            return FLIPMI_ERR_INVALID_ARG;
        }
        r = std::move(it->second);
        s->requests.erase(it);
    }
    if (pos_xyz && !r->out_pos.empty()) std::memcpy(pos_xyz, r->out_pos.data(), sizeof(float) * r->out_pos.size());
    if (vel_xyz && !r->out_vel.empty()) std::memcpy(vel_xyz, r->out_vel.data(), sizeof(float) * r->out_vel.size());
    return r->rep.status;
}

int flipmi_server_get_stats(flipmi_server *s, flipmi_server_stats *out) {
    if (!s || !out) return FLIPMI_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->m);
    *out = s->st;
    out->step_ms_p50 = percentile(s->all_step_ms, 0.5);
    out->step_ms_p99 = percentile(s->all_step_ms, 0.99);
    out->resident_bytes = (int64_t)s->resident;
    out->peak_resident_bytes = (int64_t)s->peak;
    return FLIPMI_OK;
}

int flipmi_server_destroy(flipmi_server *s) {
    if (!s) return FLIPMI_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    s->pool.clear();
    for (hipStream_t x : s->streams) (void)hipStreamDestroy(x);
    delete s;
    return FLIPMI_OK;
}

}  // extern "C"
