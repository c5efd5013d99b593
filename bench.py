#!/usr/bin/env python3
"""FLIP step throughput on one MI355X: particle-substeps/s.

Workload: the 256x128x128 double dam break (dx = 1/128, a 2 x 1 x 1 m box
with solid walls; two cuboids x in [0,0.5] and [1.5,2.0], y in [0,0.6],
full z, 8 jittered particles per cell, seed 3, gravity -9.81).  One step is
one FluidSimulation::update(1/240) with min = max substeps = 1, i.e. exactly
one substep of the whole FLIP pipeline (sort, classify, P2G, extrapolate,
assemble, MIC(0), PCG, apply, G2P + RK3 advection, collision, removal) over
every particle.  Particles and grids are resident in HBM before the timed
region; nothing is copied to the host inside it except the per-substep
scalars the step itself needs (speed maximum, histogram, PCG status).

Prints ONE JSON line.  Besides the contract keys it carries
  roofline:     the dominant kernel's algorithmic HBM bytes per launch over
                its average launch time (hipEvents on the scene's stream)
                against the MI355X peak of 8 TB/s;
  cpu_baseline: the reference CPU engine (oracle/_ref, when built) or the
                CPU oracle timed on a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# The roofline kernel is timed with HIP events on every KERNEL_SAMPLE-th PCG
# iteration of the timed region (timing every launch slowed a step by 2 %).
KERNEL_SAMPLE = 8

# Algorithmic bytes per matrix row (fluid cell) of each PCG kernel: every
# array the kernel must read or write, once.  The matrix is the FP64 diagonal
# and the three upper couplings (the lower three are the neighbours' upper
# ones), stored as the FP32 face weights the FP64 entries are rebuilt from
# exactly: 8 + 3 x 4 = 20 bytes on both paths.  FP64 path, 8-byte vectors:
#   spmv_dot:   s, matrix, z (= A s)                       -> 8 + 20 + 8
#   axpy_norm:  x r/w, s, r r/w, z                          -> 6 x 8
#   forward:    r, invdiag, 3 lower factor entries, z       -> 6 x 8
#   backward:   z r/w, invdiag, 3 upper factor entries, and
#               r for the fused z.r                         -> 7 x 8
#   update_s:   z, s r/w                                    -> 3 x 8
# Mixed path: FP32 vectors and factor, FP64 x and matrix entries:
#   spmv_dot:   z, old s, matrix, new s, q (= A s)          -> 4 + 4 + 20 + 4 + 4
#   forward with the axpy folded in: r r/w, q, invdiag, 3 lower factor
#               entries, z, and x r/w (FP64), s             -> 8 + 4 + 4 + 12 + 4 + 16 + 4
#   backward:   7 x 4
#   the direction update is folded into spmv_dot.
# A true residual (mixed path only) reads x and b (FP64) and the matrix and
# writes r (FP32): 8 + 8 + 20 + 4 = 40 bytes; the forward sweep over it
# (the same launch) adds r, invdiag, 3 lower factor entries, z: 6 x 4.
PCG_BYTES = {
    "fp64": {"k_spmv_dot": 36, "k_axpy_norm": 48, "k_forward": 48, "k_backward": 56, "k_update_s": 24},
    "mixed": {"k_spmv_dot": 36, "k_axpy_norm": 0, "k_forward": 52, "k_backward": 28, "k_update_s": 0},
}
TRUE_RESIDUAL_BYTES = 40 + 24

# HBM bytes moved per launch, measured with rocprofv3 FETCH_SIZE and WRITE_SIZE
# passes scaled by a plain 256 MiB copy (x2.0 and x1.0), on the default
# workload; the commands and counter readings are in
# profiles/traffic_pcg_kernels_double_dam_256x128x128_mixed.md.  Keyed by
# (precision, grid); other workloads report null.
MEASURED_TRAFFIC = {
    ("mixed", (256, 128, 128)): {"k_backward": 59.63e6, "k_forward": 99.55e6, "k_spmv_dot": 69.48e6,
                                 "k_true_residual": 116.03e6},
}

def parse_args(argv=None):
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--size", type=int, nargs=3, default=[256, 128, 128], metavar=("I", "J", "K"))
    p.add_argument("--dt", type=float, default=1.0 / 240.0)
    p.add_argument("--seed", type=int, default=3)
    p.add_argument("--precision", choices=["fp64", "mixed"], default="mixed",
                   help="PCG arithmetic of the measured path (the 256x128x128 workload is the mixed-precision one)")
    p.add_argument("--fp64-path", dest="fp64_path", action="store_true", default=True,
                   help="also time the reference-arithmetic FP64 PCG path (reported under config.fp64_path)")
    p.add_argument("--no-fp64-path", dest="fp64_path", action="store_false")
    p.add_argument("--profiling", type=int, default=2, choices=[0, 1, 2],
                   help="hipEvent timing: 2 = every PCG kernel launch (needed for the roofline line)")
    p.add_argument("--cpu-baseline", dest="cpu_baseline", action="store_true", default=True)
    p.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    p.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, cpus available)")
    p.add_argument("--cpu-seconds", type=float, default=20.0,
                   help="stop the CPU sample after the first substep that ends past this many seconds")
    p.add_argument("--serving", action="store_true",
                   help="the serving workload instead: Poisson arrivals of independent mixed-size scenes")
    p.add_argument("--requests", type=int, default=256, help="serving: number of requests")
    p.add_argument("--rate", type=float, default=20.0, help="serving: mean arrivals per second")
    p.add_argument("--request-steps", type=int, default=30, help="serving: update calls per request")
    p.add_argument("--max-active", type=int, default=16, help="serving: scenes stepped at once")
    p.add_argument("--budget-gib", type=float, default=200.0, help="serving: device memory budget")
    p.add_argument("--hw-queues", type=int, default=16,
                   help="serving: hardware queues per process (GPU_MAX_HW_QUEUES, at most 32; 0 = as set)")
    p.add_argument("--launch-graphs", type=int, default=1, choices=[0, 1],
                   help="serving: PCG iteration batches as captured HIP graphs")
    p.add_argument("--plain-requests", type=int, default=24,
                   help="serving: requests also run one at a time on the plain FP64 path")
    p.add_argument("--workload", choices=["dam_break_64", "pool_drop_128", "large_pool_512", "double_dam_surface_256"],
                   default=None,
                   help="one of the other workloads instead of the headline (see run_workload)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only copy 256 MiB device to device 3 times (known bytes, to scale HBM counters)")
    p.add_argument("--verbose", action="store_true")
    return p.parse_args(argv)


def log(msg, args):
    if args.verbose:
        print(msg, file=sys.stderr, flush=True)


def make_sim(fa, scene, args):
    I, J, K = args.size
    sim = fa.FluidSimulation(I, J, K, scene.dx)
    sim.addBodyForce(0.0, -9.81, 0.0)
    sim.setMinTimeStepsPerFrame(1)
    sim.setMaxTimeStepsPerFrame(1)
    sim.setPressureSolverPrecision(fa.PCG_MIXED if args.precision == "mixed" else fa.PCG_FP64)
    sim.loadMarkerParticleData(scene.positions, scene.velocities)
    sim.initialize()
    return sim


def cpu_threads(args):
    if args.cpu_threads > 0:
        return args.cpu_threads
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


class _StdoutToStderr:
    """Sends file descriptor 1 to stderr while the CPU engines run: the
    reference engine prints a banner and a log on stdout, and bench.py's
    stdout must hold only the JSON line."""

    def __enter__(self):
        import ctypes
        sys.stdout.flush()
        self._libc = ctypes.CDLL(None)
        self._libc.fflush(None)
        self._saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        self._libc.fflush(None)
        os.dup2(self._saved, 1)
        os.close(self._saved)
        return False


def cpu_baseline(scene, args, substeps=(1, 1), surface=False):
    """Times the reference engine (or, where it was not built, the CPU
    oracle) on the same scene for whole updates until --cpu-seconds pass."""
    with _StdoutToStderr():
        return _cpu_baseline(scene, args, substeps, surface)


def _cpu_baseline(scene, args, substeps=(1, 1), surface=False):
    from oracle import bindings
    I, J, K = args.size
    if bindings.ref_available():
        threads = cpu_threads(args)
        lib = bindings.Reference(threads=threads)
        kind = "reference"
    else:
        threads = 1
        lib = bindings.Oracle()
        kind = "port"
    h = lib.sim(I, J, K, scene.dx)
    h.add_body_force(0.0, -9.81, 0.0)
    h.set_substeps(*substeps)
    h.load_particles(scene.positions, scene.velocities)
    if surface:
        h.set_surface(True)
    h.initialize()
    done = 0
    work = 0
    elapsed = 0.0
    while True:
        n = h.num_particles
        t0 = time.perf_counter()
        h.update(args.dt)
        elapsed += time.perf_counter() - t0
        work += n * h.substeps()
        done += 1
        log(f"cpu update {done}: {n} particles, {elapsed:.1f} s total", args)
        if elapsed >= args.cpu_seconds or done >= 8:
            break
    h.close()
    return {
        "value": work / elapsed,
        "unit": "particle-substeps/s",
        "cores": threads,
        "kind": kind,
        "sample": (f"{done} update(s) of dt={args.dt:.6g} s of the same {I}x{J}x{K} {scene.name.replace('_', ' ')} "
                   f"({scene.num_particles} particles), FP64 PCG, {threads} thread(s), {elapsed:.1f} s"
                   + (", surface reconstruction on" if surface else "")),
    }



def run_gpu(fa, scene, args, precision, steps, warmup, profiling):
    """Warmup + timed substeps of one scene; returns throughput, latency
    percentiles, the PCG bandwidth and the dominant kernel's roofline."""
    a = argparse.Namespace(**vars(args))
    a.precision = precision
    sim = make_sim(fa, scene, a)
    sim.set_profiling(profiling)
    n = sim.getNumMarkerParticles()
    for w in range(warmup):
        st = sim.update(args.dt)
        log(f"{precision} warmup {w}: {st.pcg_iterations} PCG iterations, {st.gpu_ms:.1f} ms", args)
        n = st.num_particles
    kern = None
    if profiling >= 2:
        # The warmup timed every PCG kernel; the timed region times only the
        # dominant one, on every KERNEL_SAMPLE-th iteration (each timed launch
        # costs an event pair, which delays the launches around it).
        wt = sim.stage_times()
        kern = max(PCG_BYTES[precision], key=lambda k: wt[k][0])
        sim.set_kernel_timing([kern], every=KERNEL_SAMPLE)
    sim.stage_times(reset=True)

    lat = []
    work = 0
    rows_iters = 0
    iters = 0
    refinements = 0
    ok = True
    sim.ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        st = sim.update(args.dt)
        lat.append(time.perf_counter() - ts)
        work += n * st.substeps
        n = st.num_particles
        rows_iters += st.pressure_rows * st.pcg_iterations
        iters += st.pcg_iterations
        refinements += getattr(st, "pcg_refinements", 0)
        ok = ok and st.pressure_ok == 1
    sim.ctx.synchronize()
    elapsed = time.perf_counter() - t0
    times = sim.stage_times()
    sim.close()

    per_row = PCG_BYTES[precision]
    rows_per_launch = rows_iters / max(iters, 1)
    # PCG bandwidth: algorithmic bytes of every iteration (PCG_BYTES per
    # row) plus those of each true residual, over the PCG stage time.
    pcg_bytes = rows_iters * sum(per_row.values()) + refinements * rows_per_launch * TRUE_RESIDUAL_BYTES
    pcg_ms = times["pcg"][0]
    roof = {"kernel": None, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": None, "traffic": None}
    if kern is not None:
        kms, kcalls = times[kern]
        bytes_per_launch = per_row[kern] * rows_per_launch
        # Iterations are enqueued in batches, so a few launches after
        # convergence return at once; the average is over the launches that
        # iterated (one in KERNEL_SAMPLE is timed).
        kcalls = min(kcalls, -(-iters // KERNEL_SAMPLE))
        launch_s = kms * 1e-3 / max(kcalls, 1)
        achieved = bytes_per_launch / launch_s / 1e9 if launch_s > 0 else 0.0
        traffic = MEASURED_TRAFFIC.get((precision, tuple(args.size)), {}).get(kern)
        roof.update({
            "kernel": f"k_sweep<{'forward' if kern == 'k_forward' else 'backward'}>"
            if kern in ("k_forward", "k_backward") else kern,
            "achieved": achieved,
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_us": launch_s * 1e6,
            "launches": kcalls,
            "timed_every_nth_iteration": KERNEL_SAMPLE,
        })
    lat_ms = sorted(x * 1e3 for x in lat)
    p99 = lat_ms[min(len(lat_ms) - 1, int(round(0.99 * (len(lat_ms) - 1))))]
    return {
        "value": work / elapsed,
        "ms_per_step": elapsed * 1e3 / steps,
        "p50_ms": statistics.median(lat_ms),
        "p99_ms": p99,
        "pcg_hbm_gbs": pcg_bytes / (pcg_ms * 1e-3) / 1e9 if pcg_ms > 0 else None,
        "iters_per_step": iters / steps,
        "refinements_per_step": refinements / steps,
        "rows": round(rows_per_launch),
        "ok": ok,
        "roofline": roof,
        "stages": {k: round(v[0] / steps, 3) for k, v in times.items() if k != "pcg_iters"},
    }

def copy_calibration():
    """A plain device copy of known size: reads and writes 256 MiB per
    launch.  Its FETCH_SIZE / WRITE_SIZE readings give the scale from those
    counters to bytes."""
    import torch
    n = 256 << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.zeros(n, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 3}), flush=True)
    return 0


# The other workloads: (grid, dx, scene builder, PCG precision, tolerance,
# substep bounds, dt per update, updates).
WORKLOADS = {
    # 64^3 dam break: cuboid [0,0.4]x[0,0.5]x[0,1], 30 fps frames whose substeps
    # follow the CFL condition (1..6), 60 frames, FP64 PCG, seed 1.
    "dam_break_64": ((64, 64, 64), 1 / 64, lambda sc: sc.dam_break(64, 64, 64, 1 / 64, seed=1), "fp64", 1e-9,
                     (1, 6), 1 / 30, 60),
    # 128^3 drop into a pool: bottom 30 % plus a sphere of radius 0.15 at
    # (0.5, 0.7, 0.5), 200 substeps of 1/120 s, FP64 PCG, seed 2.
    "pool_drop_128": ((128, 128, 128), 1 / 128, lambda sc: sc.pool_drop(128, 128, 128, 1 / 128, seed=2), "fp64",
                      1e-9, (1, 1), 1 / 120, 200),
    # 512x256x256 pool, 25 % deep, with an inflow jet (cylinder of radius
    # 0.05 m along +x over x in [0.1, 0.3] m at height 0.6 m, mid depth,
    # 2 m/s), mixed PCG at 1e-6 relative, 10 substeps of 1/480 s, seed 5.
    "large_pool_512": ((512, 256, 256), 1 / 256, lambda sc: sc.pool_jet(512, 256, 256, 1 / 256, seed=5),
                       "mixed", 1e-6, (1, 1), 1 / 480, 10),
    # The headline's 256x128x128 double dam break (seed 3, mixed PCG, one
    # substep of 1/240 s per update) with surface reconstruction on, at the
    # reference's default settings: getIsomesh's mesh is built every frame.
    "double_dam_surface_256": ((256, 128, 128), 1 / 128,
                               lambda sc: sc.double_dam_break(256, 128, 128, 1 / 128, seed=3), "mixed", 1e-9,
                               (1, 1), 1 / 240, 20, {"surface": True}),
}


def run_workload(args):
    """One of WORKLOADS on the GPU (particle-substeps/s, step latency, PCG
    iterations) with the reference engine timed on a bounded sample of it."""
    import flipengine3d_amd as fa
    from flipengine3d_amd import scenes
    (I, J, K), dx, build, precision, tol, substeps, dt, updates, *extra = WORKLOADS[args.workload]
    surface = bool(extra and extra[0].get("surface"))
    t = time.perf_counter()
    scene = build(scenes)
    inflow = None
    if isinstance(scene, tuple):
        scene, inflow = scene
    log(f"{args.workload}: {scene.num_particles} particles ({time.perf_counter() - t:.1f} s)", args)
    sim = fa.FluidSimulation(I, J, K, dx, precision=fa.PCG_MIXED if precision == "mixed" else fa.PCG_FP64)
    sim.addBodyForce(0.0, -9.81, 0.0)
    sim.setMinTimeStepsPerFrame(substeps[0])
    sim.setMaxTimeStepsPerFrame(substeps[1])
    sim.setPressureSolverTolerance(tol)
    sim.loadMarkerParticleData(scene.positions, scene.velocities)
    if inflow is not None:
        sim.addMeshFluidSource(inflow.vertices, inflow.triangles, inflow.velocity)
    if surface:
        sim.enableSurfaceReconstruction()
    sim.initialize()
    if surface:
        sim.set_profiling(1)  # stage events: the surface build's time per frame
    lat, work, nsub, iters, failures = [], 0, 0, 0, 0
    n = sim.getNumMarkerParticles()
    sim.ctx.synchronize()
    t0 = time.perf_counter()
    for q in range(updates):
        ts = time.perf_counter()
        st = sim.update(dt)
        lat.append((time.perf_counter() - ts) * 1e3 / max(st.substeps, 1))
        work += n * st.substeps
        nsub += st.substeps
        iters += st.pcg_iterations
        failures += 0 if st.pressure_ok else 1
        n = st.num_particles
        if (q + 1) % 20 == 0:
            log(f"{args.workload}: {q + 1} updates, {time.perf_counter() - t0:.1f} s", args)
    elapsed = time.perf_counter() - t0
    mesh = None
    if surface:
        v, t = sim.getIsomesh()
        ms, calls = sim.stage_times()["surface"]
        mesh = {"surface_ms_per_frame": ms / max(calls, 1), "frames_meshed": calls, "vertices": int(v.shape[0]),
                "triangles": int(t.shape[0])}
    sim.close()
    lat.sort()
    out = {
        "metric": f"particle-substeps/s ({args.workload})",
        "value": work / elapsed,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "ms_per_substep": elapsed * 1e3 / nsub,
        "p50_ms": lat[len(lat) // 2],
        "p99_ms": lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))],
        "dtype": "f32 particles/grids, f64 PCG" if precision == "fp64" else
                 "f32 particles/grids and PCG, f64 dots/x/true residual",
        "data": "synthetic",
        "config": {"workload": args.workload, "grid": [I, J, K], "dx": dx, "particles": scene.num_particles,
                   "precision": precision, "tolerance": tol, "updates": updates, "substeps": nsub, "dt": dt,
                   "pcg_iterations_last_substep_mean": iters / updates, "failed_pressure_solves": failures,
                   "inflow": None if inflow is None else {"velocity": list(inflow.velocity),
                                                         "triangles": int(inflow.triangles.shape[0])},
                   "particles_end": n, "surface": mesh},
    }
    if args.cpu_baseline and scene.num_particles <= 20_000_000:
        a = argparse.Namespace(**vars(args))
        a.size, a.dt = [I, J, K], dt
        out["cpu_baseline"] = cpu_baseline(scene, a, substeps=substeps, surface=surface)
        out["vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    else:
        out["cpu_baseline"] = None
    return out


SERVING_GRIDS = [(32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96), (128, 64, 64)]


def serving_requests(args):
    """The serving workload: request r draws its grid uniformly from
    SERVING_GRIDS (dx = 1 / largest dimension), a dam-break cuboid whose x
    extent and height are U(0.2, 0.6) fractions of the domain (full z), 8
    jittered particles per cell seeded with 4 * 2^32 + r, and arrives after
    exponential gaps of mean 1 / rate (seed 4)."""
    import numpy as np
    from flipengine3d_amd import scenes
    rng = np.random.default_rng(4)
    t = 0.0
    out = []
    for r in range(args.requests):
        t += rng.exponential(1.0 / args.rate)
        I, J, K = SERVING_GRIDS[int(rng.integers(len(SERVING_GRIDS)))]
        fx, fy = rng.uniform(0.2, 0.6, size=2)
        dx = 1.0 / max(I, J, K)
        sc = scenes.dam_break(I, J, K, dx, seed=4 * 2**32 + r, box=(0.0, 0.0, 0.0, float(fx), float(fy), 1.0))
        out.append((t, (I, J, K), dx, sc))
    return out


def _serve(fa, ctx, reqs, args, arrivals=True):
    """Runs the requests through one server; with arrivals=False every
    request is submitted at once (the server's capacity).  Returns (reports,
    makespan s, server stats)."""
    import threading
    server = fa.Server(ctx, max_active=args.max_active, memory_budget=int(args.budget_gib * (1 << 30)),
                       launch_graphs=bool(args.launch_graphs))
    ids = [None] * len(reqs)

    def submitter(t0):
        for q, (ta, dims, dx, sc) in enumerate(reqs):
            delay = t0 + ta - time.perf_counter() if arrivals else 0.0
            if delay > 0:
                time.sleep(delay)
            desc = fa.scene_desc(*dims, dx, precision=fa.PCG_MIXED, substeps=(1, 1))
            ids[q] = server.submit(desc, sc.positions, sc.velocities, args.dt, args.request_steps)

    t0 = time.perf_counter()
    th = threading.Thread(target=submitter, args=(t0,))
    th.start()
    reps = []
    for q in range(len(reqs)):
        while ids[q] is None:
            time.sleep(0.001)
        rep = server.wait(ids[q])
        server.discard(ids[q])
        reps.append(rep)
        if (q + 1) % 32 == 0:
            log(f"serving: {q + 1} requests done, {time.perf_counter() - t0:.1f} s", args)
    th.join()
    makespan = time.perf_counter() - t0
    st = server.stats()
    server.close()
    return reps, makespan, st


def _plain(fa, ctx, reqs, args, precision):
    """The requests one at a time, a fresh FluidSimulation each; the clock
    covers scene creation, loading, initialize and every update, as the
    server's makespan does."""
    work, elapsed, steps = 0, 0.0, []
    for ta, dims, dx, sc in reqs:
        ts = time.perf_counter()
        sim = fa.FluidSimulation(*dims, dx, ctx=ctx, precision=precision)
        sim.addBodyForce(0.0, -9.81, 0.0)
        sim.setMinTimeStepsPerFrame(1)
        sim.setMaxTimeStepsPerFrame(1)
        sim.loadMarkerParticleData(sc.positions, sc.velocities)
        sim.initialize()
        for _ in range(args.request_steps):
            n = sim.getNumMarkerParticles()
            t1 = time.perf_counter()
            stp = sim.update(args.dt)
            steps.append((time.perf_counter() - t1) * 1e3)
            work += n * stp.substeps
        sim.close()
        elapsed += time.perf_counter() - ts
    steps.sort()
    return {"value": work / elapsed if elapsed > 0 else None, "requests": len(reqs),
            "p50_ms": steps[len(steps) // 2] if steps else None,
            "p99_ms": steps[min(len(steps) - 1, int(round(0.99 * (len(steps) - 1))))] if steps else None}


def run_serving(args):
    """Serving: requests arrive on their schedule and are admitted by the
    server (continuous batching, memory budget); value = particle-substeps of
    all requests over the makespan.  The same requests submitted all at once
    give the server's capacity.  The plain path (one scene at a time, a fresh
    FluidSimulation each, creation inside the clock as in the server's
    makespan) runs a prefix of the same requests in the same mixed precision,
    and in FP64."""
    if args.hw_queues:
        # Streams beyond the hardware queues share one; read at device init.
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(1, args.hw_queues)))
    import flipengine3d_amd as fa
    t = time.perf_counter()
    reqs = serving_requests(args)
    log(f"serving: {len(reqs)} requests built in {time.perf_counter() - t:.1f} s", args)
    ctx = fa.Context(0)
    reps, makespan, st = _serve(fa, ctx, reqs, args, arrivals=True)
    creps, cmakespan, cst = _serve(fa, ctx, reqs, args, arrivals=False)
    work = sum(r["particle_substeps"] for r in reps)
    cwork = sum(r["particle_substeps"] for r in creps)
    failed = sum(1 for r in reps if r["status"] != 0) + sum(1 for r in creps if r["status"] != 0)
    prefix = reqs[:args.plain_requests]
    plain_mixed = _plain(fa, ctx, prefix, args, fa.PCG_MIXED)
    plain_fp64 = _plain(fa, ctx, prefix, args, fa.PCG_FP64)
    value = work / makespan
    capacity = cwork / cmakespan
    nps = [sc.num_particles for _, _, _, sc in reqs]
    return {
        "metric": "serving particle-substeps/s (mixed-size scenes, Poisson arrivals); p50/p99 step latency",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "p50_ms": st["step_ms_p50"],
        "p99_ms": st["step_ms_p99"],
        "dtype": "f32 particles/grids and PCG, f64 dots/x/true residual",
        "data": "synthetic",
        "config": {
            "workload": (f"{len(reqs)} dam-break requests, grids {SERVING_GRIDS}, fill U(0.2,0.6) in x and height, "
                         f"{args.request_steps} substeps of dt={args.dt:.6g} s each, Poisson arrivals at "
                         f"{args.rate:g}/s, seed 4"),
            "particles_mean": sum(nps) / len(nps),
            "max_active": args.max_active,
            "launch_graphs": bool(args.launch_graphs),
            "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "default"),
            "budget_gib": args.budget_gib,
            "makespan_s": makespan,
            "offered_particle_substeps_per_s": sum(nps) * args.request_steps / reqs[-1][0],
            "failed_requests": failed,
            "pressure_failures": sum(r["pressure_failures"] for r in reps),
            "scenes_created": st["scenes_created"],
            "scenes_reused": st["scenes_reused"],
            "scenes_evicted": st["scenes_evicted"],
            "peak_resident_gib": st["peak_resident_bytes"] / (1 << 30),
            "queue_ms_p50": statistics.median(r["queue_ms"] for r in reps),
        },
        "capacity": {"value": capacity, "makespan_s": cmakespan, "p50_ms": cst["step_ms_p50"],
                     "p99_ms": cst["step_ms_p99"], "scenes_created": cst["scenes_created"],
                     "scenes_reused": cst["scenes_reused"]},
        "plain_mixed_path": plain_mixed,
        "plain_fp64_path": plain_fp64,
        "vs_plain": capacity / plain_mixed["value"] if plain_mixed["value"] else None,
    }


def main(argv=None):
    args = parse_args(argv)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the FLIP step runs one scene per device and this build has no "
              "multi-GPU path (a single scene would shard as x-slabs with halo exchange; serving shards "
              "whole scenes). Run with --gpus 1.", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration()
    if args.serving:
        print(json.dumps(run_serving(args)), flush=True)
        return 0
    if args.workload:
        print(json.dumps(run_workload(args)), flush=True)
        return 0

    import flipengine3d_amd as fa
    from flipengine3d_amd import scenes

    I, J, K = args.size
    dx = 1.0 / 128.0 if (I, J, K) == (256, 128, 128) else 1.0 / max(J, K)
    t = time.perf_counter()
    scene = scenes.double_dam_break(I, J, K, dx, seed=args.seed)
    log(f"scene: {scene.num_particles} particles ({time.perf_counter() - t:.1f} s)", args)

    primary = run_gpu(fa, scene, args, args.precision, args.steps, args.warmup, args.profiling)
    result = {
        "metric": "particle-substeps/s @256x128x128 + PCG HBM GB/s; p50/p99 step latency",
        "value": primary["value"],
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": primary["ms_per_step"],
        "p50_ms": primary["p50_ms"],
        "p99_ms": primary["p99_ms"],
        "pcg_hbm_gbs": primary["pcg_hbm_gbs"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 particles/grids, f64 PCG" if args.precision == "fp64" else
                 "f32 particles/grids and PCG, f64 dots/x/true residual",
        "data": "synthetic",
        "config": {
            "workload": f"double dam break {I}x{J}x{K}, dx=1/{round(1 / dx)}, seed {args.seed}, "
                        f"one substep of dt={args.dt:.6g} s per step",
            "particles": scene.num_particles,
            "pcg_precision": args.precision,
            "pcg_iterations_per_step": primary["iters_per_step"],
            "pcg_refinements_per_step": primary["refinements_per_step"],
            "pressure_rows": primary["rows"],
            "pressure_ok": primary["ok"],
            "profiling_level": args.profiling,
        },
        "roofline": primary["roofline"],
        "stage_ms_per_step": primary["stages"],
    }
    if args.fp64_path and args.precision != "fp64":
        # The reference-arithmetic FP64 PCG path on the same scene, for comparison.
        alt = run_gpu(fa, scene, args, "fp64", max(3, args.steps // 4), 1, args.profiling)
        result["config"]["fp64_path"] = {k: alt[k] for k in ("value", "ms_per_step", "p50_ms", "iters_per_step",
                                                              "pcg_hbm_gbs")}
        result["config"]["fp64_path"]["roofline_frac"] = alt["roofline"]["frac"]
    if args.cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(scene, args)
        result["config"]["gpu_over_cpu"] = result["value"] / result["cpu_baseline"]["value"]
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
