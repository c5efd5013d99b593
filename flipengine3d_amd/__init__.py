"""flipengine3d_amd -- MI355X-native FLIP fluid step.

The compute path is the HIP library ``libflipmi.so`` (built in-tree by
``flipengine3d_amd/build.py``) behind the C ABI in ``include/flipmi.h``.  This
module is a thin ctypes binding used by the tests and the benchmark; it
mirrors the reference engine's interface names (FluidSimulation,
PressureSolver, VelocityAdvector, ParticleAdvector) and raises the
reference's exception types on the same errors.  There is no CPU fallback:
loading fails loudly when the library is missing.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence, Tuple

import numpy as np

from . import scenes  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflipmi.so")

FLIPMI_OK = 0
FLIPMI_ERR_INVALID_ARG = -1
FLIPMI_ERR_NOT_INITIALIZED = -2
FLIPMI_ERR_OOM = -3
FLIPMI_ERR_NOT_CONVERGED = -4
FLIPMI_ERR_HIP = -5
FLIPMI_ERR_BUDGET = -6
FLIPMI_ERR_TIMEOUT = -7

PCG_FP64 = 0
PCG_MIXED = 1

STAGES = ["sort", "classify", "p2g", "extrapolate", "assemble", "mic0", "pcg", "apply",
          "g2p_advect", "remove", "glue", "pcg_iters", "k_spmv_dot", "k_axpy_norm", "k_forward",
          "k_backward", "k_update_s", "k_true_residual", "inflow", "surface"]
NUM_STAGES = 24

_c_int = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_double = ctypes.c_double
_vp = ctypes.c_void_p


class SceneDesc(ctypes.Structure):
    _fields_ = [("isize", _c_int), ("jsize", _c_int), ("ksize", _c_int), ("dx", _c_double),
                ("body_force", _c_double * 3), ("cfl", _c_double), ("min_substeps", _c_int),
                ("max_substeps", _c_int), ("pic_ratio", _c_double), ("pcg_tolerance", _c_double),
                ("pcg_max_iterations", _c_int), ("pcg_precision", _c_int),
                ("max_particles", _c_i64)]


class FrameStats(ctypes.Structure):
    _fields_ = [("substeps", _c_int), ("num_particles", _c_i64), ("pcg_iterations", _c_int),
                ("pcg_residual", _c_double), ("pressure_ok", _c_int), ("gpu_ms", _c_double),
                ("pressure_rows", _c_i64), ("pcg_refinements", _c_int)]


class StageTimes(ctypes.Structure):
    _fields_ = [("ms", _c_double * 24), ("calls", _c_i64 * 24)]


class PressureArgs(ctypes.Structure):
    _fields_ = [("isize", _c_int), ("jsize", _c_int), ("ksize", _c_int), ("dx", _c_double),
                ("dt", _c_double), ("tolerance", _c_double), ("acceptable_tolerance", _c_double),
                ("max_iterations", _c_int), ("precision", _c_int), ("liquid_phi", _vp),
                ("weight_u", _vp), ("weight_v", _vp), ("weight_w", _vp), ("weight_c", _vp),
                ("solid_u", _vp), ("solid_v", _vp), ("solid_w", _vp), ("u", _vp), ("v", _vp),
                ("w", _vp), ("valid_u", _vp), ("valid_v", _vp), ("valid_w", _vp),
                ("pressure_rows", _vp)]


class ServerDesc(ctypes.Structure):
    _fields_ = [("max_active", _c_int), ("memory_budget", _c_i64), ("particle_capacity_per_cell", _c_double),
                ("launch_graphs", _c_int)]


class Request(ctypes.Structure):
    _fields_ = [("scene", SceneDesc), ("pos_xyz", _vp), ("vel_xyz", _vp), ("n", _c_i64), ("dt", _c_double),
                ("steps", _c_int)]


class RequestReport(ctypes.Structure):
    _fields_ = [("status", _c_int), ("steps_done", _c_int), ("num_particles", _c_i64),
                ("particle_substeps", _c_i64), ("pressure_failures", _c_int), ("reused_scene", _c_int),
                ("queue_ms", _c_double), ("run_ms", _c_double), ("step_ms_p50", _c_double),
                ("step_ms_p99", _c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", _c_i64), ("completed", _c_i64), ("failed", _c_i64), ("queued", _c_i64),
                ("active", _c_i64), ("particle_substeps", _c_i64), ("steps", _c_i64),
                ("step_ms_p50", _c_double), ("step_ms_p99", _c_double), ("resident_bytes", _c_i64),
                ("peak_resident_bytes", _c_i64), ("scenes_created", _c_i64), ("scenes_reused", _c_i64),
                ("scenes_evicted", _c_i64)]


class SurfaceDesc(ctypes.Structure):
    _fields_ = [("enabled", _c_int), ("subdivisions", _c_int), ("polygonizer_slices", _c_int),
                ("smoothing_value", _c_double), ("smoothing_iterations", _c_int),
                ("min_polyhedron_triangles", _c_int), ("particle_scale", _c_double),
                ("obstacle_meshing_offset", _c_double), ("remove_near_domain", _c_int),
                ("remove_near_domain_distance", _c_int), ("invert_contact_normals", _c_int)]


class PressureReport(ctypes.Structure):
    _fields_ = [("iterations", _c_int), ("residual", _c_double), ("rows", _c_int),
                ("success", _c_int), ("refinements", _c_int)]


EXPORTED = [
    "flipmi_ctx_create", "flipmi_ctx_destroy", "flipmi_ctx_synchronize", "flipmi_last_error", "flipmi_scene_desc_default",
    "flipmi_scene_create", "flipmi_scene_destroy", "flipmi_scene_load_particles",
    "flipmi_scene_initialize", "flipmi_scene_update", "flipmi_scene_num_particles",
    "flipmi_scene_get_particles", "flipmi_scene_get_particles_range", "flipmi_scene_set_domain",
    "flipmi_scene_add_inflow_source", "flipmi_scene_set_inflow_source", "flipmi_scene_set_marker_jitter",
    "flipmi_scene_get_inflow_source", "flipmi_scene_add_obstacle", "flipmi_scene_set_obstacle",
    "flipmi_scene_set_outflow_source", "flipmi_scene_set_obstacle_inversed",
    "flipmi_scene_get_grid", "flipmi_scene_stage_times",
    "flipmi_scene_set_profiling", "flipmi_scene_set_kernel_timing", "flipmi_scene_set_kernel_sampling",
    "flipmi_domain_solid", "flipmi_liquid_sdf",
    "flipmi_liquid_sdf_postprocess", "flipmi_p2g", "flipmi_removal_counts",
    "flipmi_extrapolate", "flipmi_pressure_solve", "flipmi_mic0_apply", "flipmi_advect_particles",
    "flipmi_trilinear_velocity", "flipmi_server_desc_default", "flipmi_server_create", "flipmi_server_submit",
    "flipmi_server_wait", "flipmi_server_take", "flipmi_server_get_stats", "flipmi_server_destroy",
    "flipmi_extrapolate_noalloc", "flipmi_surface_desc_default", "flipmi_scene_set_surface",
    "flipmi_scene_get_isomesh", "flipmi_mesh_particles", "flipmi_mesh_counts", "flipmi_mesh_copy",
    "flipmi_mesh_destroy", "flipmi_scene_set_launch_graphs",
]

_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads libflipmi.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run flipengine3d_amd/build.py (no CPU fallback exists)")
    L = ctypes.CDLL(path)
    P = ctypes.POINTER
    L.flipmi_ctx_create.argtypes = [_c_int, P(_vp)]
    L.flipmi_ctx_destroy.argtypes = [_vp]
    L.flipmi_ctx_synchronize.argtypes = [_vp]
    L.flipmi_last_error.argtypes = [_vp]
    L.flipmi_last_error.restype = ctypes.c_char_p
    L.flipmi_scene_desc_default.argtypes = [P(SceneDesc)]
    L.flipmi_scene_create.argtypes = [_vp, P(SceneDesc), P(_vp)]
    L.flipmi_scene_destroy.argtypes = [_vp]
    L.flipmi_scene_load_particles.argtypes = [_vp, _vp, _vp, _c_i64]
    L.flipmi_scene_initialize.argtypes = [_vp]
    L.flipmi_scene_update.argtypes = [_vp, _c_double, P(FrameStats)]
    L.flipmi_scene_num_particles.argtypes = [_vp, P(_c_i64)]
    L.flipmi_scene_get_particles.argtypes = [_vp, _vp, _vp]
    L.flipmi_scene_get_particles_range.argtypes = [_vp, _c_i64, _c_i64, _vp, _vp]
    L.flipmi_scene_set_domain.argtypes = [_vp, _c_double, _c_double, _c_double, _c_double]
    L.flipmi_scene_add_inflow_source.argtypes = [_vp, _vp, _c_int, _vp, _c_int, _vp, P(_c_int)]
    L.flipmi_scene_set_inflow_source.argtypes = [_vp, _c_int, _c_int, _c_int, _c_int]
    L.flipmi_scene_set_marker_jitter.argtypes = [_vp, _c_double, _c_int, ctypes.c_uint64]
    L.flipmi_scene_get_inflow_source.argtypes = [_vp, _c_int, _vp, _vp, _vp]
    L.flipmi_scene_add_obstacle.argtypes = [_vp, _vp, _c_int, _vp, _c_int, P(_c_int)]
    L.flipmi_scene_set_obstacle.argtypes = [_vp, _c_int, _c_int]
    L.flipmi_scene_set_outflow_source.argtypes = [_vp, _c_int, _c_int, _c_int]
    L.flipmi_scene_set_obstacle_inversed.argtypes = [_vp, _c_int, _c_int]
    L.flipmi_scene_get_grid.argtypes = [_vp, _c_int, _vp]
    L.flipmi_scene_stage_times.argtypes = [_vp, P(StageTimes), _c_int]
    L.flipmi_scene_set_profiling.argtypes = [_vp, _c_int]
    L.flipmi_scene_set_kernel_timing.argtypes = [_vp, ctypes.c_uint]
    L.flipmi_scene_set_kernel_sampling.argtypes = [_vp, _c_int]
    L.flipmi_scene_set_launch_graphs.argtypes = [_vp, _c_int]
    L.flipmi_domain_solid.argtypes = [_vp, _c_int, _c_int, _c_int, _c_double] + [_vp] * 5
    L.flipmi_liquid_sdf.argtypes = [_vp, _c_int, _c_int, _c_int, _c_double, _vp, _c_i64, _c_double, _vp, _vp]
    L.flipmi_liquid_sdf_postprocess.argtypes = [_vp, _c_int, _c_int, _c_int, _c_double, _vp, _vp]
    L.flipmi_removal_counts.argtypes = [_vp, _c_int, _c_int, _c_int, _c_double, _vp, _c_i64, _vp, _vp]
    L.flipmi_p2g.argtypes = [_vp, _c_int, _c_int, _c_int, _c_double, _vp, _vp, _c_i64, _c_double] + [_vp] * 6
    L.flipmi_extrapolate.argtypes = [_vp, _c_int, _c_int, _c_int, _vp, _vp, _c_int]
    L.flipmi_extrapolate_noalloc.argtypes = [_vp, _c_int, _c_int, _c_int, _vp, _vp, _c_int, _vp]
    L.flipmi_pressure_solve.argtypes = [_vp, P(PressureArgs), P(PressureReport)]
    L.flipmi_mic0_apply.argtypes = [_vp, _c_int, _c_int, _c_int, _c_double, _c_double] + [_vp] * 8
    L.flipmi_advect_particles.argtypes = ([_vp, _c_int, _c_int, _c_int, _c_double, _vp, _vp, _vp, _vp, _c_i64,
                                          _c_double, _c_int, _vp])
    L.flipmi_trilinear_velocity.argtypes = ([_vp, _c_int, _c_int, _c_int, _c_double, _vp, _vp, _vp, _vp,
                                            _c_i64, _vp])
    L.flipmi_surface_desc_default.argtypes = [P(SurfaceDesc)]
    L.flipmi_surface_desc_default.restype = None
    L.flipmi_scene_set_surface.argtypes = [_vp, P(SurfaceDesc)]
    L.flipmi_scene_get_isomesh.argtypes = [_vp, P(_c_i64), P(_c_i64), _vp, _vp]
    L.flipmi_mesh_particles.argtypes = ([_vp, _c_int, _c_int, _c_int, _c_double, _vp, _c_i64, _vp, P(SurfaceDesc),
                                         _c_double, _vp, _c_double, P(_vp)])
    L.flipmi_mesh_counts.argtypes = [_vp, P(_c_i64), P(_c_i64)]
    L.flipmi_mesh_copy.argtypes = [_vp, _vp, _vp]
    L.flipmi_mesh_destroy.argtypes = [_vp]
    L.flipmi_server_desc_default.argtypes = [P(ServerDesc)]
    L.flipmi_server_desc_default.restype = None
    L.flipmi_server_create.argtypes = [_vp, P(ServerDesc), P(_vp)]
    L.flipmi_server_submit.argtypes = [_vp, P(Request), P(_c_i64)]
    L.flipmi_server_wait.argtypes = [_vp, _c_i64, _c_double, P(RequestReport)]
    L.flipmi_server_take.argtypes = [_vp, _c_i64, _vp, _vp]
    L.flipmi_server_get_stats.argtypes = [_vp, P(ServerStats)]
    L.flipmi_server_destroy.argtypes = [_vp]
    _lib = L
    return L


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


class FlipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"flipmi error {code}: {msg}")
        self.code = code


class Context:
    """One device context (flipmi_ctx)."""

    def __init__(self, device: int = 0):
        self.L = load_library()
        h = _vp()
        rc = self.L.flipmi_ctx_create(device, ctypes.byref(h))
        if rc != FLIPMI_OK:
            raise FlipError(rc, "flipmi_ctx_create failed (no usable HIP device?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.flipmi_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self.check(self.L.flipmi_ctx_synchronize(self.h))

    def last_error(self) -> str:
        m = self.L.flipmi_last_error(self.h)
        return m.decode() if m else ""

    def check(self, rc: int):
        if rc == FLIPMI_OK:
            return
        msg = self.last_error()
        if rc == FLIPMI_ERR_NOT_INITIALIZED:
            raise RuntimeError(msg)
        if rc == FLIPMI_ERR_INVALID_ARG:
            raise ValueError(msg)
        raise FlipError(rc, msg)

    # ---- stage drop-ins --------------------------------------------------
    def domain_solid(self, isize, jsize, ksize, dx):
        phi = np.zeros((ksize + 1, jsize + 1, isize + 1), np.float32)
        wu = np.zeros((ksize, jsize, isize + 1), np.float32)
        wv = np.zeros((ksize, jsize + 1, isize), np.float32)
        ww = np.zeros((ksize + 1, jsize, isize), np.float32)
        wc = np.zeros((ksize, jsize, isize), np.float32)
        self.check(self.L.flipmi_domain_solid(self.h, isize, jsize, ksize, dx, _ptr(phi), _ptr(wu), _ptr(wv),
                                              _ptr(ww), _ptr(wc)))
        return phi, (wu, wv, ww, wc)

    def liquid_sdf(self, isize, jsize, ksize, dx, pos, radius, solid_phi=None):
        pos = np.ascontiguousarray(pos, np.float32)
        sp = None if solid_phi is None else np.ascontiguousarray(solid_phi, np.float32)
        phi = np.zeros((ksize, jsize, isize), np.float32)
        self.check(self.L.flipmi_liquid_sdf(self.h, isize, jsize, ksize, dx, _ptr(pos), pos.shape[0], radius,
                                            _ptr(sp), _ptr(phi)))
        return phi

    def removal_counts(self, isize, jsize, ksize, dx, pos, removed=None):
        """Count grid of the removal pass's counting loop (int32 K x J x I)."""
        pos = np.ascontiguousarray(pos, np.float32)
        rm = None if removed is None else np.ascontiguousarray(removed, np.uint8)
        out = np.zeros((ksize, jsize, isize), np.int32)
        self.check(self.L.flipmi_removal_counts(self.h, isize, jsize, ksize, dx, _ptr(pos), pos.shape[0],
                                                _ptr(rm), _ptr(out)))
        return out

    def surface_desc(self, **settings) -> SurfaceDesc:
        """The reference's surface settings (enabled), with overrides."""
        d = SurfaceDesc()
        self.L.flipmi_surface_desc_default(ctypes.byref(d))
        for k, v in settings.items():
            setattr(d, k, v)
        return d

    def mesh_particles(self, isize, jsize, ksize, dx, pos, solid_phi, kernel_radius=0.0, domain_offset=None,
                       domain_scale=1.0, **settings):
        """(vertices (n, 3) float32, triangles (m, 3) int32): the surface of
        `pos` (grid coordinates) over the nodal solid SDF, as the simulation
        builds it; settings are SurfaceDesc fields."""
        pos = np.ascontiguousarray(pos, np.float32)
        sp = np.ascontiguousarray(solid_phi, np.float32)
        d = self.surface_desc(**settings)
        off = None if domain_offset is None else np.ascontiguousarray(domain_offset, np.float64)
        m = _vp()
        self.check(self.L.flipmi_mesh_particles(self.h, isize, jsize, ksize, dx, _ptr(pos), pos.shape[0], _ptr(sp),
                                                ctypes.byref(d), kernel_radius, _ptr(off), domain_scale,
                                                ctypes.byref(m)))
        try:
            nv, nt = _c_i64(0), _c_i64(0)
            self.check(self.L.flipmi_mesh_counts(m, ctypes.byref(nv), ctypes.byref(nt)))
            verts = np.zeros((nv.value, 3), np.float32)
            tris = np.zeros((nt.value, 3), np.int32)
            self.check(self.L.flipmi_mesh_copy(m, _ptr(verts), _ptr(tris)))
        finally:
            self.L.flipmi_mesh_destroy(m)
        return verts, tris

    def liquid_sdf_postprocess(self, isize, jsize, ksize, dx, phi, solid_phi):
        out = np.ascontiguousarray(phi, np.float32).copy()
        sp = np.ascontiguousarray(solid_phi, np.float32)
        self.check(self.L.flipmi_liquid_sdf_postprocess(self.h, isize, jsize, ksize, dx, _ptr(sp), _ptr(out)))
        return out

    def p2g(self, isize, jsize, ksize, dx, pos, vel, radius):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        u = np.zeros((ksize, jsize, isize + 1), np.float32)
        v = np.zeros((ksize, jsize + 1, isize), np.float32)
        w = np.zeros((ksize + 1, jsize, isize), np.float32)
        vu, vv, vw = (np.zeros(a.shape, np.uint8) for a in (u, v, w))
        self.check(self.L.flipmi_p2g(self.h, isize, jsize, ksize, dx, _ptr(pos), _ptr(vel), pos.shape[0], radius,
                                     _ptr(u), _ptr(v), _ptr(w), _ptr(vu), _ptr(vv), _ptr(vw)))
        return (u, v, w), (vu, vv, vw)

    def extrapolate(self, grid, valid, layers):
        g = np.ascontiguousarray(grid, np.float32).copy()
        vd = np.ascontiguousarray(valid, np.uint8)
        nk, nj, ni = g.shape
        self.check(self.L.flipmi_extrapolate(self.h, ni, nj, nk, _ptr(g), _ptr(vd), layers))
        return g

    def pressure_solve(self, isize, jsize, ksize, dx, dt, liquid_phi, weights, mac, valid, tol=1e-9,
                       maxiter=1000, precision=PCG_FP64, acceptable=1.0, solid_vel=None):
        """solid_vel: optional (U, V, W) solid face velocities; returned in
        'solid_vel' after the pocket conditioning."""
        lp = np.ascontiguousarray(liquid_phi, np.float32)
        sv = (None, None, None) if solid_vel is None else [np.ascontiguousarray(a, np.float32).copy()
                                                           for a in solid_vel]
        wu, wv, ww, wc = [np.ascontiguousarray(a, np.float32) for a in weights]
        u, v, w = [np.ascontiguousarray(a, np.float32).copy() for a in mac]
        vu, vv, vw = [np.ascontiguousarray(a, np.uint8).copy() for a in valid]
        rows = np.zeros(isize * jsize * ksize, np.float64)
        a = PressureArgs(isize, jsize, ksize, dx, dt, tol, acceptable, maxiter, precision, _ptr(lp), _ptr(wu),
                         _ptr(wv), _ptr(ww), _ptr(wc), _ptr(sv[0]), _ptr(sv[1]), _ptr(sv[2]), _ptr(u), _ptr(v),
                         _ptr(w), _ptr(vu),
                         _ptr(vv), _ptr(vw), _ptr(rows))
        rep = PressureReport()
        rc = self.L.flipmi_pressure_solve(self.h, ctypes.byref(a), ctypes.byref(rep))
        if rc not in (FLIPMI_OK, FLIPMI_ERR_NOT_CONVERGED):
            self.check(rc)
        return dict(ok=int(rc == FLIPMI_OK), iters=rep.iterations, err=rep.residual, rows=rep.rows,
                    soln=rows[:rep.rows], mac=(u, v, w), valid=(vu, vv, vw), refinements=rep.refinements,
                    solid_vel=None if solid_vel is None else tuple(sv))

    def mic0_apply(self, isize, jsize, ksize, dx, dt, liquid_phi, weights, r):
        lp = np.ascontiguousarray(liquid_phi, np.float32)
        wu, wv, ww, wc = [np.ascontiguousarray(a, np.float32) for a in weights]
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        inv = np.zeros_like(r)
        self.check(self.L.flipmi_mic0_apply(self.h, isize, jsize, ksize, dx, dt, _ptr(lp), _ptr(wu), _ptr(wv),
                                            _ptr(ww), _ptr(wc), _ptr(r), _ptr(z), _ptr(inv)))
        return inv, z

    def advect_particles(self, isize, jsize, ksize, dx, mac, pos, dt, order):
        u, v, w = [np.ascontiguousarray(a, np.float32) for a in mac]
        pos = np.ascontiguousarray(pos, np.float32)
        out = np.zeros_like(pos)
        self.check(self.L.flipmi_advect_particles(self.h, isize, jsize, ksize, dx, _ptr(u), _ptr(v), _ptr(w),
                                                  _ptr(pos), pos.shape[0], dt, order, _ptr(out)))
        return out

    def trilinear_velocity(self, isize, jsize, ksize, dx, mac, pos):
        u, v, w = [np.ascontiguousarray(a, np.float32) for a in mac]
        pos = np.ascontiguousarray(pos, np.float32)
        out = np.zeros_like(pos)
        self.check(self.L.flipmi_trilinear_velocity(self.h, isize, jsize, ksize, dx, _ptr(u), _ptr(v), _ptr(w),
                                                    _ptr(pos), pos.shape[0], _ptr(out)))
        return out


class FluidSimulation:
    """Device-resident scene with the reference's FluidSimulation interface
    (fluidsimulation.h:127-153): construct, add forces / load particles,
    initialize(), then update(dt) per frame."""

    _GRID_SHAPES = {
        0: lambda I, J, K: ((K, J, I), np.float32),
        1: lambda I, J, K: ((K + 1, J + 1, I + 1), np.float32),
        2: lambda I, J, K: ((K, J, I + 1), np.float32),
        3: lambda I, J, K: ((K, J + 1, I), np.float32),
        4: lambda I, J, K: ((K + 1, J, I), np.float32),
        5: lambda I, J, K: ((K, J, I + 1), np.float32),
        6: lambda I, J, K: ((K, J + 1, I), np.float32),
        7: lambda I, J, K: ((K + 1, J, I), np.float32),
        8: lambda I, J, K: ((K, J, I), np.float32),
        9: lambda I, J, K: ((K, J, I + 1), np.uint8),
        10: lambda I, J, K: ((K, J + 1, I), np.uint8),
        11: lambda I, J, K: ((K + 1, J, I), np.uint8),
        12: lambda I, J, K: ((K, J, I), np.int32),
    }

    def __init__(self, isize: int, jsize: int, ksize: int, dx: float, ctx: Optional[Context] = None,
                 precision: int = PCG_FP64, max_particles: int = 0):
        self.ctx = ctx or Context()
        self.L = self.ctx.L
        self.dims = (isize, jsize, ksize)
        self.dx = dx
        d = SceneDesc()
        self.L.flipmi_scene_desc_default(ctypes.byref(d))
        d.isize, d.jsize, d.ksize, d.dx = isize, jsize, ksize, dx
        d.pcg_precision = precision
        d.max_particles = max_particles
        self.desc = d
        self.surface = SurfaceDesc()
        self.L.flipmi_surface_desc_default(ctypes.byref(self.surface))
        self.surface.enabled = 0  # a scene starts with reconstruction off
        self.h = None
        self.last_stats = FrameStats()

    def _ensure(self):
        if self.h is None:
            h = _vp()
            self.ctx.check(self.L.flipmi_scene_create(self.ctx.h, ctypes.byref(self.desc), ctypes.byref(h)))
            self.h = h
            self._push_domain()
            self._push_surface()

    def addBodyForce(self, fx: float, fy: float, fz: float):
        if self.h is not None:
            raise RuntimeError("body forces must be set before particles are loaded")
        self.desc.body_force[0] += fx
        self.desc.body_force[1] += fy
        self.desc.body_force[2] += fz

    def setCFLConditionNumber(self, c: float):
        self.desc.cfl = c

    def setMinTimeStepsPerFrame(self, n: int):
        self.desc.min_substeps = n

    def setMaxTimeStepsPerFrame(self, n: int):
        self.desc.max_substeps = n

    def setPICFLIPRatio(self, r: float):
        self.desc.pic_ratio = r

    def setPressureSolverTolerance(self, tol: float, max_iterations: int = 1000):
        self.desc.pcg_tolerance = tol
        self.desc.pcg_max_iterations = max_iterations

    def setPressureSolverPrecision(self, precision: int):
        """PCG_FP64 (the reference's arithmetic) or PCG_MIXED."""
        if self.h is not None:
            raise RuntimeError("the solver precision must be set before particles are loaded")
        self.desc.pcg_precision = precision

    def loadMarkerParticleData(self, positions: np.ndarray, velocities: Optional[np.ndarray] = None):
        self._ensure()
        pos = np.ascontiguousarray(positions, np.float32)
        vel = None if velocities is None else np.ascontiguousarray(velocities, np.float32)
        self.ctx.check(self.L.flipmi_scene_load_particles(self.h, _ptr(pos), _ptr(vel), pos.shape[0]))

    def initialize(self):
        self._ensure()
        self.ctx.check(self.L.flipmi_scene_initialize(self.h))

    def update(self, dt: float) -> FrameStats:
        if self.h is None:
            raise RuntimeError("Error: FluidSimulation must be initialized before update.")
        st = FrameStats()
        self.ctx.check(self.L.flipmi_scene_update(self.h, dt, ctypes.byref(st)))
        self.last_stats = st
        return st

    def set_profiling(self, level=True):
        self._ensure()
        lvl = int(level) if not isinstance(level, bool) else (1 if level else 0)
        self.ctx.check(self.L.flipmi_scene_set_profiling(self.h, lvl))

    def set_launch_graphs(self, enabled=True):
        """PCG iteration batches as captured HIP graphs (bit-identical results)."""
        self._ensure()
        self.ctx.check(self.L.flipmi_scene_set_launch_graphs(self.h, int(bool(enabled))))

    def set_kernel_timing(self, kernels=None, every: int = 1):
        """PCG kernels timed at profiling level 2 (names from STAGES, e.g.
        "k_backward"); None = all.  every: time them on every n-th PCG
        iteration only."""
        self._ensure()
        mask = 0xFFFFFFFF
        if kernels is not None:
            mask = 0
            for k in kernels:
                mask |= 1 << (STAGES.index(k) - STAGES.index("k_spmv_dot"))
        self.ctx.check(self.L.flipmi_scene_set_kernel_timing(self.h, mask))
        self.ctx.check(self.L.flipmi_scene_set_kernel_sampling(self.h, int(every)))

    def stage_times(self, reset: bool = False) -> dict:
        t = StageTimes()
        self.ctx.check(self.L.flipmi_scene_stage_times(self.h, ctypes.byref(t), 1 if reset else 0))
        return {name: (t.ms[q], t.calls[q]) for q, name in enumerate(STAGES)}

    def getNumMarkerParticles(self) -> int:
        n = _c_i64()
        self.ctx.check(self.L.flipmi_scene_num_particles(self.h, ctypes.byref(n)))
        return n.value

    def particles(self) -> Tuple[np.ndarray, np.ndarray]:
        """getMarkerParticlePositionData / VelocityData: positions in domain
        coordinates."""
        n = self.getNumMarkerParticles()
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        self.ctx.check(self.L.flipmi_scene_get_particles(self.h, _ptr(pos), _ptr(vel)))
        return pos, vel

    def particles_range(self, start: int, end: int) -> Tuple[np.ndarray, np.ndarray]:
        """getMarkerParticle{Position,Velocity}DataRange (fluidsimulation.cpp:
        2314-2338); an invalid range raises, as the reference throws
        domain_error."""
        m = max(end - start, 0)
        pos = np.zeros((m, 3), np.float32)
        vel = np.zeros((m, 3), np.float32)
        self.ctx.check(self.L.flipmi_scene_get_particles_range(self.h, start, end, _ptr(pos) if m else None,
                                                               _ptr(vel) if m else None))
        return pos, vel

    # ---- static mesh obstacles ----------------------------------------------
    def addMeshObstacle(self, vertices: np.ndarray, triangles: np.ndarray) -> int:
        """MeshObject with a static mesh (updateMeshStatic) added with
        addMeshObstacle (meshobject.cpp:50-59, fluidsimulation.cpp:1994-2008);
        grid coordinates.  Returns the obstacle's index."""
        self._ensure()
        v = np.ascontiguousarray(vertices, np.float32)
        t = np.ascontiguousarray(triangles, np.int32)
        oid = _c_int(-1)
        self.ctx.check(self.L.flipmi_scene_add_obstacle(self.h, _ptr(v), v.shape[0], _ptr(t), t.shape[0],
                                                        ctypes.byref(oid)))
        return oid.value

    def setMeshObstacle(self, obstacle_id: int, enabled: bool = True):
        """MeshObject::enable()/disable(); the solid is rebuilt at the next substep."""
        self._ensure()
        self.ctx.check(self.L.flipmi_scene_set_obstacle(self.h, obstacle_id, int(enabled)))

    def setMeshObstacleInversed(self, obstacle_id: int, inversed: bool = True):
        """MeshObject::inverse() state: a container, solid outside the mesh.
        Set it before the first update (it does not rebuild an existing solid)."""
        self._ensure()
        self.ctx.check(self.L.flipmi_scene_set_obstacle_inversed(self.h, obstacle_id, int(inversed)))

    # ---- mesh inflow sources ------------------------------------------------
    def addMeshFluidSource(self, vertices: np.ndarray, triangles: np.ndarray, velocity) -> int:
        """MeshFluidSource with a static mesh (updateMeshStatic), setInflow(),
        setVelocity(velocity), added with addMeshFluidSource
        (meshfluidsource.cpp:32-44, 74-76, 114-116; fluidsimulation.cpp:
        1953-1965).  vertices: (n, 3) float32 in grid coordinates; triangles:
        (m, 3) int32.  Returns the source's index."""
        self._ensure()
        v = np.ascontiguousarray(vertices, np.float32)
        t = np.ascontiguousarray(triangles, np.int32)
        vel = (ctypes.c_double * 3)(*[float(c) for c in velocity])
        sid = _c_int(-1)
        self.ctx.check(self.L.flipmi_scene_add_inflow_source(self.h, _ptr(v), v.shape[0], _ptr(t), t.shape[0],
                                                             vel, ctypes.byref(sid)))
        return sid.value

    def setMeshFluidSource(self, source_id: int, enabled: bool = True, constrain_velocity: bool = True,
                           substep_emissions: int = 1):
        """enable()/disable(), enable/disableConstrainedFluidVelocity() and
        setSubstepEmissions() of one source (meshfluidsource.cpp:51-72, 146-156)."""
        self._ensure()
        self.ctx.check(self.L.flipmi_scene_set_inflow_source(self.h, source_id, int(enabled),
                                                             int(constrain_velocity), int(substep_emissions)))

    def meshFluidSource(self, source_id: int):
        """(sub-grid offset, nodal SDF (nk, nj, ni), cells (K, J, I)) of a
        source (MeshFluidSource::getMeshLevelSet / getCells)."""
        if self.h is None:
            raise RuntimeError("Error: FluidSimulation must be initialized first.")
        d = np.zeros(6, np.int32)
        self.ctx.check(self.L.flipmi_scene_get_inflow_source(self.h, source_id, _ptr(d), None, None))
        sdf = np.zeros((d[5], d[4], d[3]), np.float32)
        cells = np.zeros((self.desc.ksize, self.desc.jsize, self.desc.isize), np.uint8)
        self.ctx.check(self.L.flipmi_scene_get_inflow_source(self.h, source_id, _ptr(d), _ptr(sdf), _ptr(cells)))
        return tuple(int(v) for v in d[:3]), sdf, cells

    def setMeshFluidSourceOutflow(self, source_id: int, outflow: bool = True, inversed: bool = False):
        """setOutflow() / setInflow() and outflowInverse() of a source
        (meshfluidsource.cpp:74-84, 158-160)."""
        self._ensure()
        self.ctx.check(self.L.flipmi_scene_set_outflow_source(self.h, source_id, int(outflow), int(inversed)))

    def setMarkerParticleJitter(self, factor: float = 1.0, jitter_surface: bool = False, seed: int = 0):
        """setMarkerParticleJitterFactor + enable/disableJitterSurfaceMarkerParticles
        (fluidsimulation.cpp:185-210) for inflow particles; `seed` keys the
        jitter draws."""
        self._ensure()
        self.ctx.check(self.L.flipmi_scene_set_marker_jitter(self.h, float(factor), int(jitter_surface),
                                                             int(seed)))

    def setDomainOffset(self, x: float, y: float, z: float):
        """fluidsimulation.cpp:307-318 (stored in float, like vmath::vec3)."""
        self._domain_offset = tuple(float(np.float32(v)) for v in (x, y, z))
        self._push_domain()

    def getDomainOffset(self):
        return getattr(self, "_domain_offset", (0.0, 0.0, 0.0))

    def setDomainScale(self, scale: float):
        """fluidsimulation.cpp:324-330."""
        self._domain_scale = float(scale)
        self._push_domain()

    def getDomainScale(self) -> float:
        return getattr(self, "_domain_scale", 1.0)

    # -- surface reconstruction (fluidsimulation.cpp:168-540) ------------------
    def _push_surface(self):
        if self.h is not None:
            self.ctx.check(self.L.flipmi_scene_set_surface(self.h, ctypes.byref(self.surface)))

    def _set_surface(self, **kv):
        old = {k: getattr(self.surface, k) for k in kv}
        for k, v in kv.items():
            setattr(self.surface, k, v)
        try:
            self._push_surface()
        except (FlipError, ValueError):
            for k, v in old.items():
                setattr(self.surface, k, v)
            raise

    def enableSurfaceReconstruction(self):
        self._set_surface(enabled=1)

    def disableSurfaceReconstruction(self):
        self._set_surface(enabled=0)

    def isSurfaceReconstructionEnabled(self) -> bool:
        return bool(self.surface.enabled)

    def setSurfaceSubdivisionLevel(self, n: int):
        self._set_surface(subdivisions=n)

    def setNumPolygonizerSlices(self, n: int):
        self._set_surface(polygonizer_slices=n)

    def setSurfaceSmoothingValue(self, s: float):
        self._set_surface(smoothing_value=s)

    def setSurfaceSmoothingIterations(self, n: int):
        self._set_surface(smoothing_iterations=n)

    def setMinPolyhedronTriangleCount(self, n: int):
        self._set_surface(min_polyhedron_triangles=n)

    def setMarkerParticleScale(self, s: float):
        self._set_surface(particle_scale=s)

    def setObstacleMeshingOffset(self, s: float):
        self._set_surface(obstacle_meshing_offset=s)

    def enableRemoveSurfaceNearDomain(self):
        self._set_surface(remove_near_domain=1)

    def disableRemoveSurfaceNearDomain(self):
        self._set_surface(remove_near_domain=0)

    def setRemoveSurfaceNearDomainDistance(self, n: int):
        self._set_surface(remove_near_domain_distance=n)

    def enableInvertedContactNormals(self):
        self._set_surface(invert_contact_normals=1)

    def disableInvertedContactNormals(self):
        self._set_surface(invert_contact_normals=0)

    def getIsomesh(self):
        """(vertices (n, 3) float32 in domain coordinates, triangles (m, 3)
        int32) of the last surface mesh (fluidsimulation.cpp:2543-2546)."""
        nv, nt = _c_i64(0), _c_i64(0)
        if self.h is None:
            return np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32)
        self.ctx.check(self.L.flipmi_scene_get_isomesh(self.h, ctypes.byref(nv), ctypes.byref(nt), None, None))
        verts = np.zeros((nv.value, 3), np.float32)
        tris = np.zeros((nt.value, 3), np.int32)
        self.ctx.check(self.L.flipmi_scene_get_isomesh(self.h, ctypes.byref(nv), ctypes.byref(nt), _ptr(verts),
                                                        _ptr(tris)))
        return verts, tris

    def _push_domain(self):
        if self.h is None:
            return  # pushed when the scene is created
        o = self.getDomainOffset()
        self.ctx.check(self.L.flipmi_scene_set_domain(self.h, o[0], o[1], o[2], self.getDomainScale()))

    def grid(self, which: int) -> np.ndarray:
        shape, dt = self._GRID_SHAPES[which](*self.dims)
        out = np.zeros(shape, dt)
        self.ctx.check(self.L.flipmi_scene_get_grid(self.h, which, _ptr(out)))
        return out

    def close(self):
        if self.h is not None:
            self.L.flipmi_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def scene_desc(isize: int, jsize: int, ksize: int, dx: float, body_force=(0.0, -9.81, 0.0),
               precision: int = PCG_FP64, substeps=(1, 6), cfl: float = 5.0, pic_ratio: float = 0.05,
               tolerance: float = 1e-9, max_iterations: int = 1000) -> SceneDesc:
    """flipmi_scene_desc with the reference's defaults (FluidSimulation's
    constructor and the setters the workloads use)."""
    d = SceneDesc()
    load_library().flipmi_scene_desc_default(ctypes.byref(d))
    d.isize, d.jsize, d.ksize, d.dx = isize, jsize, ksize, dx
    d.body_force[0], d.body_force[1], d.body_force[2] = body_force
    d.pcg_precision = precision
    d.min_substeps, d.max_substeps = substeps
    d.cfl, d.pic_ratio = cfl, pic_ratio
    d.pcg_tolerance, d.pcg_max_iterations = tolerance, max_iterations
    return d


class Server:
    """The serving loop (flipmi_server): independent scenes admitted by a
    memory budget and stepped concurrently, one HIP stream per slot."""

    def __init__(self, ctx: Optional[Context] = None, max_active: int = 8, memory_budget: int = 200 << 30,
                 particle_capacity_per_cell: float = 3.0, launch_graphs: bool = True):
        self.ctx = ctx or Context()
        self.L = self.ctx.L
        d = ServerDesc()
        self.L.flipmi_server_desc_default(ctypes.byref(d))
        d.max_active, d.memory_budget, d.particle_capacity_per_cell = max_active, memory_budget, particle_capacity_per_cell
        d.launch_graphs = int(bool(launch_graphs))
        h = _vp()
        self.ctx.check(self.L.flipmi_server_create(self.ctx.h, ctypes.byref(d), ctypes.byref(h)))
        self.h = h

    def submit(self, desc: SceneDesc, positions: np.ndarray, velocities: Optional[np.ndarray], dt: float,
               steps: int) -> int:
        pos = np.ascontiguousarray(positions, np.float32)
        vel = None if velocities is None else np.ascontiguousarray(velocities, np.float32)
        r = Request(desc, _ptr(pos), _ptr(vel), pos.shape[0], dt, steps)
        rid = _c_i64()
        self.ctx.check(self.L.flipmi_server_submit(self.h, ctypes.byref(r), ctypes.byref(rid)))
        return rid.value

    def wait(self, rid: int, timeout: float = -1.0) -> dict:
        rep = RequestReport()
        rc = self.L.flipmi_server_wait(self.h, rid, timeout, ctypes.byref(rep))
        if rc == FLIPMI_ERR_TIMEOUT:
            raise TimeoutError(f"request {rid} not done after {timeout} s")
        out = {name: getattr(rep, name) for name, _ in RequestReport._fields_}
        if rc != FLIPMI_OK:
            out["error"] = self.L.flipmi_last_error(self.ctx.h).decode()
        return out

    def take(self, rid: int, n: int) -> Tuple[np.ndarray, np.ndarray]:
        """The final particles of a finished request (n = its report's
        num_particles); releases the request."""
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        self.L.flipmi_server_take(self.h, rid, _ptr(pos) if n else None, _ptr(vel) if n else None)
        return pos, vel

    def discard(self, rid: int):
        """Releases a finished request without copying its particles."""
        self.L.flipmi_server_take(self.h, rid, None, None)

    def result(self, rid: int, timeout: float = -1.0):
        rep = self.wait(rid, timeout)
        pos, vel = self.take(rid, rep["num_particles"])
        return rep, pos, vel

    def stats(self) -> dict:
        st = ServerStats()
        self.ctx.check(self.L.flipmi_server_get_stats(self.h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in ServerStats._fields_}

    def close(self):
        if self.h is not None:
            self.L.flipmi_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
