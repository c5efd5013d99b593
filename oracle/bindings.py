"""ctypes bindings for the two CPU checkers -- TEST INFRASTRUCTURE.

``Oracle`` wraps oracle/liboracle.so (this repository's CPU restatement of
the reference FLIP step) and ``Reference`` wraps oracle/_ref/libflipref.so
(the reference engine compiled from its own sources plus
oracle/ref_harness.cpp).  Both expose the same methods so a test can run a
stage on either and compare.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libflipref.so")

_c_int = ctypes.c_int
_c_double = ctypes.c_double
_vp = ctypes.c_void_p


def build(ref: bool = True) -> None:
    """Compiles the checkers (make in oracle/)."""
    targets = ["oracle", "ref"] if ref else ["oracle"]
    subprocess.run(["make", "-s", "-C", HERE, "-j8"] + targets, check=True)


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


class _Base:
    prefix = ""

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = ctypes.CDLL(path)
        p = self.prefix
        L = self.lib
        getattr(L, p + "sim_create").restype = _vp
        self._sim_create = getattr(L, p + "sim_create")
        for name in ["sim_destroy", "sim_initialize", "sim_num_particles", "sim_frame_substeps"]:
            getattr(L, p + name).argtypes = [_vp]
        getattr(L, p + "sim_update").argtypes = [_vp, _c_double]
        getattr(L, p + "sim_add_body_force").argtypes = [_vp, _c_double, _c_double, _c_double]
        getattr(L, p + "sim_set_cfl").argtypes = [_vp, _c_double]
        getattr(L, p + "sim_set_substeps").argtypes = [_vp, _c_int, _c_int]
        getattr(L, p + "sim_set_pic_ratio").argtypes = [_vp, _c_double]
        getattr(L, p + "sim_load_particles").argtypes = [_vp, _vp, _vp, _c_int]
        getattr(L, p + "sim_get_particles").argtypes = [_vp, _vp, _vp]
        getattr(L, p + "sim_set_domain").argtypes = [_vp, _c_double, _c_double, _c_double, _c_double]
        getattr(L, p + "sim_get_particle_range").argtypes = [_vp, _c_int, _c_int, _vp, _vp]
        getattr(L, p + "sim_get_float_grid").argtypes = [_vp, _c_int, _vp]
        getattr(L, p + "sim_get_bool_grid").argtypes = [_vp, _c_int, _vp]
        getattr(L, p + "domain_solid").argtypes = [_c_int] * 3 + [_c_double] + [_vp] * 5
        getattr(L, p + "liquid_sdf").argtypes = [_c_int] * 3 + [_c_double, _vp, _c_int, _c_double, _vp, _vp]
        getattr(L, p + "p2g").argtypes = [_c_int] * 3 + [_c_double, _vp, _vp, _c_int, _c_double] + [_vp] * 6
        getattr(L, p + "extrapolate").argtypes = [_c_int] * 3 + [_vp, _vp, _c_int]
        getattr(L, p + "mic0_apply").argtypes = [_c_int] * 3 + [_c_double, _c_double] + [_vp] * 8
        getattr(L, p + "advect_particles").argtypes = ([_c_int] * 3 + [_c_double] + [_vp] * 4
                                                      + [_c_int, _c_double, _c_int, _vp])
        getattr(L, p + "trilinear_velocity").argtypes = [_c_int] * 3 + [_c_double] + [_vp] * 4 + [_c_int, _vp]
        getattr(L, p + "sim_add_inflow").argtypes = [_vp, _vp, _c_int, _vp, _c_int] + [_c_double] * 3
        getattr(L, p + "sim_set_inflow").argtypes = [_vp] + [_c_int] * 4
        getattr(L, p + "sim_set_outflow").argtypes = [_vp] + [_c_int] * 3
        getattr(L, p + "sim_set_obstacle_inversed").argtypes = [_vp] + [_c_int] * 2
        getattr(L, p + "sim_inflow_sdf").argtypes = [_vp, _c_int, _vp, _vp]
        getattr(L, p + "sim_isomesh").argtypes = [_vp] * 5
        surface = [_vp, _c_int, _c_int, _c_double, _c_int, _c_double, _c_double, _c_int, _c_int, _c_int]
        getattr(L, p + "sim_set_surface").argtypes = surface + ([_c_double] if p == "orc_" else [])

    # -- whole-step driver -------------------------------------------------
    def sim_create(self, isize, jsize, ksize, dx):
        return self._sim_create(isize, jsize, ksize, dx)

    def sim(self, isize, jsize, ksize, dx):
        return _SimHandle(self, isize, jsize, ksize, dx)

    # -- stages ------------------------------------------------------------
    def domain_solid(self, isize, jsize, ksize, dx):
        phi = np.zeros((ksize + 1, jsize + 1, isize + 1), np.float32)
        wu = np.zeros((ksize, jsize, isize + 1), np.float32)
        wv = np.zeros((ksize, jsize + 1, isize), np.float32)
        ww = np.zeros((ksize + 1, jsize, isize), np.float32)
        wc = np.zeros((ksize, jsize, isize), np.float32)
        getattr(self.lib, self.prefix + "domain_solid")(isize, jsize, ksize, dx, _ptr(phi), _ptr(wu),
                                                        _ptr(wv), _ptr(ww), _ptr(wc))
        return phi, (wu, wv, ww, wc)

    def liquid_sdf(self, isize, jsize, ksize, dx, pos, radius, solid_phi=None):
        pos = np.ascontiguousarray(pos, np.float32)
        phi = np.zeros((ksize, jsize, isize), np.float32)
        sp = None if solid_phi is None else np.ascontiguousarray(solid_phi, np.float32)
        getattr(self.lib, self.prefix + "liquid_sdf")(isize, jsize, ksize, dx, _ptr(pos), pos.shape[0],
                                                      radius, _ptr(sp), _ptr(phi))
        return phi

    def p2g(self, isize, jsize, ksize, dx, pos, vel, radius):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        u = np.zeros((ksize, jsize, isize + 1), np.float32)
        v = np.zeros((ksize, jsize + 1, isize), np.float32)
        w = np.zeros((ksize + 1, jsize, isize), np.float32)
        vu = np.zeros(u.shape, np.uint8)
        vv = np.zeros(v.shape, np.uint8)
        vw = np.zeros(w.shape, np.uint8)
        getattr(self.lib, self.prefix + "p2g")(isize, jsize, ksize, dx, _ptr(pos), _ptr(vel), pos.shape[0],
                                               radius, _ptr(u), _ptr(v), _ptr(w), _ptr(vu), _ptr(vv), _ptr(vw))
        return (u, v, w), (vu, vv, vw)

    def extrapolate(self, grid, valid, layers):
        g = np.ascontiguousarray(grid, np.float32).copy()
        vd = np.ascontiguousarray(valid, np.uint8).copy()
        nk, nj, ni = g.shape
        getattr(self.lib, self.prefix + "extrapolate")(ni, nj, nk, _ptr(g), _ptr(vd), layers)
        return g

    def mic0_apply(self, isize, jsize, ksize, dx, dt, liquid_phi, weights, r):
        wu, wv, ww, wc = [np.ascontiguousarray(a, np.float32) for a in weights]
        lp = np.ascontiguousarray(liquid_phi, np.float32)
        r = np.ascontiguousarray(r, np.float64)
        inv = np.zeros_like(r)
        z = np.zeros_like(r)
        getattr(self.lib, self.prefix + "mic0_apply")(isize, jsize, ksize, dx, dt, _ptr(lp), _ptr(wu),
                                                      _ptr(wv), _ptr(ww), _ptr(wc), _ptr(r), _ptr(inv), _ptr(z))
        return inv, z

    def advect_particles(self, isize, jsize, ksize, dx, mac, pos, dt, order):
        u, v, w = [np.ascontiguousarray(a, np.float32) for a in mac]
        pos = np.ascontiguousarray(pos, np.float32)
        out = np.zeros_like(pos)
        getattr(self.lib, self.prefix + "advect_particles")(isize, jsize, ksize, dx, _ptr(u), _ptr(v), _ptr(w),
                                                            _ptr(pos), pos.shape[0], dt, order, _ptr(out))
        return out

    def trilinear_velocity(self, isize, jsize, ksize, dx, mac, pos):
        u, v, w = [np.ascontiguousarray(a, np.float32) for a in mac]
        pos = np.ascontiguousarray(pos, np.float32)
        out = np.zeros_like(pos)
        getattr(self.lib, self.prefix + "trilinear_velocity")(isize, jsize, ksize, dx, _ptr(u), _ptr(v),
                                                              _ptr(w), _ptr(pos), pos.shape[0], _ptr(out))
        return out


class _SimHandle:
    """Owns one FluidSimulation (reference) or Sim (oracle) instance."""

    def __init__(self, lib: _Base, isize, jsize, ksize, dx):
        self.L = lib
        self.dims = (isize, jsize, ksize)
        self.dx = dx
        self.h = lib.sim_create(isize, jsize, ksize, dx)

    def _f(self, name):
        return getattr(self.L.lib, self.L.prefix + name)

    def close(self):
        if self.h:
            self._f("sim_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_body_force(self, fx, fy, fz):
        self._f("sim_add_body_force")(self.h, fx, fy, fz)

    def set_cfl(self, c):
        self._f("sim_set_cfl")(self.h, c)

    def set_substeps(self, mn, mx):
        self._f("sim_set_substeps")(self.h, mn, mx)

    def set_pic_ratio(self, r):
        self._f("sim_set_pic_ratio")(self.h, r)

    def load_particles(self, pos, vel):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        self._f("sim_load_particles")(self.h, _ptr(pos), _ptr(vel), pos.shape[0])

    def initialize(self):
        return self._f("sim_initialize")(self.h)

    def update(self, dt):
        rc = self._f("sim_update")(self.h, dt)
        if rc != 0:
            raise RuntimeError("update failed")

    @property
    def num_particles(self):
        return self._f("sim_num_particles")(self.h)

    def particles(self):
        n = self.num_particles
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        self._f("sim_get_particles")(self.h, _ptr(pos), _ptr(vel))
        return pos, vel

    def set_domain(self, offset, scale):
        """setDomainOffset + setDomainScale (fluidsimulation.cpp:307-330)."""
        self._f("sim_set_domain")(self.h, float(offset[0]), float(offset[1]), float(offset[2]), float(scale))

    def particle_range(self, start, end):
        """getMarkerParticle{Position,Velocity}DataRange: positions in domain
        coordinates; None for a range the reference rejects."""
        n = max(end - start, 0)
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        rc = self._f("sim_get_particle_range")(self.h, start, end, _ptr(pos) if n else None, _ptr(vel) if n else None)
        return None if rc != 0 else (pos, vel)

    def substeps(self):
        return self._f("sim_frame_substeps")(self.h)

    # -- mesh inflow sources ---------------------------------------------
    def add_inflow(self, verts, tris, velocity):
        """MeshFluidSource with a static mesh, inflow, setVelocity; returns its index."""
        verts = np.ascontiguousarray(verts, np.float32)
        tris = np.ascontiguousarray(tris, np.int32)
        return self._f("sim_add_inflow")(self.h, _ptr(verts), verts.shape[0], _ptr(tris), tris.shape[0],
                                         float(velocity[0]), float(velocity[1]), float(velocity[2]))

    def set_inflow(self, sid, enabled=True, constrain=True, emissions=1):
        self._f("sim_set_inflow")(self.h, sid, int(enabled), int(constrain), int(emissions))

    def add_obstacle(self, verts, tris, jitter=None):
        """MeshObject with a static mesh added with addMeshObstacle; the oracle
        also takes the ray jitter of its inside test (None: keyed draw)."""
        verts = np.ascontiguousarray(verts, np.float32)
        tris = np.ascontiguousarray(tris, np.int32)
        if self.L.prefix == "orc_":
            jit = None if jitter is None else np.ascontiguousarray(jitter, np.float32)
            return self._f("sim_add_obstacle")(self.h, _ptr(verts), verts.shape[0], _ptr(tris), tris.shape[0],
                                               _ptr(jit))
        return self._f("sim_add_obstacle")(self.h, _ptr(verts), verts.shape[0], _ptr(tris), tris.shape[0])

    def set_obstacle(self, oid, enabled):
        self._f("sim_set_obstacle")(self.h, oid, int(enabled))

    def set_obstacle_inversed(self, oid, inversed=True):
        self._f("sim_set_obstacle_inversed")(self.h, oid, int(inversed))

    def set_outflow(self, sid, outflow=True, inversed=False):
        """setOutflow() / outflowInverse() of a source."""
        self._f("sim_set_outflow")(self.h, sid, int(outflow), int(inversed))

    def set_jitter(self, factor, surface=False, seed=0):
        """setMarkerParticleJitterFactor + jitter-surface switch (the oracle
        also takes the seed of its jitter draws)."""
        if self.L.prefix == "orc_":
            self._f("sim_set_jitter")(self.h, float(factor), int(surface), int(seed))
        else:
            self._f("sim_set_jitter")(self.h, float(factor), int(surface))

    def inflow_sdf(self, sid):
        """(grid offset (i, j, k), nodal phi (nk, nj, ni)) of a source after its first update."""
        dims = np.zeros(6, np.int32)
        if self._f("sim_inflow_sdf")(self.h, sid, _ptr(dims), None) != 0:
            return None
        phi = np.zeros((dims[5], dims[4], dims[3]), np.float32)
        self._f("sim_inflow_sdf")(self.h, sid, _ptr(dims), _ptr(phi))
        return tuple(int(v) for v in dims[:3]), phi

    _FSHAPES = {
        0: lambda I, J, K: (K, J, I),
        1: lambda I, J, K: (K + 1, J + 1, I + 1),
        2: lambda I, J, K: (K, J, I + 1),
        3: lambda I, J, K: (K, J + 1, I),
        4: lambda I, J, K: (K + 1, J, I),
        5: lambda I, J, K: (K, J, I + 1),
        6: lambda I, J, K: (K, J + 1, I),
        7: lambda I, J, K: (K + 1, J, I),
        8: lambda I, J, K: (K, J, I),
    }

    def float_grid(self, which):
        shape = self._FSHAPES[which](*self.dims)
        out = np.zeros(shape, np.float32)
        self._f("sim_get_float_grid")(self.h, which, _ptr(out))
        return out

    def valid_grid(self, which):
        shape = self._FSHAPES[2 + which](*self.dims)
        out = np.zeros(shape, np.uint8)
        self._f("sim_get_bool_grid")(self.h, which, _ptr(out))
        return out

    def set_surface(self, enabled=True, subdivisions=1, smooth_value=0.5, smooth_iterations=2,
                    particle_scale=3.0, obstacle_offset=0.0, remove_near_domain=False, remove_distance=0,
                    invert_contact=False):
        """Surface reconstruction settings (the engine's setters; meshing
        after the first substep of every frame)."""
        args = [self.h, int(enabled), subdivisions, smooth_value, smooth_iterations, particle_scale,
                obstacle_offset, int(remove_near_domain), remove_distance, int(invert_contact)]
        if self.L.prefix == "orc_":
            args.append(0.08)
        self._f("sim_set_surface")(*args)

    def isomesh(self):
        """(vertices (n, 3) float32, triangles (m, 3) int32) of getIsomesh()."""
        nv, nt = ctypes.c_int(0), ctypes.c_int(0)
        f = self._f("sim_isomesh")
        f(self.h, ctypes.byref(nv), ctypes.byref(nt), None, None)
        verts = np.zeros((nv.value, 3), np.float32)
        tris = np.zeros((nt.value, 3), np.int32)
        f(self.h, ctypes.byref(nv), ctypes.byref(nt), _ptr(verts), _ptr(tris))
        return verts, tris

    def count_grid(self):
        """(count grid of the last removal pass, particles that pass removed
        for their speed, speed removals since creation).  Oracle only."""
        I, J, K = self.dims
        out = np.zeros((K, J, I), np.int32)
        last, total = ctypes.c_int(0), ctypes.c_int(0)
        self._f("sim_get_count_grid")(self.h, _ptr(out), ctypes.byref(last), ctypes.byref(total))
        return out, last.value, total.value


class Oracle(_Base):
    prefix = "orc_"

    def __init__(self, path: str = ORACLE_SO):
        super().__init__(path)
        L = self.lib
        L.orc_sim_create.argtypes = [_c_int] * 3 + [_c_double]
        L.orc_sim_set_pcg.argtypes = [_vp, _c_int, _c_double, _c_int]
        L.orc_sim_pressure_info.argtypes = [_vp, _vp, _vp, _vp]
        L.orc_pressure_solve.argtypes = ([_c_int] * 3 + [_c_double, _c_double, _c_double, _c_int, _c_int]
                                         + [_vp] * 15 + [_vp] * 5)
        L.orc_sim_set_jitter.argtypes = [_vp, _c_double, _c_int, ctypes.c_uint64]
        L.orc_mesh_nodes_inside.argtypes = [_vp, _c_int, _vp, _c_int, _c_double] + [_c_int] * 3 + [_vp, _vp]
        L.orc_mesh_sdf.argtypes = [_vp, _c_int, _vp, _c_int] + [_c_int] * 3 + [_c_double, _c_int, _vp, _vp]
        L.orc_sim_inflow_cells.argtypes = [_vp, _c_int, _vp]
        L.orc_ray_jitter.argtypes = [ctypes.c_uint64, _c_int, _c_int, _c_double, _vp]
        L.orc_sim_add_obstacle.argtypes = [_vp, _vp, _c_int, _vp, _c_int, _vp]
        L.orc_sim_set_obstacle.argtypes = [_vp, _c_int, _c_int]
        L.orc_inflow_build.argtypes = ([_vp, _c_int, _vp, _c_int] + [_c_int] * 3
                                       + [_c_double, ctypes.c_uint64, _c_int, _vp, _vp, _vp])
        L.orc_sim_get_count_grid.argtypes = [_vp, _vp, _vp, _vp]
        L.orc_removal_counts.argtypes = [_c_int] * 3 + [_c_double, _vp, _c_int, _vp, _vp]

    def surface_mesh(self, isize, jsize, ksize, dx, pos, solid_phi, subdivisions=1, smooth_value=0.5,
                     smooth_iterations=2, particle_scale=3.0, obstacle_offset=0.0, remove_near_domain=False,
                     remove_distance=0, invert_contact=False, contact_threshold=0.08, domain_offset=(0, 0, 0),
                     domain_scale=1.0):
        """(vertices, triangles) of the surface of `pos` over the nodal solid
        SDF, as getIsomesh() returns it."""
        L = self.lib
        L.orc_surface_mesh.restype = _vp
        L.orc_surface_mesh.argtypes = ([_c_int] * 3 + [_c_double, _vp, _c_int, _vp, _c_int, _c_double, _c_int,
                                       _c_double, _c_double, _c_int, _c_int, _c_int, _c_double, _vp, _c_double])
        L.orc_mesh_get.argtypes = [_vp] * 5
        L.orc_mesh_free.argtypes = [_vp]
        pos = np.ascontiguousarray(pos, np.float32)
        sp = np.ascontiguousarray(solid_phi, np.float32)
        off = np.asarray(domain_offset, np.float32)
        h = L.orc_surface_mesh(isize, jsize, ksize, dx, _ptr(pos), pos.shape[0], _ptr(sp), subdivisions,
                               smooth_value, smooth_iterations, particle_scale, obstacle_offset,
                               int(remove_near_domain), remove_distance, int(invert_contact), contact_threshold,
                               _ptr(off), domain_scale)
        try:
            nv, nt = ctypes.c_int(0), ctypes.c_int(0)
            L.orc_mesh_get(h, ctypes.byref(nv), ctypes.byref(nt), None, None)
            verts = np.zeros((nv.value, 3), np.float32)
            tris = np.zeros((nt.value, 3), np.int32)
            L.orc_mesh_get(h, ctypes.byref(nv), ctypes.byref(nt), _ptr(verts), _ptr(tris))
        finally:
            L.orc_mesh_free(h)
        return verts, tris

    def removal_counts(self, isize, jsize, ksize, dx, pos, removed=None):
        """Count grid of the removal pass's counting loop on given positions
        (removed: optional per-particle flags of the solid test)."""
        pos = np.ascontiguousarray(pos, np.float32)
        rm = None if removed is None else np.ascontiguousarray(removed, np.uint8)
        out = np.zeros((ksize, jsize, isize), np.int32)
        self.lib.orc_removal_counts(isize, jsize, ksize, dx, _ptr(pos), pos.shape[0], _ptr(rm), _ptr(out))
        return out

    def _mesh_args(self, verts, tris):
        verts = np.ascontiguousarray(verts, np.float32)
        tris = np.ascontiguousarray(tris, np.int32)
        return verts, tris

    def mesh_nodes_inside(self, verts, tris, dx, dims, jitter):
        """Nodes (nk, nj, ni) of an ni x nj x nk node grid inside a closed mesh."""
        verts, tris = self._mesh_args(verts, tris)
        ni, nj, nk = dims
        out = np.zeros((nk, nj, ni), np.uint8)
        jit = np.ascontiguousarray(jitter, np.float32)
        rc = self.lib.orc_mesh_nodes_inside(_ptr(verts), verts.shape[0], _ptr(tris), tris.shape[0], dx,
                                            ni, nj, nk, _ptr(jit), _ptr(out))
        return None if rc != 0 else out

    def mesh_sdf(self, verts, tris, isize, jsize, ksize, dx, band, jitter):
        verts, tris = self._mesh_args(verts, tris)
        phi = np.zeros((ksize + 1, jsize + 1, isize + 1), np.float32)
        jit = np.ascontiguousarray(jitter, np.float32)
        rc = self.lib.orc_mesh_sdf(_ptr(verts), verts.shape[0], _ptr(tris), tris.shape[0], isize, jsize,
                                   ksize, dx, band, _ptr(jit), _ptr(phi))
        return None if rc != 0 else phi

    def inflow_cells(self, sim, sid):
        I, J, K = sim.dims
        out = np.zeros((K, J, I), np.uint8)
        if self.lib.orc_sim_inflow_cells(sim.h, sid, _ptr(out)) != 0:
            return None
        return out

    def inflow_build(self, verts, tris, dims, dx, seed=0, sid=0):
        """(sub-grid offset, nodal SDF (nk, nj, ni), cells (K, J, I)) of a source."""
        verts, tris = self._mesh_args(verts, tris)
        I, J, K = dims
        d = np.zeros(6, np.int32)
        L = self.lib
        if L.orc_inflow_build(_ptr(verts), verts.shape[0], _ptr(tris), tris.shape[0], I, J, K, dx, seed, sid,
                              _ptr(d), None, None) != 0:
            return None
        phi = np.zeros((d[5], d[4], d[3]), np.float32)
        cells = np.zeros((K, J, I), np.uint8)
        L.orc_inflow_build(_ptr(verts), verts.shape[0], _ptr(tris), tris.shape[0], I, J, K, dx, seed, sid,
                           _ptr(d), _ptr(phi), _ptr(cells))
        return tuple(int(v) for v in d[:3]), phi, cells

    def ray_jitter(self, seed, sid, which, dx):
        out = np.zeros(3, np.float32)
        self.lib.orc_ray_jitter(seed, sid, which, dx, _ptr(out))
        return out

    def pressure_solve(self, isize, jsize, ksize, dx, dt, liquid_phi, solid_phi, weights, mac, valid,
                       tol=1e-9, maxiter=1000, mode=0, solid_vel=None):
        return _pressure_solve(self.lib.orc_pressure_solve, True, isize, jsize, ksize, dx, dt, liquid_phi,
                               solid_phi, weights, mac, valid, tol, maxiter, mode, solid_vel)


class Reference(_Base):
    prefix = "ref_"

    def __init__(self, path: str = REF_SO, threads: int = 1):
        super().__init__(path)
        L = self.lib
        L.ref_sim_create.argtypes = [_c_int] * 3 + [_c_double, _c_int]
        L.ref_set_thread_count.argtypes = [_c_int]
        L.ref_pressure_solve.argtypes = ([_c_int] * 3 + [_c_double, _c_double, _c_double, _c_int]
                                         + [_vp] * 15 + [_vp] * 5)
        L.ref_sim_pressure_status.argtypes = [_vp, ctypes.c_char_p, _c_int]
        L.ref_sim_set_jitter.argtypes = [_vp, _c_double, _c_int]
        L.ref_mesh_nodes_inside.argtypes = ([_vp, _c_int, _vp, _c_int, _c_double] + [_c_int] * 3
                                            + [ctypes.c_uint, _vp, _vp])
        L.ref_mesh_sdf.argtypes = ([_vp, _c_int, _vp, _c_int] + [_c_int] * 3
                                   + [_c_double, _c_int, ctypes.c_uint, _vp, _vp])
        L.ref_sim_inflow_cells.argtypes = [_vp, _c_int, ctypes.c_uint, _vp, _vp]
        L.ref_sim_add_obstacle.argtypes = [_vp, _vp, _c_int, _vp, _c_int]
        L.ref_sim_set_obstacle.argtypes = [_vp, _c_int, _c_int]
        L.ref_predict_ray_jitters.argtypes = [ctypes.c_uint, _c_double, _c_int, _vp]
        L.ref_seed_rand.argtypes = [ctypes.c_uint]
        self.threads = threads

    def sim_create(self, isize, jsize, ksize, dx):
        return self.lib.ref_sim_create(isize, jsize, ksize, dx, self.threads)

    def set_threads(self, n):
        self.threads = n
        self.lib.ref_set_thread_count(n)

    def predict_ray_jitters(self, seed, dx, count):
        """The ray jitters of the next `count` inside tests after srand(seed);
        reseeds rand() with `seed` afterwards."""
        out = np.zeros((count, 3), np.float32)
        self.lib.ref_predict_ray_jitters(seed, dx, count, _ptr(out))
        self.lib.ref_seed_rand(seed)
        return out

    def mesh_nodes_inside(self, verts, tris, dx, dims, seed):
        """(nodes, ray jitter) of getGridNodesInsideTriangleMesh with rand() seeded."""
        verts = np.ascontiguousarray(verts, np.float32)
        tris = np.ascontiguousarray(tris, np.int32)
        ni, nj, nk = dims
        out = np.zeros((nk, nj, ni), np.uint8)
        jit = np.zeros(3, np.float32)
        self.lib.ref_mesh_nodes_inside(_ptr(verts), verts.shape[0], _ptr(tris), tris.shape[0], dx, ni, nj, nk,
                                       seed, _ptr(jit), _ptr(out))
        return out, jit

    def mesh_sdf(self, verts, tris, isize, jsize, ksize, dx, band, seed):
        verts = np.ascontiguousarray(verts, np.float32)
        tris = np.ascontiguousarray(tris, np.int32)
        phi = np.zeros((ksize + 1, jsize + 1, isize + 1), np.float32)
        jit = np.zeros(3, np.float32)
        self.lib.ref_mesh_sdf(_ptr(verts), verts.shape[0], _ptr(tris), tris.shape[0], isize, jsize, ksize,
                              dx, band, seed, _ptr(jit), _ptr(phi))
        return phi, jit

    def inflow_cells(self, sim, sid, seed):
        I, J, K = sim.dims
        out = np.zeros((K, J, I), np.uint8)
        jit = np.zeros(3, np.float32)
        self.lib.ref_sim_inflow_cells(sim.h, sid, seed, _ptr(jit), _ptr(out))
        return out, jit

    def pressure_solve(self, isize, jsize, ksize, dx, dt, liquid_phi, solid_phi, weights, mac, valid,
                       tol=1e-9, maxiter=1000, mode=0, solid_vel=None):
        return _pressure_solve(self.lib.ref_pressure_solve, False, isize, jsize, ksize, dx, dt, liquid_phi,
                               solid_phi, weights, mac, valid, tol, maxiter, mode, solid_vel)


def _pressure_solve(fn, with_mode, isize, jsize, ksize, dx, dt, liquid_phi, solid_phi, weights, mac, valid,
                    tol, maxiter, mode, solid_vel=None):
    """solid_vel: optional (U, V, W) solid face velocities; the returned
    'solid_vel' holds them after the solver's pocket conditioning."""
    lp = np.ascontiguousarray(liquid_phi, np.float32)
    sp = np.ascontiguousarray(solid_phi, np.float32)
    wu, wv, ww, wc = [np.ascontiguousarray(a, np.float32) for a in weights]
    u, v, w = [np.ascontiguousarray(a, np.float32).copy() for a in mac]
    vu, vv, vw = [np.ascontiguousarray(a, np.uint8).copy() for a in valid]
    sv = (None, None, None) if solid_vel is None else [np.ascontiguousarray(a, np.float32).copy() for a in solid_vel]
    nmax = isize * jsize * ksize
    rhs = np.zeros(nmax, np.float64)
    soln = np.zeros(nmax, np.float64)
    n = ctypes.c_int(0)
    it = ctypes.c_int(0)
    err = ctypes.c_double(0)
    args = [isize, jsize, ksize, dx, dt, tol, maxiter]
    if with_mode:
        args.append(mode)
    ok = fn(*args, _ptr(lp), _ptr(sp), _ptr(wu), _ptr(wv), _ptr(ww), _ptr(wc), _ptr(u), _ptr(v), _ptr(w),
            _ptr(vu), _ptr(vv), _ptr(vw), _ptr(sv[0]), _ptr(sv[1]), _ptr(sv[2]), _ptr(rhs), _ptr(soln),
            ctypes.byref(n), ctypes.byref(it), ctypes.byref(err))
    m = n.value
    return dict(ok=int(ok), rhs=rhs[:m], soln=soln[:m], iters=it.value, err=err.value,
                mac=(u, v, w), valid=(vu, vv, vw), solid_vel=None if solid_vel is None else tuple(sv))
