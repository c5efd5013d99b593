"""flipmi_extrapolate_noalloc: the extrapolation entry point that allocates
nothing (its BFS status grid is a fixed per-device array) gives the same
bits as flipmi_extrapolate on the same inputs, over two calls in a row with
the second one smaller, so the scratch array's leftover contents from the
first call cannot leak into the second."""

import ctypes

import numpy as np
import pytest

import flipengine3d_amd as fa
from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu

HIP_H2D, HIP_D2H = 1, 2


class _Hip:
    """hipMalloc/hipMemcpy/hipFree of the HIP runtime libflipmi.so links
    (the test is the caller that owns the device buffers)."""

    def __init__(self):
        fa.load_library()
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]

    def upload(self, a: np.ndarray) -> ctypes.c_void_p:
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(a.nbytes, 1)) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, HIP_H2D) == 0
        return p

    def download(self, p, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, HIP_D2H) == 0
        return out

    def free(self, p):
        assert self.L.hipFree(p) == 0


def _case(shape, seed):
    """A MAC-component-like grid: velocities everywhere, valid inside a blob,
    so the BFS has several layers to fill."""
    rng = np.random.default_rng(seed)
    nk, nj, ni = shape
    grid = rng.standard_normal(shape).astype(np.float32)
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    c = np.array([nk, nj, ni]) * np.array([0.4, 0.35, 0.5])
    r2 = ((k - c[0]) ** 2 + (j - c[1]) ** 2 + (i - c[2]) ** 2)
    valid = ((r2 < (0.3 * min(shape)) ** 2) | (rng.random(shape) < 0.02)).astype(np.uint8)
    return grid, valid


def _noalloc(ctx, hip, grid, valid, layers):
    nk, nj, ni = grid.shape
    dg, dv = hip.upload(grid), hip.upload(valid)
    try:
        rc = ctx.L.flipmi_extrapolate_noalloc(ctx.h, ni, nj, nk, dg, dv, layers, None)
        assert rc == fa.FLIPMI_OK, ctx.last_error()
        ctx.synchronize()
        return hip.download(dg, grid)
    finally:
        hip.free(dg)
        hip.free(dv)


def test_noalloc_matches_extrapolate_twice_then_smaller(gpu):
    hip = _Hip()
    cases = [((29, 33, 41), 7, 1), ((9, 12, 17), 3, 2)]
    for shape, layers, seed in cases:
        print(f"noalloc extrapolate shape={shape} layers={layers} seed={seed}", flush=True)
        grid, valid = _case(shape, seed)
        want = gpu.extrapolate(grid, valid, layers)
        got = _noalloc(gpu, hip, grid, valid, layers)
        assert bits_equal(got, want), f"shape {shape}: {int(np.count_nonzero(got != want))} faces differ"
        # The BFS did fill faces: the comparison is not of untouched inputs.
        assert int(np.count_nonzero(got != grid)) > 0


def test_noalloc_rejects_oversized_grid(gpu):
    # Checked on the host before anything is enqueued.
    limit = 64 * 1024 * 1024
    rc = gpu.L.flipmi_extrapolate_noalloc(gpu.h, 1024, 1024, limit // (1024 * 1024) + 1,
                                          ctypes.c_void_p(256), ctypes.c_void_p(256), 7, None)
    assert rc == fa.FLIPMI_ERR_INVALID_ARG
    assert "scratch limit" in gpu.last_error()
