"""The allocation-free entry point mph_scene_apply_matrix_noalloc (GPU).

It computes the product of mph_scene_apply_matrix on caller-owned device arrays,
with the factored product's records in the library's fixed scratch array instead of
the scene's buffers.  Its result must be bit-identical to mph_scene_apply_matrix on
the same scene and vector, on two calls in a row where the second scene is smaller
than the first (the scratch keeps the first call's contents), for both operator
forms and both precisions; a scene too large for the scratch is refused.
The device arrays are the caller's: they are allocated here through the HIP runtime
that the library itself links.
"""
import ctypes as C

import numpy as np
import pytest

from mphimplicit_amd import abi, scenes
from mphimplicit_amd.scene import Scene, lib
from test_gpu_parity import SCENES

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DeviceArray:
    """A caller-owned device buffer of doubles."""

    def __init__(self, hip, n):
        self.hip, self.n, self.p = hip, n, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), 8 * n) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, 8 * self.n, H2D) == 0

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, 8 * self.n, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _noalloc(hip, scene, precision, x):
    xd, yd = DeviceArray(hip, x.size), DeviceArray(hip, x.size)
    try:
        xd.put(x)
        yd.put(np.full(x.size, np.nan))  # every entry must be written
        rc = lib().mph_scene_apply_matrix_noalloc(scene.h, precision, xd.p, yd.p)
        assert rc == 0, lib().mph_last_error().decode()
        return yd.get()
    finally:
        xd.free()
        yd.free()


def _scene(name, form, precision):
    p, g = SCENES[name]()
    p.operator_form = form
    p.precision = precision
    s = Scene(p, g)
    s.step(1)
    return s


CASES = [  # (scene, operator form, scene precision, product precision); larger first
    ("3d_4096", abi.OPERATOR_FACTORED, abi.PRECISION_FP64, abi.PRECISION_FP64),
    ("tiny3d", abi.OPERATOR_FACTORED, abi.PRECISION_MIXED, abi.PRECISION_FP64),
    ("tiny3d", abi.OPERATOR_FACTORED, abi.PRECISION_MIXED, abi.PRECISION_MIXED),
    ("2d_40x40", abi.OPERATOR_ASSEMBLED, abi.PRECISION_MIXED, abi.PRECISION_MIXED),
    ("tiny2d", abi.OPERATOR_ASSEMBLED, abi.PRECISION_FP64, abi.PRECISION_FP64),
]


def test_noalloc_equals_apply_matrix():
    hip = _hip()
    sizes = []
    for name, form, sprec, prec in CASES:
        print("[noalloc] %s form %d precision %d/%d" % (name, form, sprec, prec), flush=True)
        s = _scene(name, form, sprec)
        n = 3 * int(s.scalar("FluidCount"))
        sizes.append(int(s.scalar("ParticleCount")))
        x = np.random.default_rng(len(sizes)).standard_normal(n)
        ref = s.apply_matrix(x, prec)
        got = _noalloc(hip, s, prec, x)
        assert np.array_equal(got.view(np.int64), ref.view(np.int64)), (name, np.abs(got - ref).max())
        again = _noalloc(hip, s, prec, x)  # a repeated call on the same scratch
        assert np.array_equal(again.view(np.int64), ref.view(np.int64))
        s.close()
    assert sizes[0] > sizes[1]  # the second factored call is on a smaller scene


def test_noalloc_rejects_too_large():
    """A factored FP64 scene of more than 1,797,558 particles does not fit the
    scratch: MPH_E_ARG with a message naming the limit, before any launch."""
    p, g = scenes.config_2d(nx=1400, ny=1400, box=(1500, 1450), steps=2)
    assert g.np > 1797558
    p.operator_form = abi.OPERATOR_FACTORED
    s = Scene(p, g)
    hip = _hip()
    n = 3 * int(s.scalar("FluidCount"))
    xd, yd = DeviceArray(hip, n), DeviceArray(hip, n)
    try:
        rc = lib().mph_scene_apply_matrix_noalloc(s.h, abi.PRECISION_FP64, xd.p, yd.p)
        assert rc == abi.E_ARG
        assert "1797558" in lib().mph_last_error().decode()
    finally:
        xd.free()
        yd.free()
        s.close()
