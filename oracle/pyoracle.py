"""ctypes wrapper of the CPU oracle (oracle/libmph_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from mphimplicit_amd.abi import Grid, Params, StepStats

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmph_oracle.so")

DOT_SERIAL = 0
DOT_TREE = 1
DOT_OMP = 2

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle lib`")
        L = C.CDLL(LIB_PATH)
        L.mpho_create.restype = C.c_void_p
        L.mpho_create.argtypes = [C.POINTER(Params), C.POINTER(Grid), C.c_int]
        L.mpho_step.argtypes = [C.c_void_p, C.c_int, C.POINTER(StepStats)]
        L.mpho_run_phase.argtypes = [C.c_void_p, C.c_int]
        L.mpho_field_size.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_size_t), C.POINTER(C.c_int)]
        L.mpho_read.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_size_t]
        L.mpho_write.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_size_t]
        L.mpho_scalar.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]
        L.mpho_set_option.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.mpho_precondition.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.mpho_destroy.argtypes = [C.c_void_p]
        L.mpho_last_error.restype = C.c_char_p
        L.mpho_dot.restype = C.c_double
        L.mpho_dot.argtypes = [C.c_int, C.c_long, C.c_void_p, C.c_void_p]
        L.mpho_bitonic_sort.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        L.mpho_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


_DTYPES = {4: None, 8: None}


def set_threads(n: int) -> None:
    lib().mpho_set_threads(int(n))


def dot(mode: int, x: np.ndarray, y: np.ndarray) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().mpho_dot(mode, x.size, x.ctypes.data, y.ctypes.data)


def bitonic_sort(key: np.ndarray, val: np.ndarray):
    n = key.size
    power = n.bit_length() - 1
    assert 1 << power == n
    k = np.ascontiguousarray(key, dtype=np.int32).copy()
    v = np.ascontiguousarray(val, dtype=np.int32).copy()
    lib().mpho_bitonic_sort(power, k.ctypes.data, v.ctypes.data)
    return k, v


def make_grid(time, spacing, dmin, dmax, typ, pos, vel):
    """Build an mph_grid whose arrays stay alive with the returned tuple."""
    typ = np.ascontiguousarray(typ, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
    g = Grid()
    g.time = float(time)
    g.np = int(typ.size)
    g.spacing = float(spacing)
    for d in range(3):
        g.domain_min[d] = float(dmin[d])
        g.domain_max[d] = float(dmax[d])
    g.type = typ.ctypes.data_as(C.POINTER(C.c_int))
    g.pos = pos.ctypes.data_as(C.POINTER(C.c_double))
    g.vel = vel.ctypes.data_as(C.POINTER(C.c_double))
    return g, (typ, pos, vel)


class OracleScene:
    """One scene of the CPU restatement; same call surface as the HIP scene."""

    INT_FIELDS = {
        "Property", "ParticleIndex", "CellIndex", "CellParticle", "CellFluidParticleBegin",
        "CellFluidParticleEnd", "CellWallParticleBegin", "CellWallParticleEnd", "NeighborCount",
        "NeighborFluidCount", "NeighborCountP", "NeighborInd", "NeighborIndP", "CsrIndA", "CsrIndC",
    }
    LONG_FIELDS = {"NeighborPtr", "NeighborPtrP", "CsrPtrA", "CsrPtrC"}

    def __init__(self, params: Params, grid_data, dot_mode: int = DOT_SERIAL):
        g, keep = make_grid(
            grid_data.time, grid_data.spacing, grid_data.domain_min, grid_data.domain_max,
            grid_data.type, grid_data.pos, grid_data.vel,
        )
        self._keep = keep
        self.params = params.copy()
        h = lib().mpho_create(C.byref(self.params), C.byref(g), dot_mode)
        if not h:
            raise RuntimeError("oracle create failed: %s" % lib().mpho_last_error().decode())
        self.h = C.c_void_p(h)

    def close(self):
        if self.h:
            lib().mpho_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError("oracle error %d: %s" % (rc, lib().mpho_last_error().decode()))

    def step(self, n: int = 1) -> StepStats:
        st = StepStats()
        self._check(lib().mpho_step(self.h, n, C.byref(st)))
        return st

    def run_phase(self, phase: int) -> None:
        self._check(lib().mpho_run_phase(self.h, phase))

    def dtype_of(self, name: str, esize: int):
        if name in self.INT_FIELDS:
            return np.int32
        if name in self.LONG_FIELDS:
            return np.int64
        return np.float64 if esize == 8 else np.int32

    def read(self, name: str) -> np.ndarray:
        nb = C.c_size_t()
        es = C.c_int()
        self._check(lib().mpho_field_size(self.h, name.encode(), C.byref(nb), C.byref(es)))
        out = np.empty(nb.value // max(es.value, 1), dtype=self.dtype_of(name, es.value))
        self._check(lib().mpho_read(self.h, name.encode(), out.ctypes.data, nb.value))
        return out

    def write(self, name: str, arr: np.ndarray) -> None:
        nb = C.c_size_t()
        es = C.c_int()
        self._check(lib().mpho_field_size(self.h, name.encode(), C.byref(nb), C.byref(es)))
        a = np.ascontiguousarray(arr, dtype=self.dtype_of(name, es.value))
        if a.nbytes != nb.value:
            raise ValueError("size mismatch for %s: %d vs %d" % (name, a.nbytes, nb.value))
        self._check(lib().mpho_write(self.h, name.encode(), a.ctypes.data, a.nbytes))

    def scalar(self, name: str) -> float:
        v = C.c_double()
        self._check(lib().mpho_scalar(self.h, name.encode(), C.byref(v)))
        return v.value

    def set_option(self, name: str, value: float) -> None:
        self._check(lib().mpho_set_option(self.h, name.encode(), float(value)))

    def precondition(self, q: np.ndarray) -> np.ndarray:
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.full_like(q, -1.0)
        self._check(lib().mpho_precondition(self.h, q.ctypes.data, out.ctypes.data))
        return out
