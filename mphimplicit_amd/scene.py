"""Python binding of libmph_hip.so (the C ABI of include/mph_abi.h).

The library is loaded from this package directory; if it is missing, every entry
point raises, because the time stepper has no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi
from .abi import Grid, Params, StepStats

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmph_hip.so")

_lib = None


class MphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("mph error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Load libmph_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libmph_hip.so is not built (run __graft_entry__.build() or "
                "`make -C mphimplicit_amd/csrc`); there is no CPU fallback"
            )
        L = C.CDLL(LIB_PATH)
        vp, cp, i, d = C.c_void_p, C.c_char_p, C.c_int, C.c_double
        L.mph_last_error.restype = cp
        L.mph_params_default.argtypes = [C.POINTER(Params)]
        L.mph_params_default.restype = None
        L.mph_read_data_file.argtypes = [cp, C.POINTER(Params)]
        L.mph_read_grid_header.argtypes = [cp, C.POINTER(i)]
        L.mph_read_grid_file.argtypes = [cp, i, C.POINTER(d), C.POINTER(d), vp, vp, vp, vp, vp]
        L.mph_scene_create.argtypes = [C.POINTER(Params), C.POINTER(Grid), i, C.POINTER(vp)]
        L.mph_scene_step.argtypes = [vp, i, C.POINTER(StepStats)]
        L.mph_scene_run_phase.argtypes = [vp, i]
        L.mph_scene_field_size.argtypes = [vp, cp, C.POINTER(C.c_size_t), C.POINTER(i)]
        L.mph_scene_read.argtypes = [vp, cp, vp, C.c_size_t]
        L.mph_scene_write.argtypes = [vp, cp, vp, C.c_size_t]
        L.mph_scene_scalar.argtypes = [vp, cp, C.POINTER(d)]
        L.mph_scene_domain.argtypes = [vp, C.POINTER(d), vp, vp]
        L.mph_scene_write_prof.argtypes = [vp, cp]
        L.mph_scene_write_vtk.argtypes = [vp, cp]
        L.mph_scene_set_timing.argtypes = [vp, i]
        L.mph_scene_set_option.argtypes = [vp, cp, d]
        L.mph_scene_destroy.argtypes = [vp]
        L.mph_scene_device_bytes.argtypes = [vp, C.POINTER(C.c_ulonglong)]
        L.mph_scene_apply_matrix.argtypes = [vp, i, C.POINTER(d), C.POINTER(d)]
        L.mph_scene_apply_matrix_noalloc.argtypes = [vp, i, vp, vp]  # device pointers
        L.mph_batch_create.argtypes = [i, C.POINTER(vp)]
        L.mph_batch_add.argtypes = [vp, vp]
        L.mph_batch_remove.argtypes = [vp, vp]
        L.mph_batch_size.argtypes = [vp]
        L.mph_batch_step.argtypes = [vp, vp]
        L.mph_batch_destroy.argtypes = [vp]
        L.mph_probe_bandwidth.argtypes = [i, C.c_ulonglong, i, C.POINTER(d), C.POINTER(d)]
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise MphError(rc, lib().mph_last_error().decode())


INT_FIELDS = {
    "Property", "ParticleIndex", "CellIndex", "CellParticle", "CellFluidParticleBegin", "CellFluidParticleEnd",
    "CellWallParticleBegin", "CellWallParticleEnd", "NeighborCount", "NeighborFluidCount", "NeighborCountP",
    "NeighborInd", "NeighborIndP", "CsrIndA", "CsrIndC", "ACol",
}
LONG_FIELDS = {"NeighborPtr", "NeighborPtrP", "CsrPtrA", "CsrPtrC"}


def field_dtype(name: str, esize: int):
    if name in INT_FIELDS:
        return np.int32
    if name in LONG_FIELDS:
        return np.int64
    return np.float64 if esize == 8 else np.int32


def make_grid(time, spacing, dmin, dmax, typ, pos, vel):
    """An mph_grid over numpy arrays; keep the returned arrays alive while in use."""
    typ = np.ascontiguousarray(typ, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
    g = Grid()
    g.time = float(time)
    g.np = int(typ.size)
    g.spacing = float(spacing)
    for k in range(3):
        g.domain_min[k] = float(dmin[k])
        g.domain_max[k] = float(dmax[k])
    g.type = typ.ctypes.data_as(C.POINTER(C.c_int))
    g.pos = pos.ctypes.data_as(C.POINTER(C.c_double))
    g.vel = vel.ctypes.data_as(C.POINTER(C.c_double))
    return g, (typ, pos, vel)


def read_data_file(path: str, dim: int = 3) -> Params:
    """mph_read_data_file: parse a *.data file with the library's own parser."""
    p = Params()
    lib().mph_params_default(C.byref(p))
    p.dim = dim
    _check(lib().mph_read_data_file(path.encode(), C.byref(p)))
    return p


class Scene:
    """One MPH-I scene resident on a HIP device (mph_scene_* of the C ABI)."""

    def __init__(self, params: Params, grid_data, device: int = 0):
        g, keep = make_grid(
            grid_data.time, grid_data.spacing, grid_data.domain_min, grid_data.domain_max,
            grid_data.type, grid_data.pos, grid_data.vel,
        )
        self._keep = keep
        self.params = params.copy()
        h = C.c_void_p()
        _check(lib().mph_scene_create(C.byref(self.params), C.byref(g), int(device), C.byref(h)))
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            _check(lib().mph_scene_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, n: int = 1) -> StepStats:
        st = StepStats()
        _check(lib().mph_scene_step(self.h, int(n), C.byref(st)))
        return st

    def run_phase(self, phase: int) -> None:
        _check(lib().mph_scene_run_phase(self.h, int(phase)))

    def set_timing(self, on: int) -> None:
        _check(lib().mph_scene_set_timing(self.h, int(on)))

    def set_option(self, name: str, value: float) -> None:
        """mph_scene_set_option: a schedule choice that never changes a result, or
        ConvergenceCheck (the reference's CONVERGENCE_CHECK record of each solve)."""
        _check(lib().mph_scene_set_option(self.h, name.encode(), float(value)))

    def field_bytes(self, name: str) -> int:
        """Current size of a named field in bytes (mph_scene_field_size)."""
        nb = C.c_size_t()
        es = C.c_int()
        _check(lib().mph_scene_field_size(self.h, name.encode(), C.byref(nb), C.byref(es)))
        return int(nb.value)

    def read(self, name: str) -> np.ndarray:
        nb = C.c_size_t()
        es = C.c_int()
        _check(lib().mph_scene_field_size(self.h, name.encode(), C.byref(nb), C.byref(es)))
        out = np.empty(nb.value // max(es.value, 1), dtype=field_dtype(name, es.value))
        _check(lib().mph_scene_read(self.h, name.encode(), out.ctypes.data, nb.value))
        return out

    def write(self, name: str, arr: np.ndarray) -> None:
        nb = C.c_size_t()
        es = C.c_int()
        _check(lib().mph_scene_field_size(self.h, name.encode(), C.byref(nb), C.byref(es)))
        a = np.ascontiguousarray(arr, dtype=field_dtype(name, es.value))
        if a.nbytes != nb.value:
            raise ValueError("size mismatch for %s: %d vs %d bytes" % (name, a.nbytes, nb.value))
        _check(lib().mph_scene_write(self.h, name.encode(), a.ctypes.data, a.nbytes))

    def scalar(self, name: str) -> float:
        v = C.c_double()
        _check(lib().mph_scene_scalar(self.h, name.encode(), C.byref(v)))
        return v.value

    def write_vtk(self, path: str) -> None:
        _check(lib().mph_scene_write_vtk(self.h, path.encode()))

    def write_prof(self, path: str) -> None:
        _check(lib().mph_scene_write_prof(self.h, path.encode()))

    def device_bytes(self) -> int:
        v = C.c_ulonglong()
        _check(lib().mph_scene_device_bytes(self.h, C.byref(v)))
        return v.value

    def apply_matrix(self, x: np.ndarray, precision: int = abi.PRECISION_FP64) -> np.ndarray:
        """y = A x with the last assembled particle matrix, in the scene's operator
        form (mph_scene_apply_matrix); x has 3*FluidCount entries."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        _check(lib().mph_scene_apply_matrix(self.h, int(precision), x.ctypes.data_as(C.POINTER(C.c_double)),
                                            y.ctypes.data_as(C.POINTER(C.c_double))))
        return y


def probe_bandwidth(device: int = 0, nbytes: int = 8 << 30, reps: int = 5) -> tuple[float, float]:
    """mph_probe_bandwidth: (streaming-read GB/s, copy GB/s) measured on `device`."""
    r, c = C.c_double(), C.c_double()
    _check(lib().mph_probe_bandwidth(int(device), int(nbytes), int(reps), C.byref(r), C.byref(c)))
    return r.value, c.value


class Batch:
    """mph_batch_*: many scenes advanced one time step together (lockstep CR).

    The batch does not own its scenes; keep them alive while they are members."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().mph_batch_create(int(device), C.byref(h)))
        self.h = h
        self.members: list[Scene] = []

    def add(self, scene: Scene) -> None:
        _check(lib().mph_batch_add(self.h, scene.h))
        self.members.append(scene)

    def remove(self, scene: Scene) -> None:
        _check(lib().mph_batch_remove(self.h, scene.h))
        self.members.remove(scene)

    def __len__(self) -> int:
        return int(lib().mph_batch_size(self.h))

    def step(self) -> list:
        """One time step of every member; returns one StepStats per member.

        A member whose solve failed on its own (a non-finite CR residual) has a
        non-zero ``status`` and did not advance; the others did, and no exception is
        raised for it.  Remove such a member before the next step.  Any other failure
        raises MphError."""
        n = len(self.members)
        arr = (StepStats * max(n, 1))()
        rc = lib().mph_batch_step(self.h, C.cast(arr, C.c_void_p))
        stats = [arr[k] for k in range(n)]
        if rc == abi.E_STATE and any(st.status != 0 for st in stats):
            return stats
        _check(rc)
        return stats

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().mph_batch_destroy(self.h)  # releases the members' membership
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
