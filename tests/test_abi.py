"""The C-ABI library on the host (no GPU needed): exports, file parsers, writers.

* every function declared in include/*.h is exported by libmph_hip.so;
* mph_read_data_file parses the reference's *.data files (the four samples and the
  synthetic scenes) to the same values as the Python parser;
* mph_read_grid_file reads back exactly what was written with %.17e;
* the VTK and PROF writers produce byte-identical files to the reference's own
  writeVtkFile / writeProfFile (fixtures written by the reference solver) when fed
  the same values (taken from the oracle, which matches the reference bit for bit).
"""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import pyoracle
from mphimplicit_amd import abi, files
from mphimplicit_amd.scene import lib, read_data_file

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(mph_\w+)\s*\(", text))
    return sorted(names)


def test_all_declared_symbols_exported():
    L = C.CDLL(os.path.join(ROOT, "mphimplicit_amd", "libmph_hip.so"))
    names = declared_functions()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def fixture(name):
    return np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False)


@pytest.mark.parametrize("name", ["cube3d", "squaredrop2d", "bridge2d", "glued2d", "tiny2d", "tiny3d"])
def test_data_parsers_agree(name, tmp_path):
    fx = fixture(name)
    path = tmp_path / "box.data"
    path.write_text(str(fx["data_text"]))
    dim = int(fx["dim"])
    a = read_data_file(str(path), dim)
    b = files.parse_data_text(str(fx["data_text"]), dim)
    for f, _ in abi.Params._fields_:
        va, vb = np.ctypeslib.as_array(getattr(a, f)) if hasattr(getattr(a, f), "_length_") else getattr(a, f), \
            np.ctypeslib.as_array(getattr(b, f)) if hasattr(getattr(b, f), "_length_") else getattr(b, f)
        assert np.array_equal(va, vb), f


def test_grid_round_trip(tmp_path):
    fx = fixture("tiny3d")
    h = fx["in_hdr"]
    g = files.GridData(float(h[0]), float(h[1]), h[2:5], h[5:8], fx["in_type"], fx["in_pos"], fx["in_vel"])
    path = str(tmp_path / "t.grid")
    files.write_grid_file(path, g)
    n = C.c_int()
    assert lib().mph_read_grid_header(path.encode(), C.byref(n)) == 0 and n.value == g.np
    t, sp = C.c_double(), C.c_double()
    dmin, dmax = np.zeros(3), np.zeros(3)
    typ = np.zeros(g.np, np.int32)
    pos, vel = np.zeros((g.np, 3)), np.zeros((g.np, 3))
    rc = lib().mph_read_grid_file(path.encode(), g.np, C.byref(t), C.byref(sp), dmin.ctypes.data, dmax.ctypes.data,
                                  typ.ctypes.data, pos.ctypes.data, vel.ctypes.data)
    assert rc == 0
    assert np.array_equal(typ, g.type) and np.array_equal(pos, g.pos) and np.array_equal(vel, g.vel)
    assert np.array_equal(dmin, g.domain_min) and np.array_equal(dmax, g.domain_max) and sp.value == g.spacing


def write_vtk(path, o):
    L = lib()
    arr = {n: o.read(n) for n in ["Property", "Position", "Velocity", "Force", "VolStrainP", "DivergenceP",
                                  "PressureP", "VirialPressureAtParticle", "NeighborCount"]}
    rc = L.mph_write_vtk_arrays(path.encode(), C.c_int(arr["Property"].size),
                                *[C.c_void_p(arr[n].ctypes.data) for n in ["Property", "Position", "Velocity",
                                                                          "Force", "VolStrainP", "DivergenceP",
                                                                          "PressureP", "VirialPressureAtParticle",
                                                                          "NeighborCount"]])
    assert rc == 0


def write_prof(path, o, g):
    L = lib()
    typ, pos, vel = o.read("Property"), o.read("Position"), o.read("Velocity")
    dmin = np.ascontiguousarray(g.domain_min, dtype=np.float64)
    dmax = np.ascontiguousarray(g.domain_max, dtype=np.float64)
    rc = L.mph_write_prof_arrays(path.encode(), C.c_double(o.scalar("Time")), C.c_int(typ.size),
                                 C.c_double(g.spacing), C.c_void_p(dmin.ctypes.data), C.c_void_p(dmax.ctypes.data),
                                 C.c_void_p(typ.ctypes.data), C.c_void_p(pos.ctypes.data),
                                 C.c_void_p(vel.ctypes.data))
    assert rc == 0


@pytest.mark.parametrize("name", ["tiny2d", "tiny3d"])
def test_writers_byte_identical(name, tmp_path):
    fx = fixture(name)
    dim = int(fx["dim"])
    p = files.parse_data_text(str(fx["data_text"]), dim)
    h = fx["in_hdr"]
    g = files.GridData(float(h[0]), float(h[1]), h[2:5].copy(), h[5:8].copy(), fx["in_type"].copy(),
                       fx["in_pos"].copy(), fx["in_vel"].copy())
    o = pyoracle.OracleScene(p, g, pyoracle.DOT_SERIAL)
    for s in range(int(fx["steps"])):
        tag = "step%03d" % s
        prof = str(tmp_path / ("%s.prof" % tag))
        write_prof(prof, o, g)
        assert open(prof, "rb").read() == fx["%s/file/ref.prof" % tag].tobytes(), tag
        o.step(1)  # every step of these fixtures is a VTK step
        vtk = str(tmp_path / ("%s.vtk" % tag))
        write_vtk(vtk, o)
        assert open(vtk, "rb").read() == fx["%s/file/ref.vtk" % tag].tobytes(), tag


def test_no_device_fails_loudly():
    """Without a usable HIP device the library reports an error: there is no CPU
    path behind the boundary.  (Skipped where a GPU is present.)"""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = lib()
    h = C.c_void_p()
    assert L.mph_batch_create(0, C.byref(h)) < 0
    assert L.mph_last_error().decode() != ""
    fx = fixture("tiny3d")
    p = files.parse_data_text(str(fx["data_text"]), 3)
    hdr = fx["in_hdr"]
    g = files.GridData(float(hdr[0]), float(hdr[1]), hdr[2:5].copy(), hdr[5:8].copy(), fx["in_type"].copy(),
                       fx["in_pos"].copy(), fx["in_vel"].copy())
    from mphimplicit_amd.scene import MphError, Scene

    with pytest.raises(MphError):
        Scene(p, g)
