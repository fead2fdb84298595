"""HIP path vs the CPU oracle on the GPU.

1. Bit for bit, against the oracle run with the GPU's dot-product tree (DOT_TREE):
   every other operation of the step is the reference's own arithmetic in both, so
   all fields, the multigrid hierarchy, the CR iteration counts and scalars must be
   bit-identical after every phase, and whole trajectories must stay identical.
2. Against the oracle in the reference's serial dot order (= the reference binary,
   pinned by tests/test_oracle_golden.py): the two differ only in the rounding of
   the CR inner products, so positions and velocities after many steps agree to a
   stated relative tolerance, and CR iteration counts agree.
Scenes are seeded (mphimplicit_amd.scenes) or the reference's own samples (inputs
from tests/golden/*.npz), small enough for the oracle to finish in seconds.
"""
import os

import numpy as np
import pytest

import pyoracle
from mphimplicit_amd import abi, files, scenes
from mphimplicit_amd.scene import MphError, Scene
from parity import diff, step_by_phase

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def golden_scene(name):
    fx = np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False)
    p = files.parse_data_text(str(fx["data_text"]), int(fx["dim"]))
    h = fx["in_hdr"]
    g = files.GridData(float(h[0]), float(h[1]), h[2:5].copy(), h[5:8].copy(), fx["in_type"].copy(),
                       fx["in_pos"].copy(), fx["in_vel"].copy())
    return p, g


def _with(make, **kw):
    """A scene with some parameters changed (paths the default scenes do not reach)."""
    def f():
        p, g = make()
        for k, v in kw.items():
            setattr(p, k, v)
        return p, g
    return f


SCENES = {
    "tiny2d": lambda: scenes.tiny_2d(),
    "tiny3d": lambda: scenes.tiny_3d(),
    "2d_40x40": lambda: scenes.config_2d(),
    "3d_4096": lambda: scenes.config_3d_drop(n=16, box=30, lift=4, steps=30),
    "cube3d_sample": lambda: golden_scene("cube3d"),
    "bridge2d_sample": lambda: golden_scene("bridge2d"),
    "glued2d_sample": lambda: golden_scene("glued2d"),
    # plain CR without the multigrid preconditioner (the reference built without
    # MULTIGRID_SOLVER)
    "3d_4096_nomg": _with(lambda: scenes.config_3d_drop(n=16, box=30, lift=4, steps=30), multigrid=0),
    # ~70 P-neighbours per particle: more than the LDS staging of C Lambda C^T holds
    "tiny3d_wideP": _with(lambda: scenes.tiny_3d(), radius_ratio_p=2.6),
    # ~400 fluid neighbours per particle: rows longer than the Galerkin LDS staging
    "3d_4096_wideV": _with(lambda: scenes.config_3d_drop(n=16, box=30, lift=4, steps=30), radius_ratio_a=4.6,
                           radius_ratio_v=4.6),
    # fluid filling the domain in x and y: neighbours through the periodic boundary
    "periodic3d": lambda: scenes.periodic_3d(),
    # 39,330 level-0 rows: level 0 takes the one-thread-per-row multigrid sweeps, the
    # coarser levels the lane-split ones (below 32768 rows), as on the 1M scene
    "2d_160k": lambda: scenes.config_2d(nx=400, ny=400, box=(440, 420)),
    # translating and rotating walls of both wall types (initializeWall, calculateWall)
    "tiny2d_moving": lambda: scenes.tiny_2d_moving(steps=5),
    "tiny3d_moving": lambda: scenes.tiny_3d_moving(steps=5),
}


def _pair(name, tol=None, mode=pyoracle.DOT_TREE):
    p, g = SCENES[name]()
    if tol is not None:
        p.cr_tolerance = tol
    print("[gpu parity] scene %s np=%d" % (name, g.np), flush=True)
    return Scene(p, g), pyoracle.OracleScene(p, g, mode)


@pytest.mark.parametrize("name", list(SCENES))
def test_init_sequence_bitwise(name):
    gpu, orc = _pair(name)
    for n in ["CellIndex", "CellParticle", "ParticleIndex", "NeighborInd", "NeighborIndP", "DensityA",
              "GravityCenter", "VolStrainP", "Position"]:
        m = diff(n, gpu.read(n), orc.read(n))
        assert m is None, m
    for n in ["N0a", "N0p", "Swa", "Swg", "Swp", "Swv", "R2g", "MaxRadius"]:
        assert gpu.scalar(n) == orc.scalar(n), n


@pytest.mark.parametrize("name", list(SCENES))
def test_phases_bitwise(name):
    gpu, orc = _pair(name)
    steps = 1 if name == "glued2d_sample" else 2
    bad = step_by_phase(gpu, orc, steps=steps)
    assert not bad, "\n".join(bad[:10])


@pytest.mark.parametrize("name,steps,tol", [("2d_40x40", 12, 1e-16), ("3d_4096", 8, 1e-16),
                                            ("cube3d_sample", 3, 1e-12), ("2d_160k", 2, 1e-16),
                                            ("tiny3d_moving", 5, 1e-16)])
def test_trajectory_bitwise(name, steps, tol):
    """mph_scene_step over several steps (incl. the first step's VTK diagnostics)."""
    gpu, orc = _pair(name, tol)
    for s in range(steps):
        a, b = gpu.step(1), orc.step(1)
        assert a.cr_iterations == b.cr_iterations, (s, a.cr_iterations, b.cr_iterations)
        assert (a.rr0, a.rr, a.rs0, a.rs) == (b.rr0, b.rr, b.rs0, b.rs), s
    for n in ["Position", "Velocity", "Force", "ParticleIndex", "NeighborInd", "CsrCofA", "VectorB",
              "DivergenceP", "PressureP", "VirialPressureAtParticle"]:
        m = diff(n, gpu.read(n), orc.read(n))
        assert m is None, m


def test_neighbor_search_paths():
    """The one-scan neighbour search sizes its scratch by the previous search's
    largest counts. Compressing the particles by 0.85 about the centre makes the
    lists longer than that, so the search falls back to the count and fill passes;
    the search after it fits again. Lists and runs stay bit-identical to the
    oracle's on every path."""
    gpu, orc = _pair("3d_4096")
    assert gpu.scalar("NeighborOneScan") == 0.0  # the first search has no counts to size by
    gpu.step(1), orc.step(1)
    assert gpu.scalar("NeighborOneScan") == 1.0
    pos = gpu.read("Position")
    assert pos.tobytes() == orc.read("Position").tobytes()
    xyz = pos.reshape(-1, 3)
    c = 0.5 * (xyz.min(axis=0) + xyz.max(axis=0))
    packed = (c + 0.85 * (xyz - c)).reshape(-1)
    for sc in (gpu, orc):
        sc.write("Position", packed)
    for expect in (0.0, 1.0):  # overflow: fill pass; then scratch sized for the denser lists
        for sc in (gpu, orc):
            sc.run_phase(abi.PHASE_CELLS)
            sc.run_phase(abi.PHASE_NEIGHBOR)
        assert gpu.scalar("NeighborOneScan") == expect
        for n in ["CellParticle", "NeighborInd", "NeighborIndP", "Position"]:
            m = diff(n, gpu.read(n), orc.read(n))
            assert m is None, m


@pytest.mark.parametrize("name,steps", [("2d_40x40", 20), ("3d_4096", 10), ("bridge2d_sample", 3)])
def test_against_reference_dot_order(name, steps):
    """GPU vs the reference's serial dot order: positions and velocities agree to
    1e-10 relative to the largest velocity (resp. 1e-12 of the domain), particle
    order and CR iteration counts are identical."""
    gpu, orc = _pair(name, 1e-16, pyoracle.DOT_SERIAL)
    for s in range(steps):
        a, b = gpu.step(1), orc.step(1)
        assert abs(a.cr_iterations - b.cr_iterations) <= 1, (s, a.cr_iterations, b.cr_iterations)
    assert np.array_equal(gpu.read("ParticleIndex"), orc.read("ParticleIndex"))
    vg, vo = gpu.read("Velocity"), orc.read("Velocity")
    xg, xo = gpu.read("Position"), orc.read("Position")
    vmax = np.abs(vo).max()
    assert np.abs(vg - vo).max() <= 1e-10 * vmax
    assert np.abs(xg - xo).max() <= 1e-12


def _fields(sc):
    sc.run_phase(abi.PHASE_DIAGNOSTICS)
    return {n: sc.read(n) for n in ("Velocity", "Position", "PressureP")}


@pytest.mark.parametrize("name,steps", [("2d_40x40", 10), ("3d_4096", 6)])
def test_mixed_precision_against_fp64(name, steps):
    """MPH_PRECISION_MIXED (FP32 particle and multigrid matrices inside FP64
    iterative refinement) against the FP64 oracle in the reference's serial order.

    Both stop at the same rule, ||b - A x||/||b|| < 1e-8, and a solution at that
    residual is itself only ~1e-7 from the converged one (the implicit system is
    stiff), so the bar is accuracy, not agreement: after several steps the mixed
    run's error against a converged FP64 run (oracle at rr/rr0 < 1e-28) must be at
    most twice the plain FP64 run's own error at 1e-8, per field, and below 1e-6 of
    the field's maximum (3e-9 for positions, in metres)."""
    p, g = SCENES[name]()
    p.cr_tolerance = 1e-16
    orc = pyoracle.OracleScene(p, g, pyoracle.DOT_SERIAL)
    pt = p.copy()
    pt.cr_tolerance = 1e-28
    tight = pyoracle.OracleScene(pt, g, pyoracle.DOT_SERIAL)
    pm = p.copy()
    pm.precision = abi.PRECISION_MIXED
    gpu = Scene(pm, g)
    for s in range(steps):
        a = gpu.step(1)
        orc.step(1)
        tight.step(1)
        assert a.rr / a.rr0 < 1e-16, (s, a.rr, a.rr0)
    assert gpu.scalar("RefinementSteps") >= 1
    assert np.array_equal(gpu.read("ParticleIndex"), tight.read("ParticleIndex"))
    fm, ff, ft = _fields(gpu), _fields(orc), _fields(tight)
    for n, bound in (("Velocity", 1e-6), ("PressureP", 1e-6), ("Position", 3e-9)):
        scale = 1.0 if n == "Position" else np.abs(ft[n]).max()
        em = np.abs(fm[n] - ft[n]).max() / scale
        ef = np.abs(ff[n] - ft[n]).max() / scale
        print("[mixed] %s %s: mixed err %.3e, fp64 err %.3e" % (name, n, em, ef), flush=True)
        assert em <= 2.0 * ef + 1e-12, (n, em, ef)
        assert em <= bound, (n, em)


@pytest.mark.parametrize("name,lane_rows", [("3d_4096", 0), ("3d_4096", 1 << 30), ("2d_160k", 1 << 30)])
def test_sweep_paths_bitwise(name, lane_rows):
    """The multigrid smoothing sweeps exist as one thread per row and as rows split
    over lanes (the default below 32768 rows).  The option LaneRowsMax moves every
    level of a scene to one form (0: one thread per row everywhere, as level 0 of
    the 1M 3D scene; 2^30: lanes everywhere): the step stays bit-identical to the
    oracle either way."""
    gpu, orc = _pair(name)
    gpu.set_option("LaneRowsMax", lane_rows)
    bad = step_by_phase(gpu, orc, steps=1, diagnostics=False)
    assert not bad, "\n".join(bad[:10])


def _diverging(name):
    """A scene whose implicit system is not finite: an infinite shear viscosity makes
    the harmonic-mean viscosities inf/inf = NaN, so A and b hold NaN while positions,
    cells and lists stay finite (the viscosity enters the implicit terms only)."""
    p, g = SCENES[name]()
    for t in range(4):
        p.shear_viscosity[t] = float("inf")
    return p, g


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64, abi.PRECISION_MIXED])
def test_nonfinite_residual_fails_fast(precision):
    """rr/rr0 < tol is false for NaN, so the reference would iterate to its cap of
    3*Nf; the solve reports MPH_E_STATE at the first check instead."""
    p, g = _diverging("tiny3d")
    p.precision = precision
    s = Scene(p, g)
    with pytest.raises(MphError) as e:
        s.step(1)
    assert e.value.code == abi.E_STATE and "non-finite" in str(e.value)
    s.close()
