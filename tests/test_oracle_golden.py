"""The CPU oracle against the reference's own results, bit for bit.

Fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py) hold the
SHA-256 digest of every array the unmodified reference solver produced after each
phase of each time step, for the reference's four sample cases (Cube 3D; the
SquareDropSkip, Bridge (surface tension) and GluedLightSquare (two fluids) 2D cases)
and two seeded synthetic scenes.  The oracle in serial dot order must reproduce all
of them exactly: cell order, neighbour lists, explicit fields, the CSR A and C,
b, P, the multigrid hierarchy, M^-1 applied to a fixed vector, CR iteration counts,
rr0/rr/rs0/rs, the solved velocities and the output-step diagnostics.
"""
import hashlib
import os

import numpy as np
import pytest

import pyoracle
from mphimplicit_amd import abi, files

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["tiny2d", "tiny3d", "squaredrop2d", "bridge2d", "glued2d", "cube3d", "tiny2d_moving", "tiny3d_moving"]

# probe array -> (phase after which it is read, oracle field)
AFTER = {
    abi.PHASE_KINEMATICS: {"Position_kin": "Position", "Velocity_kin": "Velocity"},
    abi.PHASE_CELLS: {
        "ParticleIndex": "ParticleIndex", "CellIndex": "CellIndex", "CellParticle": "CellParticle",
        "CellFluidParticleBegin": "CellFluidParticleBegin", "CellFluidParticleEnd": "CellFluidParticleEnd",
        "CellWallParticleBegin": "CellWallParticleBegin", "CellWallParticleEnd": "CellWallParticleEnd",
        "Position_sort": "Position", "Velocity_sort": "Velocity", "Property_sort": "Property",
    },
    abi.PHASE_NEIGHBOR: {n: n for n in ["NeighborCount", "NeighborFluidCount", "NeighborCountP", "NeighborPtr",
                                        "NeighborPtrP", "NeighborInd", "NeighborIndP"]},
    abi.PHASE_EXPLICIT: dict({n: n for n in ["DensityA", "GravityCenter", "VolStrainP", "PressureA", "Mass",
                                              "Kappa", "Lambda", "Mu", "Force"]}, VelocityExpl="Velocity"),
    abi.PHASE_MATRIX_A: {"CsrPtrA": "CsrPtrA", "CsrIndA": "CsrIndA", "CsrCofA_visc": "CsrCofA",
                         "VectorB_visc": "VectorB"},
    abi.PHASE_MATRIX_C: {n: n for n in ["CsrPtrC", "CsrIndC", "CsrCofC", "VectorP"]},
    abi.PHASE_MULTIPLY_C: {"CsrCofA": "CsrCofA", "VectorB": "VectorB"},
    abi.PHASE_MULTIGRID: {"InvDiagA": "InvDiagA"},
    abi.PHASE_SOLVE: {"VelocitySolved": "Velocity"},
    abi.PHASE_DIAGNOSTICS: {"DivergenceP": "DivergenceP", "PressureP": "PressureP",
                            "VirialPressureAtParticle": "VirialPressureAtParticle",
                            "VirialStressAtParticle": "VirialStressAtParticle", "ForceVtk": "Force",
                            "NeighborCountVtk": "NeighborCount"},
    abi.PHASE_ADVANCE: {"Position_end": "Position", "Velocity_end": "Velocity"},
}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load_case(name):
    fx = np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False)
    dim = int(fx["dim"])
    params = files.parse_data_text(str(fx["data_text"]), dim)
    h = fx["in_hdr"]
    grid = files.GridData(float(h[0]), float(h[1]), h[2:5].copy(), h[5:8].copy(), fx["in_type"].copy(),
                          fx["in_pos"].copy(), fx["in_vel"].copy())
    return fx, params, grid


def check(fx, tag, probe_name, arr):
    key = "%s/%s/sha" % (tag, probe_name)
    assert key in fx, "fixture lacks %s" % key
    assert arr.size == int(fx["%s/%s/len" % (tag, probe_name)]), "%s %s: size" % (tag, probe_name)
    assert sha(arr) == str(fx[key]), "%s %s differs from the reference" % (tag, probe_name)


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference(name):
    fx, params, grid = load_case(name)
    o = pyoracle.OracleScene(params, grid, pyoracle.DOT_SERIAL)
    consts = fx["init/consts"]
    for k, s in enumerate(["N0a", "N0p", "Swa", "Swg", "Swp", "Swv", "R2g", "MaxRadius"]):
        assert o.scalar(s) == consts[k], s
    # the init sequence (cell sort, neighbour search, densities) of main.cpp:369-375
    check(fx, "init", "ParticleIndex", o.read("ParticleIndex"))
    check(fx, "init", "CellIndex", o.read("CellIndex"))
    check(fx, "init", "NeighborInd", o.read("NeighborInd"))
    check(fx, "init", "Position_sorted", o.read("Position"))
    steps = int(fx["steps"])
    for s in range(steps):
        tag = "step%03d" % s
        for ph in range(abi.PHASE_KINEMATICS, abi.PHASE_ADVANCE + 1):
            if ph == abi.PHASE_DIAGNOSTICS and "%s/DivergenceP/sha" % tag not in fx:
                continue
            o.run_phase(ph)
            for pn, on in AFTER[ph].items():
                check(fx, tag, pn, o.read(on))
            if ph == abi.PHASE_MULTIGRID:
                depth = int(fx["%s/MultiGridDepth" % tag][0])
                assert int(o.scalar("MultiGridDepth")) == depth
                for l in range(depth):
                    check(fx, tag, "MGCofA_L%d" % l, o.read("MGCofA_L%d" % l))
                    check(fx, tag, "MGInvD_L%d" % l, o.read("MGInvD_L%d" % l))
                check(fx, tag, "PrecondOut", o.precondition(fx["%s/PrecondIn" % tag]))
            if ph == abi.PHASE_SOLVE:
                cr = fx["%s/CR" % tag]
                got = np.array([o.scalar(n) for n in ["CRIterations", "rr0", "rr", "rs0", "rs"]])
                assert np.array_equal(got.view(np.int64), cr.view(np.int64)), (got, cr)
    o.close()
