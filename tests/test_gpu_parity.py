"""HIP path vs the CPU oracle, bit for bit.

The oracle runs with the GPU's dot-product tree (DOT_TREE); every other operation of
the step is the reference's own arithmetic in both, so all fields, the multigrid
hierarchy, the CR iteration counts and scalars must be bit-identical after every
phase.  Scenes are seeded and small enough for the oracle to finish in seconds.
"""
import sys

import numpy as np
import pytest

import pyoracle
from mphimplicit_amd import abi, scenes
from mphimplicit_amd.scene import Scene
from parity import step_by_phase, diff

pytestmark = pytest.mark.gpu

SCENES = {
    "tiny2d": lambda: scenes.tiny_2d(),
    "tiny3d": lambda: scenes.tiny_3d(),
    "2d_40x40": lambda: scenes.config_2d(),
}


def _pair(name, tol=None):
    p, g = SCENES[name]()
    if tol is not None:
        p.cr_tolerance = tol
    print("[gpu parity] scene %s np=%d" % (name, g.np), flush=True)
    gpu = Scene(p, g)
    orc = pyoracle.OracleScene(p, g, pyoracle.DOT_TREE)
    return gpu, orc


@pytest.mark.parametrize("name", list(SCENES))
def test_init_sequence_bitwise(name):
    gpu, orc = _pair(name)
    for n in ["CellIndex", "CellParticle", "ParticleIndex", "NeighborInd", "NeighborIndP", "DensityA",
              "GravityCenter", "VolStrainP", "Position"]:
        assert diff(n, gpu.read(n), orc.read(n)) is None, diff(n, gpu.read(n), orc.read(n))
    for n in ["N0a", "N0p", "Swa", "Swg", "Swp", "Swv", "R2g", "MaxRadius"]:
        assert gpu.scalar(n) == orc.scalar(n), n


@pytest.mark.parametrize("name", list(SCENES))
def test_phases_bitwise(name):
    gpu, orc = _pair(name)
    bad = step_by_phase(gpu, orc, steps=2)
    assert not bad, "\n".join(bad[:10])
