"""Batched executor (mph_batch_*) on the GPU: a scene's results are the same alone
and in a batch, bit for bit.

Scenes of different dimensions, sizes, multigrid depths and physics (surface
tension, two fluids) are advanced together; the same inputs are advanced alone
with mph_scene_step; every step's CR iteration count and residual scalars and the
final fields must be identical.  Members that stop early (end_time), leave the
batch or join later must not change the others.  Scenes of the factored operator
form share a batch with assembled ones (the products split into one launch set per
form).  The single-scene path itself is pinned to the oracle by test_gpu_parity.py
and test_gpu_factored.py.
"""
import numpy as np
import pytest

from mphimplicit_amd import abi
from mphimplicit_amd.scene import Batch, Scene
from test_gpu_parity import SCENES

pytestmark = pytest.mark.gpu

FIELDS = ["Position", "Velocity", "ParticleIndex", "NeighborInd", "CsrCofA", "VectorB", "PressureP"]
MIX = ["tiny2d", "2d_40x40", "3d_4096", "cube3d_sample", "bridge2d_sample", "glued2d_sample", "tiny3d"]


def make(name, precision, tol=1e-12, end_steps=None, form=abi.OPERATOR_ASSEMBLED):
    p, g = SCENES[name]()
    p.cr_tolerance = tol
    p.precision = precision
    p.operator_form = form
    if end_steps is not None:
        p.end_time = g.time + (end_steps - 1) * p.dt
    else:
        p.end_time = 1e30
    return p, g


def alone(name, precision, steps, end_steps=None, form=abi.OPERATOR_ASSEMBLED):
    p, g = make(name, precision, end_steps=end_steps, form=form)
    s = Scene(p, g)
    stats = [s.step(1) for _ in range(steps)]
    return s, stats


def same(a, b):
    for n in FIELDS:
        if n == "CsrCofA" and not a.scalar("AssembledBlocks"):
            n = "MGCofA_L0"  # factored scenes build no blocks; compare their level 0 instead
        x, y = a.read(n), b.read(n)
        assert x.shape == y.shape and np.array_equal(x.view(np.uint8), y.view(np.uint8)), n


def forms_of(kind):
    """Operator form per member: all assembled, all factored, or alternating (a
    batch whose launches split between the two forms)."""
    A, F = abi.OPERATOR_ASSEMBLED, abi.OPERATOR_FACTORED
    return [A if kind == "assembled" else F if kind == "factored" else (A, F)[k % 2] for k in range(len(MIX))]


@pytest.mark.parametrize("kind", ["assembled", "alternating"])
@pytest.mark.parametrize("precision", [abi.PRECISION_FP64, abi.PRECISION_MIXED])
def test_batch_equals_alone(precision, kind):
    steps = 3
    forms = forms_of(kind)
    print("[batch] precision %d, %d scenes, %s operator forms" % (precision, len(MIX), kind), flush=True)
    members = [Scene(*make(n, precision, form=f)) for n, f in zip(MIX, forms)]
    with Batch() as b:
        for s in members:
            b.add(s)
        assert len(b) == len(MIX)
        hist = [b.step() for _ in range(steps)]
    # members build their lists with the one-scan search too (scratch sized at creation)
    assert all(s.scalar("NeighborOneScan") == 1.0 for s in members)
    # the batch's wall time per group of phases, the same in every member's record
    for st in hist:
        split = [st[0].ms_phase[k] for k in (abi.PHASE_NEIGHBOR, abi.PHASE_MULTIGRID, abi.PHASE_SOLVE,
                                             abi.PHASE_ADVANCE)]
        assert all(v > 0 for v in split) and sum(split) <= st[0].ms_total * (1 + 1e-9)
        assert all(list(x.ms_phase) == list(st[0].ms_phase) for x in st)
    for k, name in enumerate(MIX):
        ref, rstats = alone(name, precision, steps, form=forms[k])
        for t in range(steps):
            a, r = hist[t][k], rstats[t]
            assert a.steps_done == 1
            assert a.cr_iterations == r.cr_iterations, (name, t, a.cr_iterations, r.cr_iterations)
            assert (a.rr0, a.rr, a.rs0, a.rs) == (r.rr0, r.rr, r.rs0, r.rs), (name, t)
        same(members[k], ref)
        assert members[k].scalar("Time") == ref.scalar("Time")


def test_batch_membership_changes():
    """A member that reaches its end time idles; a member removed after one step and
    one added after one step still match their solo runs."""
    prec = abi.PRECISION_FP64
    stop = Scene(*make("2d_40x40", prec, end_steps=1))    # runs one step only
    stay = Scene(*make("3d_4096", prec))
    leave = Scene(*make("cube3d_sample", prec))
    late = Scene(*make("bridge2d_sample", prec))
    with Batch() as b:
        for s in (stop, stay, leave):
            b.add(s)
        st = b.step()
        assert [x.steps_done for x in st] == [1, 1, 1]
        b.remove(leave)
        b.add(late)
        st = b.step()
        assert [x.steps_done for x in st] == [0, 1, 1]
        st = b.step()
    same(stop, alone("2d_40x40", prec, 3, end_steps=1)[0])
    same(stay, alone("3d_4096", prec, 3)[0])
    same(leave, alone("cube3d_sample", prec, 1)[0])
    same(late, alone("bridge2d_sample", prec, 2)[0])


def test_member_cannot_be_destroyed():
    """mph_scene_destroy refuses a scene that a batch still holds (MPH_E_STATE)."""
    from mphimplicit_amd.scene import MphError

    s = Scene(*make("tiny2d", abi.PRECISION_FP64))
    b = Batch()
    b.add(s)
    with pytest.raises(MphError) as e:
        s.close()
    assert e.value.code == -5
    b.remove(s)
    s.close()
    b.close()


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64, abi.PRECISION_MIXED])
def test_nonfinite_member_fails_alone(precision):
    """A member whose system is not finite leaves the lockstep solve at once with
    status MPH_E_STATE and does not advance; the other members step as alone."""
    from test_gpu_parity import _diverging
    names = ["tiny2d", "tiny3d", "2d_40x40"]
    p, g = _diverging("tiny3d")
    p.precision = precision
    p.end_time = 1e30
    bad = Scene(p, g)
    members, alone_runs = [], []
    for n in names:
        mp, mg = make(n, precision)
        members.append(Scene(mp, mg))
        alone_runs.append(alone(n, precision, 2)[0])
    with Batch() as b:
        for s in members[:1] + [bad] + members[1:]:
            b.add(s)
        st = b.step()
        assert st[1].status == abi.E_STATE and st[1].steps_done == 0
        assert all(x.status == 0 and x.steps_done == 1 for k, x in enumerate(st) if k != 1)
        b.remove(bad)
        st = b.step()
        assert all(x.status == 0 and x.steps_done == 1 for x in st)
    for s, a in zip(members, alone_runs):
        same(s, a)
