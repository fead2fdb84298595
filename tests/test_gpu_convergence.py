"""The reference's CONVERGENCE_CHECK record on the GPU (scene option ConvergenceCheck).

With the option on, each FP64 solve keeps rr and r.s of every CR iteration that did
not stop (source/main.cpp:4348-4351) and, after the loop, r.r and r.s of the true
residual b - A x (main.cpp:4383-4397).  tests/test_oracle_convergence.py pins the
oracle's record to the reference built with the switch, line for line.  Here the HIP
record must equal the oracle's bit for bit in the GPU's dot order (DOT_TREE), in every
CR stop mode, and the option must leave the trajectory unchanged.
"""
import numpy as np
import pytest

import pyoracle
from mphimplicit_amd import abi
from mphimplicit_amd.scene import Batch, MphError, Scene
from test_gpu_parity import SCENES
from test_oracle_convergence import load_case

pytestmark = pytest.mark.gpu

A, F = abi.OPERATOR_ASSEMBLED, abi.OPERATOR_FACTORED


def _bits(v):
    return np.float64(v).view(np.int64)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("name,steps", [("tiny2d", 3), ("tiny3d", 3), ("tiny3d_moving", 2)])
def test_record_bitwise_against_oracle(name, steps, mode):
    print("[convergence] %s CrStop %d" % (name, mode), flush=True)
    _, p, g = load_case(name)
    p.operator_form = A
    o = pyoracle.OracleScene(p, g, pyoracle.DOT_TREE)
    o.set_option("ConvergenceCheck", 1)
    s = Scene(p, g)
    s.set_option("CrStop", mode)
    s.set_option("ConvergenceCheck", 1)
    for k in range(steps):
        st, so = s.step(1), o.step(1)
        assert st.cr_iterations == so.cr_iterations, k
        h, ho = s.read("ConvergenceHistory"), o.read("ConvergenceHistory")
        assert h.size == 2 * st.cr_iterations
        assert h.tobytes() == ho.tobytes(), (k, np.flatnonzero(h != ho)[:8])
        for n in ("rr0", "rs0", "rr", "rs", "CheckRR", "CheckRS"):
            assert _bits(s.scalar(n)) == _bits(o.scalar(n)), (k, n, s.scalar(n), o.scalar(n))
    s.close()
    o.close()


@pytest.mark.parametrize("name,form", [("3d_4096_nomg", A), ("3d_4096", F), ("2d_40x40", F)])
def test_record_leaves_the_trajectory(name, form):
    """The same steps with and without the record: identical positions, velocities and
    CR scalars.  Without multigrid the record's r.s after the loop is taken against b
    (main.cpp:4391); the assembled scene is also checked against the oracle bit for bit."""
    print("[convergence] %s form %d" % (name, form), flush=True)
    runs = []
    for check in (0, 1):
        p, g = SCENES[name]()
        p.operator_form = form
        s = Scene(p, g)
        s.set_option("ConvergenceCheck", check)
        stats = [s.step(1) for _ in range(2)]
        runs.append(([(t.cr_iterations, t.rr0, t.rr, t.rs0, t.rs) for t in stats],
                     s.read("Position"), s.read("Velocity"), s.read("ConvergenceHistory"),
                     s.scalar("CheckRR"), s.scalar("CheckRS")))
        s.close()
    (a, pa, va, ha, _, _), (b, pb, vb, hb, crr, crs) = runs
    assert a == b
    assert pa.tobytes() == pb.tobytes() and va.tobytes() == vb.tobytes()
    assert ha.size == 0 and hb.size == 2 * b[-1][0]
    h = hb.reshape(-1, 2)
    tol = SCENES[name]()[0].cr_tolerance
    # every kept row is an iteration that did not stop; the loop stopped after the last
    assert np.all(h[:, 0] / b[-1][1] >= tol) and b[-1][2] / b[-1][1] < tol
    # the true residual is the loop's residual up to rounding
    assert abs(crr - b[-1][2]) <= 1e-6 * b[-1][2]
    if form == A:
        p, g = SCENES[name]()
        o = pyoracle.OracleScene(p, g, pyoracle.DOT_TREE)
        o.set_option("ConvergenceCheck", 1)
        o.step(2)
        assert o.read("ConvergenceHistory").tobytes() == hb.tobytes()
        assert _bits(o.scalar("CheckRR")) == _bits(crr) and _bits(o.scalar("CheckRS")) == _bits(crs)
        o.close()


def test_option_refused_where_it_has_no_loop():
    """Mixed precision has no FP64 CR loop to record; a batch solves its members
    itself; a value other than 0 or 1 is out of range."""
    p, g = SCENES["tiny2d"]()
    p.precision = abi.PRECISION_MIXED
    s = Scene(p, g)
    with pytest.raises(MphError):
        s.set_option("ConvergenceCheck", 1)
    s.close()
    p, g = SCENES["tiny2d"]()
    s = Scene(p, g)
    with pytest.raises(MphError):
        s.set_option("ConvergenceCheck", 2)
    s.set_option("ConvergenceCheck", 1)
    b = Batch()
    with pytest.raises(MphError):
        b.add(s)
    s.set_option("ConvergenceCheck", 0)
    b.add(s)
    with pytest.raises(MphError):
        s.set_option("ConvergenceCheck", 1)
    b.remove(s)
    b.close()
    s.close()
