"""The CR loop's stop test on the device (scene option CrStop).

The reference tests rr/rr0 < tol on the host before every iteration
(source/main.cpp:4344-4346). The three modes here make the same decision:

- CrStop 0 reads rr back for every iteration, as the reference does.
- CrStop 1 (the default) moves the test into the r.r reduction on the device. After
  the stop, a flag turns every kernel of the one iteration still queued into a no-op.
  The host enqueues iteration k+1 before it waits for iteration k.
- CrStop 2 does the same, but replays iterations 1, 2, ... from two captured HIP
  graphs, one per ping-pong parity of the V-cycle.

The arithmetic is identical in all three modes, so trajectories, CR scalars and
iteration counts must be bit-identical across them. test_gpu_parity.py checks the
default mode against the oracle bit for bit.
"""
import numpy as np
import pytest

from mphimplicit_amd import abi
from mphimplicit_amd.scene import MphError, Scene
from test_gpu_parity import SCENES, _diverging

pytestmark = pytest.mark.gpu

A, F = abi.OPERATOR_ASSEMBLED, abi.OPERATOR_FACTORED


def _run(name, form, mode, steps, tol=1e-16, maxit=0, timing=0):
    p, g = SCENES[name]()
    p.operator_form = form
    p.cr_tolerance = tol
    p.max_iterations = maxit
    s = Scene(p, g)
    s.set_option("CrStop", mode)
    if timing:
        s.set_timing(timing)
    stats = []
    for _ in range(steps):
        st = s.step(1)
        stats.append((st.cr_iterations, st.rr0, st.rr, st.rs0, st.rs))
    fields = {n: s.read(n) for n in ("Position", "Velocity", "ParticleIndex")}
    graphs = {n: s.scalar(n) for n in ("CrGraphLaunches", "CrCaptures", "CrInstantiations", "CrCaptureMs",
                                       "CrGraphError")}
    s.close()
    return stats, fields, graphs


def _same(a, b, what):
    sa, fa, _ = a
    sb, fb, _ = b
    assert sa == sb, (what, sa, sb)
    for n in fa:
        assert fa[n].tobytes() == fb[n].tobytes(), (what, n)


@pytest.mark.parametrize("name,form,steps", [("2d_40x40", A, 4), ("3d_4096", F, 3), ("3d_4096", A, 2),
                                             ("3d_4096_nomg", F, 2), ("periodic3d", F, 2), ("tiny3d", A, 3),
                                             ("2d_160k", F, 2), ("tiny3d_moving", F, 3)])
def test_stop_modes_bitwise(name, form, steps):
    print("[crstop] %s form %d, %d steps" % (name, form, steps), flush=True)
    host = _run(name, form, 0, steps)
    dev = _run(name, form, 1, steps)
    graph = _run(name, form, 2, steps)
    _same(host, dev, "device stop")
    _same(host, graph, "graphs")
    g = graph[2]
    print("[crstop] %s: iterations %s, graphs %s" % (name, [s[0] for s in host[0]], g), flush=True)
    assert g["CrGraphError"] == 0
    assert dev[2]["CrGraphLaunches"] == 0
    # the variant exploration of the first V-cycles runs eagerly; later solves replay
    if sum(s[0] for s in host[0][1:]) > steps:
        assert g["CrGraphLaunches"] > 0 and g["CrCaptures"] >= 2


@pytest.mark.parametrize("maxit", [1, 2, 5])
def test_stop_modes_iteration_cap(maxit):
    """max_iterations caps the loop before the tolerance is met: every mode stops after
    the same number of iterations with the same iterate."""
    runs = [_run("3d_4096", F, m, 3, maxit=maxit) for m in (0, 1, 2)]
    for st in runs[0][0]:
        assert st[0] == maxit
    _same(runs[0], runs[1], "device stop")
    _same(runs[0], runs[2], "graphs")


def test_stop_modes_zero_iterations():
    """A tolerance the initial residual already meets: no iteration runs, in any mode."""
    runs = [_run("tiny3d", A, m, 2, tol=1e30) for m in (0, 1, 2)]
    assert all(st[0] == 0 for st in runs[0][0])
    _same(runs[0], runs[1], "device stop")
    _same(runs[0], runs[2], "graphs")


def test_stop_modes_with_launch_timing():
    """Per-launch SpMV timing (events around every product) keeps the device stop but
    launches the iterations eagerly: same bits, no graph launches."""
    plain = _run("3d_4096", F, 2, 2)
    timed = _run("3d_4096", F, 2, 2, timing=2)
    _same(plain, timed, "timed")
    assert timed[2]["CrGraphLaunches"] == 0


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_stop_modes_nonfinite(mode):
    """A non-finite system fails fast in every mode, and the scene still reports it
    (the stop flag is off again after the failed solve)."""
    p, g = _diverging("tiny3d")
    s = Scene(p, g)
    s.set_option("CrStop", mode)
    with pytest.raises(MphError) as e:
        s.step(1)
    assert e.value.code == abi.E_STATE and "non-finite" in str(e.value)
    s.close()


def test_option_checked():
    p, g = SCENES["tiny3d"]()
    s = Scene(p, g)
    with pytest.raises(MphError):
        s.set_option("CrStop", 3)
    s.close()
