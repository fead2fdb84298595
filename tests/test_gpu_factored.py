"""The factored particle matrix (MPH_OPERATOR_FACTORED) on the GPU.

The factored form rebuilds the product A x from one scalar per block slot and per
RadiusP pair plus the positions. Its algebra is pinned on the CPU against the
oracle's assembled matrix by test_oracle_math.py::test_factored_form_identity.
The same terms are associated differently, so the form is not bit-identical. The
bar is therefore the floating-point one:

1. Product. After one step of identical inputs, the assembled matrices of an
   ASSEMBLED and a FACTORED scene are bit-identical. Their products with a random
   vector agree normwise: |y_f - y_a| <= 1e-11 max(|A| |x|), with |A| from the
   oracle's CSR values. With the FP32-stored values of mixed precision the bound is
   1e-6, since both forms round the same matrix to FP32.
2. Fields. After several steps solved to a tight residual, velocity and pressure
   agree with the oracle to 1e-8 of the field's maximum, the stated FP64 tolerance,
   and positions to 1e-11 m. Particle order is identical and CR iteration counts
   differ by at most one.
3. Mixed precision with the factored form meets the accuracy bar of the assembled
   mixed test: at most twice the plain FP64 run's own error against a converged
   oracle.
"""
import numpy as np
import pytest

import pyoracle
from mphimplicit_amd import abi
from mphimplicit_amd.scene import MphError, Scene
from test_gpu_parity import SCENES, _fields

pytestmark = pytest.mark.gpu

PRODUCT_SCENES = ["tiny2d", "tiny3d", "2d_40x40", "3d_4096", "cube3d_sample", "bridge2d_sample", "glued2d_sample",
                  "3d_4096_nomg", "tiny3d_wideP", "3d_4096_wideV", "periodic3d"]


def _scene(name, form, precision=abi.PRECISION_FP64, tol=None):
    p, g = SCENES[name]()
    p.operator_form = form
    p.precision = precision
    if tol is not None:
        p.cr_tolerance = tol
    return Scene(p, g), p, g


def _abs_product(orc, x):
    """|A| |x| from the oracle's CSR matrix."""
    ptr, ind, val = orc.read("CsrPtrA"), orc.read("CsrIndA"), orc.read("CsrCofA")
    rows = np.repeat(np.arange(ptr.size - 1), np.diff(ptr))
    out = np.zeros(ptr.size - 1)
    np.add.at(out, rows, np.abs(val) * np.abs(x[ind]))
    return out


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64, abi.PRECISION_MIXED])
@pytest.mark.parametrize("name", PRODUCT_SCENES)
def test_factored_product(name, precision):
    print("[factored] product %s precision %d" % (name, precision), flush=True)
    a, p, g = _scene(name, abi.OPERATOR_ASSEMBLED, precision)
    f, _, _ = _scene(name, abi.OPERATOR_FACTORED, precision)
    orc = pyoracle.OracleScene(p, g, pyoracle.DOT_TREE)
    a.step(1)
    f.step(1)
    orc.step(1)
    # the matrices are those of the same particles in the same order
    assert np.array_equal(a.read("ParticleIndex"), f.read("ParticleIndex"))
    if f.scalar("AssembledBlocks"):  # level 0 too narrow for the factored build
        assert np.array_equal(a.read("CsrCofA").view(np.uint64), f.read("CsrCofA").view(np.uint64))
    else:  # multigrid level 0 and InvDiagA built from the factored scalars
        with pytest.raises(MphError):
            f.read("CsrCofA")
        for n in ("MGCofA_L0", "InvDiagA") if p.multigrid else ():
            ra, rf = a.read(n), f.read(n)
            err = np.abs(rf - ra).max() / np.abs(ra).max()
            print("[factored] %s %s: max rel diff %.3e" % (name, n, err), flush=True)
            assert ra.shape == rf.shape and err <= 1e-12, (n, err)
    nf = int(a.scalar("FluidCount"))
    rng = np.random.default_rng(11)
    x = rng.standard_normal(3 * nf)
    scale = _abs_product(orc, x).max()
    bound = 1e-6 if precision == abi.PRECISION_MIXED else 1e-11
    for prec in sorted({abi.PRECISION_FP64, precision}):
        ya, yf = a.apply_matrix(x, prec), f.apply_matrix(x, prec)
        err = np.abs(yf - ya).max() / scale
        print("[factored] %s values %s: max|yf-ya| / max(|A||x|) = %.3e" % (name, "fp32" if prec else "fp64", err),
              flush=True)
        assert err <= (1e-11 if prec == abi.PRECISION_FP64 else bound), err
    # the FP64 assembled product against the oracle's CSR matrix (numpy order)
    ptr, ind, val = orc.read("CsrPtrA"), orc.read("CsrIndA"), orc.read("CsrCofA")
    rows = np.repeat(np.arange(ptr.size - 1), np.diff(ptr))
    yo = np.zeros(ptr.size - 1)
    np.add.at(yo, rows, val * x[ind])
    assert np.abs(a.apply_matrix(x) - yo).max() <= 1e-12 * scale


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64, abi.PRECISION_MIXED])
@pytest.mark.parametrize("name", PRODUCT_SCENES + ["2d_160k"])
def test_cell_staged_pass_bitwise(name, precision):
    """The FP64 second pass exists as the chunk kernel (every gather through the
    caches) and as the cell-staged kernel (the row cell's visited cells in LDS, the
    default).  Same entries, lanes and order of summation: the products are
    bit-identical, and so are whole steps (the CR scalars and velocities) with either
    kernel, in plain FP64 and in mixed precision (whose outer residuals take the FP64
    pass; its FP32 products are the chunk kernel's either way); the periodic scene
    takes the wrapped displacements."""
    outs = []
    precs = (abi.PRECISION_FP64, abi.PRECISION_MIXED) if precision == abi.PRECISION_MIXED else (abi.PRECISION_FP64,)
    for cells in (0, 1):  # chunk kernel, then the cell-staged one
        p, g = SCENES[name]()
        p.operator_form = abi.OPERATOR_FACTORED
        p.precision = precision
        p.cr_tolerance = 1e-16
        s = Scene(p, g)
        s.set_option("FactoredCells", cells)
        st = s.step(1)
        nf = int(s.scalar("FluidCount"))
        x = np.random.default_rng(5).standard_normal(3 * nf)
        outs.append(([s.apply_matrix(x, q) for q in precs], st.cr_iterations, st.rr, s.read("Velocity")))
        if cells == 0:  # the cell kernel without the staged places of this step: the table search
            s.set_option("FactoredCells", 1)
            ys = s.apply_matrix(x)
        s.close()
    y0, it0, rr0, v0 = outs[0]
    assert np.array_equal(y0[0].view(np.int64), ys.view(np.int64)), np.abs(y0[0] - ys).max()
    for k, (y1, it1, rr1, v1) in enumerate(outs[1:]):
        for a, b in zip(y0, y1):
            assert np.array_equal(a.view(np.int64), b.view(np.int64)), (k, np.abs(a - b).max())
        assert it0 == it1 and rr0 == rr1, k
        assert np.array_equal(v0.view(np.int64), v1.view(np.int64)), k


# scenes whose solves converge to rr/rr0 < 1e-24 (glued2d_sample stops at the
# reference's iteration cap 3*nf near 2e-9, where even two dot orders of the oracle
# differ by 4e-7: its operator is checked by the product test only)
@pytest.mark.parametrize("name,steps", [("2d_40x40", 10), ("3d_4096", 6), ("bridge2d_sample", 3),
                                        ("cube3d_sample", 3), ("tiny3d_wideP", 4), ("3d_4096_wideV", 2),
                                        ("periodic3d", 5)])
def test_factored_fields_against_oracle(name, steps):
    print("[factored] fields %s, %d steps" % (name, steps), flush=True)
    gpu, p, g = _scene(name, abi.OPERATOR_FACTORED, tol=1e-24)
    orc = pyoracle.OracleScene(p, g, pyoracle.DOT_SERIAL)
    for s in range(steps):
        a, b = gpu.step(1), orc.step(1)
        assert abs(a.cr_iterations - b.cr_iterations) <= 1, (s, a.cr_iterations, b.cr_iterations)
    assert np.array_equal(gpu.read("ParticleIndex"), orc.read("ParticleIndex"))
    fg, fo = _fields(gpu), _fields(orc)
    for n in ("Velocity", "PressureP"):
        err = np.abs(fg[n] - fo[n]).max() / np.abs(fo[n]).max()
        print("[factored] %s %s: rel err %.3e" % (name, n, err), flush=True)
        assert err <= 1e-8, (n, err)
    assert np.abs(fg["Position"] - fo["Position"]).max() <= 1e-11


@pytest.mark.parametrize("name,steps", [("2d_40x40", 10), ("3d_4096", 6)])
def test_factored_mixed_against_fp64(name, steps):
    """As test_gpu_parity.py::test_mixed_precision_against_fp64, with the factored
    form in both the FP32 inner solves and the FP64 outer residual."""
    p, g = SCENES[name]()
    p.cr_tolerance = 1e-16
    orc = pyoracle.OracleScene(p, g, pyoracle.DOT_SERIAL)
    pt = p.copy()
    pt.cr_tolerance = 1e-28
    tight = pyoracle.OracleScene(pt, g, pyoracle.DOT_SERIAL)
    pm = p.copy()
    pm.precision = abi.PRECISION_MIXED
    pm.operator_form = abi.OPERATOR_FACTORED
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
        print("[factored mixed] %s %s: mixed err %.3e, fp64 err %.3e" % (name, n, em, ef), flush=True)
        assert em <= 2.0 * ef + 1e-12, (n, em, ef)
        assert em <= bound, (n, em)


def test_factored_state_checks():
    """apply_matrix needs a matrix of the current particles, and FP32 values only in
    mixed scenes."""
    f, _, _ = _scene("tiny3d", abi.OPERATOR_FACTORED)
    x = np.zeros(3 * int(f.scalar("FluidCount")))
    with pytest.raises(MphError) as e:
        f.apply_matrix(x)  # nothing assembled yet
    assert e.value.code == -5
    f.step(1)
    assert np.all(f.apply_matrix(x) == 0.0)
    with pytest.raises(MphError) as e:
        f.apply_matrix(x, abi.PRECISION_MIXED)
    assert e.value.code == -5
    f.run_phase(abi.PHASE_KINEMATICS)  # particles moved: the matrix is stale
    with pytest.raises(MphError) as e:
        f.apply_matrix(x)
    assert e.value.code == -5
