"""The drop-in executable end to end on the GPU.

mphimplicit_amd/mph_run takes the reference executable's arguments. It is given the
reference's Cube sample (box.data / box.grid, EndTime shortened to 13 steps) and runs
the whole time loop on the GPU, writing *.prof and *.vtk files and the log.

The CPU oracle, in the GPU's dot-product order and therefore bit-identical in every
field, is stepped with the same output schedule. The reference-format writers produce
the files it would write. Every file mph_run wrote must be byte-identical to the
corresponding oracle file, and every logged CR iteration count must equal the
oracle's.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle
from mphimplicit_amd import files
from test_abi import write_prof, write_vtk

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
RUN = os.path.join(os.path.dirname(HERE), "mphimplicit_amd", "mph_run")


def test_mph_run_files_match_oracle(tmp_path):
    fx = np.load(os.path.join(HERE, "golden", "cube3d.npz"), allow_pickle=False)
    text = re.sub(r"(?m)^EndTime\s+\S+", "EndTime\t\t\t\t0.0125", str(fx["data_text"]))
    (tmp_path / "box.data").write_text(text)
    h = fx["in_hdr"]
    g = files.GridData(float(h[0]), float(h[1]), h[2:5].copy(), h[5:8].copy(), fx["in_type"].copy(),
                       fx["in_pos"].copy(), fx["in_vel"].copy())
    files.write_grid_file(str(tmp_path / "box.grid"), g)
    print("[cli] mph_run on the cube3d sample, 13 steps", flush=True)
    r = subprocess.run([RUN, "box.data", "box.grid", "out%03d.prof", "out%03d.vtk", "box.log", "1", "--dim", "3"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    gpu_iters = [int(m) for m in re.findall(r"iter,=(\d+)", (tmp_path / "box.log").read_text())]

    p = files.parse_data_text(text, 3)
    o = pyoracle.OracleScene(p, g, pyoracle.DOT_TREE)
    dt = p.dt
    time = o.scalar("Time")
    output_next = vtk_next = time
    step, checked, iters = 0, 0, []
    while time < p.end_time + 1.0e-5 * dt:
        if time + 1.0e-5 * dt >= output_next:  # the prof file is written before the step
            ref = tmp_path / ("oracle%03d.prof" % step)
            write_prof(str(ref), o, g)
            assert (tmp_path / ("out%03d.prof" % step)).read_bytes() == ref.read_bytes(), step
            output_next += p.output_interval
            checked += 1
        vtk_due = time + 1.0e-5 * dt >= vtk_next
        st = o.step(1)
        iters.append(st.cr_iterations)
        assert bool(st.vtk_steps) == vtk_due
        if vtk_due:  # the vtk file is written after the step, with its diagnostics
            vtk_next += p.vtk_output_interval
            ref = tmp_path / ("oracle%03d.vtk" % step)
            write_vtk(str(ref), o)
            assert (tmp_path / ("out%03d.vtk" % step)).read_bytes() == ref.read_bytes(), step
            checked += 1
        time = o.scalar("Time")
        step += 1
    assert step == 13 and checked == 3
    assert sorted(x.name for x in tmp_path.glob("out[0-9]*")) == ["out000.prof", "out000.vtk", "out010.vtk"]
    assert (tmp_path / "output.vtk").exists()  # the initial state, as main.cpp:374
    assert gpu_iters == iters


def test_mph_run_convergence_check_lines(tmp_path):
    """--convergence-check logs the lines of the reference built with CONVERGENCE_CHECK
    (main.cpp:4350 per iteration, 4381, 4396).  On tiny3d the GPU's dot order prints
    the same six digits as the reference's serial order, so the log must hold exactly
    the reference's lines, kept in tests/golden/convergence.npz."""
    from test_oracle_convergence import load_case

    fx, p, g = load_case("tiny3d")
    (tmp_path / "box.data").write_text(str(fx["tiny3d/data_text"]))
    files.write_grid_file(str(tmp_path / "box.grid"), g)
    print("[cli] mph_run --convergence-check on tiny3d", flush=True)
    r = subprocess.run([RUN, "box.data", "box.grid", "out%03d.prof", "out%03d.vtk", "box.log", "1", "--dim", "3",
                        "--convergence-check"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    got = [ln for ln in (tmp_path / "box.log").read_text().splitlines() if re.match(r"line:\d+, iter,=", ln)]
    want = [str(s) for s in fx["tiny3d/lines"]]
    assert len(got) == len(want)
    for a, b in zip(got, want):
        assert a == b
