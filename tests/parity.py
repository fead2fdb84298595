"""Shared helpers of the parity tests: bitwise comparison of named fields between
the HIP scene and the CPU oracle, phase by phase."""
from __future__ import annotations

import numpy as np

from mphimplicit_amd import abi

PHASE_FIELDS = {
    abi.PHASE_KINEMATICS: ["Position", "Velocity", "Force"],
    abi.PHASE_CELLS: [
        "CellIndex", "CellParticle", "ParticleIndex", "Property", "Position", "Velocity",
        "CellFluidParticleBegin", "CellFluidParticleEnd", "CellWallParticleBegin", "CellWallParticleEnd",
    ],
    abi.PHASE_NEIGHBOR: [
        "NeighborCount", "NeighborFluidCount", "NeighborCountP", "NeighborPtr", "NeighborPtrP",
        "NeighborInd", "NeighborIndP",
    ],
    abi.PHASE_EXPLICIT: [
        "DensityA", "GravityCenter", "VolStrainP", "PressureA", "Mass", "Kappa", "Lambda", "Mu", "Force",
        "Velocity",
    ],
    abi.PHASE_MATRIX_A: ["CsrPtrA", "CsrIndA", "CsrCofA", "VectorB"],
    abi.PHASE_MATRIX_C: ["CsrPtrC", "CsrIndC", "CsrCofC", "VectorP"],
    abi.PHASE_MULTIPLY_C: ["CsrCofA", "VectorB"],
    abi.PHASE_MULTIGRID: ["InvDiagA"],
    abi.PHASE_SOLVE: ["Velocity"],
    abi.PHASE_DIAGNOSTICS: ["DivergenceP", "PressureP", "VirialPressureAtParticle", "VirialStressAtParticle", "Force"],
}
SOLVE_SCALARS = ["CRIterations", "rr0", "rr", "rs0", "rs"]


def bits(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    if a.dtype == np.float64:
        return a.view(np.int64)
    if a.dtype == np.float32:
        return a.view(np.int32)
    return a


def diff(name: str, a: np.ndarray, b: np.ndarray) -> str | None:
    """None if bit-identical, else a short description of the first difference."""
    if a.shape != b.shape:
        return "%s: shape %s vs %s" % (name, a.shape, b.shape)
    ba, bb = bits(a), bits(b)
    bad = np.nonzero(ba != bb)[0]
    if bad.size == 0:
        return None
    k = int(bad[0])
    msg = "%s: %d/%d differ, first at %d: %r vs %r" % (name, bad.size, a.size, k, a[k], b[k])
    if a.dtype.kind == "f":
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
        msg += " (max rel %.3e)" % np.nanmax(rel)
    return msg


def compare_phase(gpu, orc, phase: int, extra=()) -> list[str]:
    out = []
    names = list(PHASE_FIELDS.get(phase, [])) + list(extra)
    if phase == abi.PHASE_MULTIGRID:
        d_g, d_o = int(gpu.scalar("MultiGridDepth")), int(orc.scalar("MultiGridDepth"))
        if d_g != d_o:
            return ["MultiGridDepth %d vs %d" % (d_g, d_o)]
        for l in range(d_o):
            names += ["MGCofA_L%d" % l, "MGInvD_L%d" % l]
    for n in names:
        m = diff(n, gpu.read(n), orc.read(n))
        if m:
            out.append("phase %s %s" % (abi.PHASE_NAMES[phase], m))
    if phase == abi.PHASE_SOLVE:
        for n in SOLVE_SCALARS:
            a, b = gpu.scalar(n), orc.scalar(n)
            if np.float64(a).view(np.int64) != np.float64(b).view(np.int64):
                out.append("phase solve %s: %r vs %r" % (n, a, b))
    return out


def step_by_phase(gpu, orc, steps: int, diagnostics: bool = True) -> list[str]:
    """Run `steps` time steps phase by phase on both, comparing after each phase."""
    bad = []
    for s in range(steps):
        for ph in range(abi.PHASE_KINEMATICS, abi.PHASE_SOLVE + 1):
            gpu.run_phase(ph)
            orc.run_phase(ph)
            bad += ["step %d %s" % (s, m) for m in compare_phase(gpu, orc, ph)]
            if bad:
                return bad
        if diagnostics:
            gpu.run_phase(abi.PHASE_DIAGNOSTICS)
            orc.run_phase(abi.PHASE_DIAGNOSTICS)
            bad += ["step %d %s" % (s, m) for m in compare_phase(gpu, orc, abi.PHASE_DIAGNOSTICS)]
            if bad:
                return bad
        gpu.run_phase(abi.PHASE_ADVANCE)
        orc.run_phase(abi.PHASE_ADVANCE)
    return bad
