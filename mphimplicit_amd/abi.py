"""ctypes mirror of include/mph_abi.h (structs and constants).

The structs are shared by the HIP library wrapper (``mphimplicit_amd.scene``) and by
the test harness of the CPU oracle, which takes the same parameter block.
"""
from __future__ import annotations

import ctypes as C

TYPE_COUNT = 4

OK = 0
E_ARG = -1
E_IO = -2
E_DEVICE = -3
E_NOMEM = -4
E_STATE = -5
E_FIELD = -6
E_NOTIMPL = -7

PRECISION_FP64 = 0
PRECISION_MIXED = 1
OPERATOR_ASSEMBLED = 0
OPERATOR_FACTORED = 1

PHASE_KINEMATICS = 1
PHASE_CELLS = 2
PHASE_NEIGHBOR = 3
PHASE_EXPLICIT = 4
PHASE_MATRIX_A = 5
PHASE_MATRIX_C = 6
PHASE_MULTIPLY_C = 7
PHASE_MULTIGRID = 8
PHASE_SOLVE = 9
PHASE_DIAGNOSTICS = 10
PHASE_ADVANCE = 11

PHASE_NAMES = {
    PHASE_KINEMATICS: "kinematics",
    PHASE_CELLS: "cells",
    PHASE_NEIGHBOR: "neighbor",
    PHASE_EXPLICIT: "explicit",
    PHASE_MATRIX_A: "matrix_a",
    PHASE_MATRIX_C: "matrix_c",
    PHASE_MULTIPLY_C: "multiply_c",
    PHASE_MULTIGRID: "multigrid",
    PHASE_SOLVE: "solve",
    PHASE_DIAGNOSTICS: "diagnostics",
    PHASE_ADVANCE: "advance",
}

_D4 = C.c_double * TYPE_COUNT
_D44 = (C.c_double * TYPE_COUNT) * TYPE_COUNT
_D3 = C.c_double * 3
_D43 = (C.c_double * 3) * TYPE_COUNT


class Params(C.Structure):
    """mph_params: the *.data keywords plus the solver switches."""

    _fields_ = [
        ("dim", C.c_int),
        ("dt", C.c_double),
        ("output_interval", C.c_double),
        ("vtk_output_interval", C.c_double),
        ("end_time", C.c_double),
        ("radius_ratio_a", C.c_double),
        ("radius_ratio_p", C.c_double),
        ("radius_ratio_v", C.c_double),
        ("density", _D4),
        ("bulk_modulus", _D4),
        ("bulk_viscosity", _D4),
        ("shear_viscosity", _D4),
        ("surface_tension", _D4),
        ("interaction_ratio", _D44),
        ("gravity", _D3),
        ("wall_center", _D43),
        ("wall_velocity", _D43),
        ("wall_omega", _D43),
        ("cr_tolerance", C.c_double),
        ("multigrid", C.c_int),
        ("precision", C.c_int),
        ("max_iterations", C.c_int),
        ("operator_form", C.c_int),
    ]

    def copy(self) -> "Params":
        out = Params()
        C.pointer(out)[0] = self
        return out


class Grid(C.Structure):
    """mph_grid: particles and domain (the content of a *.grid file)."""

    _fields_ = [
        ("time", C.c_double),
        ("np", C.c_int),
        ("spacing", C.c_double),
        ("domain_min", _D3),
        ("domain_max", _D3),
        ("type", C.POINTER(C.c_int)),
        ("pos", C.POINTER(C.c_double)),
        ("vel", C.POINTER(C.c_double)),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("steps_done", C.c_int),
        ("cr_iterations", C.c_int),
        ("cr_iterations_total", C.c_int),
        ("rr0", C.c_double),
        ("rr", C.c_double),
        ("rs0", C.c_double),
        ("rs", C.c_double),
        ("vtk_steps", C.c_int),
        ("ms_phase", C.c_double * 12),
        ("ms_total", C.c_double),
        ("status", C.c_int),
    ]


def default_params(dim: int = 3) -> Params:
    """Same values as mph_params_default(): the reference's static initialisers
    (source/main.cpp:65-75: everything 0 except Dt=1e100), multigrid on,
    CR tolerance rr/rr0 < 1e-12 (source/main.cpp:4346)."""
    p = Params()
    p.dim = dim
    p.dt = 1.0e100
    p.cr_tolerance = 1.0e-12
    p.multigrid = 1
    p.precision = PRECISION_FP64
    p.max_iterations = 0
    p.operator_form = OPERATOR_ASSEMBLED
    return p
