"""Readers and writers for the reference's text formats.

* ``*.data``: keyword file parsed by readDataFile (source/main.cpp:523-563).  Each
  line is tried against the keyword patterns in order; lines starting with '#' are
  comments; anything else is reported as an invalid line and ignored.
* ``*.grid`` / ``*.prof``: line 1 time, line 2 ``count spacing xmin xmax ymin ymax
  zmin zmax``, then one particle per line ``type x y z vx vy vz``
  (readGridFile 565-694, writeProfFile 696-720).

The C++ host library has its own readers (``mph_read_data_file`` /
``mph_read_grid_file``); these Python versions serve scripts and tests.  Python's
float() is correctly rounded, like glibc's strtod behind sscanf("%lf").
"""
from __future__ import annotations

import re
from dataclasses import dataclass

import numpy as np

from .abi import Params, default_params

_NUM = r"([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)"

# (keyword pattern, field, count) in the order of source/main.cpp:536-555
_DATA_KEYS = [
    ("Dt", "dt", 1),
    ("OutputInterval", "output_interval", 1),
    ("VtkOutputInterval", "vtk_output_interval", 1),
    ("EndTime", "end_time", 1),
    ("RadiusRatioA", "radius_ratio_a", 1),
    ("RadiusRatioP", "radius_ratio_p", 1),
    ("RadiusRatioV", "radius_ratio_v", 1),
    ("Density", "density", 4),
    ("BulkModulus", "bulk_modulus", 4),
    ("BulkViscosity", "bulk_viscosity", 4),
    ("ShearViscosity", "shear_viscosity", 4),
    ("SurfaceTension", "surface_tension", 4),
    ("InteractionRatio(Type0)", ("interaction_ratio", 0), 4),
    ("InteractionRatio(Type1)", ("interaction_ratio", 1), 4),
    ("InteractionRatio(Type2)", ("interaction_ratio", 2), 4),
    ("InteractionRatio(Type3)", ("interaction_ratio", 3), 4),
    ("Gravity", "gravity", 3),
]
_WALL_RE = {
    w: re.compile(
        r"^\s*Wall%d\s*Center\s*%s\s+%s\s+%s\s*Velocity\s*%s\s+%s\s+%s\s*Omega\s*%s\s+%s\s+%s"
        % ((w,) + (_NUM,) * 9)
    )
    for w in (2, 3)
}


def _match(keyword: str, count: int, line: str):
    m = re.match(r"^\s*" + re.escape(keyword) + r"\s*" + r"\s+".join([_NUM] * count), line)
    return [float(g) for g in m.groups()] if m else None


def parse_data_text(text: str, dim: int = 3, params: Params | None = None) -> Params:
    """Parse the content of a *.data file into an mph_params block."""
    p = params if params is not None else default_params(dim)
    for line in text.splitlines():
        if line.startswith("#"):
            continue
        done = False
        for key, field, count in _DATA_KEYS:
            vals = _match(key, count, line)
            if vals is None:
                continue
            if isinstance(field, tuple):
                row = getattr(p, field[0])[field[1]]
                for k, v in enumerate(vals):
                    row[k] = v
            elif count == 1:
                setattr(p, field, vals[0])
            else:
                arr = getattr(p, field)
                for k, v in enumerate(vals):
                    arr[k] = v
            done = True
            break
        if done:
            continue
        for w, rx in _WALL_RE.items():
            m = rx.match(line)
            if m:
                v = [float(g) for g in m.groups()]
                for d in range(3):
                    p.wall_center[w][d] = v[d]
                    p.wall_velocity[w][d] = v[3 + d]
                    p.wall_omega[w][d] = v[6 + d]
                done = True
                break
    return p


def read_data_file(path: str, dim: int = 3, params: Params | None = None) -> Params:
    with open(path) as f:
        return parse_data_text(f.read(), dim, params)


@dataclass
class GridData:
    time: float
    spacing: float
    domain_min: np.ndarray  # (3,)
    domain_max: np.ndarray  # (3,)
    type: np.ndarray  # (np,) int32
    pos: np.ndarray  # (np, 3) float64
    vel: np.ndarray  # (np, 3) float64

    @property
    def np(self) -> int:
        return int(self.type.shape[0])


def read_grid_file(path: str) -> GridData:
    with open(path) as f:
        time = float(f.readline().split()[0])
        h = f.readline().split()
        n = int(h[0])
        spacing = float(h[1])
        dmin = np.array([float(h[2]), float(h[4]), float(h[6])])
        dmax = np.array([float(h[3]), float(h[5]), float(h[7])])
        rows = np.loadtxt(f, max_rows=n, ndmin=2) if n > 0 else np.zeros((0, 7))
    return GridData(
        time,
        spacing,
        dmin,
        dmax,
        rows[:, 0].astype(np.int32),
        np.ascontiguousarray(rows[:, 1:4]),
        np.ascontiguousarray(rows[:, 4:7]),
    )


def write_grid_file(path: str, g: GridData, fmt: str = "%.17e") -> None:
    """Write a *.grid file.  The default %.17e round-trips every double exactly,
    so the reference binary and this library read identical inputs."""
    with open(path, "w") as f:
        f.write((fmt % g.time) + "\n")
        f.write(
            "%d %s  %s %s %s  %s %s %s\n"
            % (
                g.np,
                fmt % g.spacing,
                fmt % g.domain_min[0],
                fmt % g.domain_max[0],
                fmt % g.domain_min[1],
                fmt % g.domain_max[1],
                fmt % g.domain_min[2],
                fmt % g.domain_max[2],
            )
        )
        line = "%d  " + " ".join([fmt] * 3) + "  " + " ".join([fmt] * 3) + "\n"
        for i in range(g.np):
            f.write(line % ((int(g.type[i]),) + tuple(g.pos[i]) + tuple(g.vel[i])))


def write_data_file(path: str, p: Params) -> None:
    """Write a *.data file that readDataFile parses back into ``p``."""
    r = repr

    def four(a):
        return "  ".join(r(float(a[k])) for k in range(4))

    lines = [
        "Dt %s" % r(p.dt),
        "OutputInterval %s" % r(p.output_interval),
        "VtkOutputInterval %s" % r(p.vtk_output_interval),
        "EndTime %s" % r(p.end_time),
        "RadiusRatioA %s" % r(p.radius_ratio_a),
        "RadiusRatioP %s" % r(p.radius_ratio_p),
        "RadiusRatioV %s" % r(p.radius_ratio_v),
        "Density %s" % four(p.density),
        "BulkModulus %s" % four(p.bulk_modulus),
        "BulkViscosity %s" % four(p.bulk_viscosity),
        "ShearViscosity %s" % four(p.shear_viscosity),
        "SurfaceTension %s" % four(p.surface_tension),
    ]
    for t in range(4):
        lines.append("InteractionRatio(Type%d) %s" % (t, four(p.interaction_ratio[t])))
    lines.append("Gravity %s %s %s" % tuple(r(float(p.gravity[d])) for d in range(3)))
    for w in (2, 3):
        v = [p.wall_center[w][d] for d in range(3)] + [p.wall_velocity[w][d] for d in range(3)]
        v += [p.wall_omega[w][d] for d in range(3)]
        lines.append("Wall%d  Center %s %s %s  Velocity %s %s %s  Omega %s %s %s" % ((w,) + tuple(r(float(x)) for x in v)))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
