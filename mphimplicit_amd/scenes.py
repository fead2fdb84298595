"""Seeded synthetic scenes: the benchmark configurations and small test scenes.

Lattice particles sit at ``lower + (i + 0.5) * l0`` like the reference's
pre-processor (generator/generator.cpp:170-193).  Fluid particles get i.i.d. jitter
U(-0.01 l0, +0.01 l0) per coordinate (x and y only in 2D) from SplitMix64 seeded
by the scene seed; wall particles stay on the lattice.  Parameters the
configurations leave open come from the reference's only 3D sample
(results3d/Cube*/box.data): density 1000, bulk modulus 1e6, bulk viscosity 1e4,
no surface tension, interaction ratios 1, static walls; radius ratios 3.5 / 1.7 /
3.5.
"""
from __future__ import annotations

import math

import numpy as np

from .abi import Params, default_params
from .files import GridData

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """n uniforms in [0,1) from the SplitMix64 stream of `seed` (53-bit mantissa)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def lattice(lo, hi, l0: float, dim: int) -> np.ndarray:
    """Lattice points of the box [lo, hi) (in units of l0) at (i+0.5)*l0."""
    axes = []
    for d in range(3 if dim == 3 else 2):
        n = int(round(hi[d] - lo[d]))
        axes.append(lo[d] * l0 + (np.arange(n) + 0.5) * l0)
    if dim == 2:
        gx, gy = np.meshgrid(axes[0], axes[1], indexing="ij")
        return np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], axis=1)
    gx, gy, gz = np.meshgrid(axes[0], axes[1], axes[2], indexing="ij")
    return np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)


def box_walls(inner_lo, inner_hi, layers: int, l0: float, dim: int, open_top: bool) -> np.ndarray:
    """Wall lattice of `layers` cells around [inner_lo, inner_hi) (lattice units);
    the last axis is vertical; an open top has no lid."""
    lo = [a - layers for a in inner_lo]
    hi = [b + layers for b in inner_hi]
    if open_top:
        hi[dim - 1] = inner_hi[dim - 1]
    pts = lattice(lo, hi, l0, dim)
    u = pts[:, :dim] / l0
    inside = np.ones(len(pts), dtype=bool)
    for d in range(dim):
        inside &= (u[:, d] > inner_lo[d]) & (u[:, d] < inner_hi[d])
    return pts[~inside]


def base_params(dim: int, mu: float, gravity, dt: float, steps: int) -> Params:
    p = default_params(dim)
    p.dt = dt
    p.end_time = (steps - 1) * dt  # the loop runs floor(EndTime/Dt)+1 steps
    p.output_interval = 1.0e9
    p.vtk_output_interval = 1.0e9
    p.radius_ratio_a = 3.5
    p.radius_ratio_p = 1.7
    p.radius_ratio_v = 3.5
    for t in range(4):
        p.density[t] = 1000.0
        p.bulk_modulus[t] = 1.0e6
        p.bulk_viscosity[t] = 1.0e4
        p.shear_viscosity[t] = mu
        p.surface_tension[t] = 0.0
        for u in range(4):
            p.interaction_ratio[t][u] = 1.0
    for d in range(3):
        p.gravity[d] = float(gravity[d])
    return p


def assemble(fluid, walls, vel_fluid, l0, dmin, dmax, seed: int, dim: int, jitter=0.01) -> GridData:
    nf = len(fluid)
    fluid = fluid.copy()
    if jitter:
        ncomp = 2 if dim == 2 else 3
        u = splitmix64_uniform(seed, nf * ncomp).reshape(nf, ncomp)
        fluid[:, :ncomp] += (2.0 * u - 1.0) * (jitter * l0)
    pos = np.concatenate([fluid, walls])
    vel = np.zeros_like(pos)
    vel[:nf] = vel_fluid
    typ = np.concatenate([np.zeros(nf, dtype=np.int32), np.full(len(walls), 2, dtype=np.int32)])
    return GridData(0.0, l0, np.array(dmin, float), np.array(dmax, float), typ, pos, vel)


def config_2d(nx=40, ny=40, box=(100, 60), seed=1, steps=200, mu=0.1):
    """Config 1: 2D fluid rectangle in the left corner of a 4-layer U-box."""
    l0 = 0.01
    fluid = lattice([0, 0], [nx, ny], l0, 2)
    walls = box_walls([0, 0], [box[0], box[1]], 4, l0, 2, open_top=True)
    pad = 5
    dmin = [-(4 + pad) * l0, -(4 + pad) * l0, -0.05]
    dmax = [(box[0] + 4 + pad) * l0, (box[1] + 40) * l0, 0.05]
    g = assemble(fluid, walls, 0.0, l0, dmin, dmax, seed, 2)
    p = base_params(2, mu, (0.0, -9.8, 0.0), 1.0e-3, steps)
    return p, g


def config_3d_drop(n=40, box=80, lift=16, seed=1, steps=100, mu=0.1):
    """Config 2: fluid cube of n^3 dropped in an open-top box of box^3 cells
    including 4-layer walls (118,016 walls at box=80)."""
    l0 = 0.01
    inner = box - 8
    lo_f = [4 + (inner - n) / 2.0, 4 + (inner - n) / 2.0, 4 + lift]
    fluid = lattice(lo_f, [lo_f[0] + n, lo_f[1] + n, lo_f[2] + n], l0, 3)
    walls = box_walls([4, 4, 4], [box - 4, box - 4, box], 4, l0, 3, open_top=True)
    dmin = [-0.05] * 3
    dmax = [box * l0 + 0.05] * 3
    g = assemble(fluid, walls, 0.0, l0, dmin, dmax, seed, 3)
    p = base_params(3, mu, (0.0, 0.0, -9.8), 5.0e-4, steps)
    return p, g


def periodic_3d(n=12, depth=6, seed=9, steps=5, mu=0.1):
    """A fluid layer of n x n x depth particles on a 4-layer floor, filling the
    domain in x and y: the domain is periodic there, so particles at opposite
    sides are neighbours through the boundary (test scene)."""
    l0 = 0.01
    fluid = lattice([0, 0, 4], [n, n, 4 + depth], l0, 3)
    walls = lattice([0, 0, 0], [n, n, 4], l0, 3)
    dmin = [0.0, 0.0, -0.05]
    dmax = [n * l0, n * l0, (4 + depth + 10) * l0]
    g = assemble(fluid, walls, 0.0, l0, dmin, dmax, seed, 3)
    p = base_params(3, mu, (0.0, 0.0, -9.8), 5.0e-4, steps)
    return p, g


def config_3d_slab(n=100, inner=140, seed=1, steps=50, mu=1.0, amplitude=0.1):
    """Config 3: n^3 fluid resting on the floor of a closed 4-layer box with an
    inner^3 interior, seeded Taylor-Green initial velocity of the given amplitude."""
    l0 = 0.01
    off = 4 + (inner - n) / 2.0
    lo_f = [off, off, 4]
    fluid = lattice(lo_f, [off + n, off + n, 4 + n], l0, 3)
    walls = box_walls([4, 4, 4], [4 + inner, 4 + inner, 4 + inner], 4, l0, 3, open_top=False)
    k = 2.0 * math.pi / (n * l0)
    ph = splitmix64_uniform(seed + 7919, 3) * 2.0 * math.pi
    x, y, z = fluid[:, 0], fluid[:, 1], fluid[:, 2]
    vel = np.zeros_like(fluid)
    vel[:, 0] = amplitude * np.sin(k * x + ph[0]) * np.cos(k * y + ph[1]) * np.cos(k * z + ph[2])
    vel[:, 1] = -amplitude * np.cos(k * x + ph[0]) * np.sin(k * y + ph[1]) * np.cos(k * z + ph[2])
    dmin = [-0.05] * 3
    dmax = [(inner + 8) * l0 + 0.05] * 3
    g = assemble(fluid, walls, vel, l0, dmin, dmax, seed, 3)
    p = base_params(3, mu, (0.0, 0.0, -9.8), 5.0e-4, steps)
    return p, g


def tiny_2d(seed=3, steps=5):
    """12x12 fluid in a small U-box (test scene)."""
    return config_2d(nx=12, ny=12, box=(20, 16), seed=seed, steps=steps)


def tiny_3d(seed=5, steps=5, lift=2):
    """6^3 fluid dropped in a 14^3 open box (test scene)."""
    return config_3d_drop(n=6, box=14, lift=lift, seed=seed, steps=steps)


def moving_walls(p: Params, g: GridData, dim: int) -> tuple[Params, GridData]:
    """Give the walls rigid motion (test scenes of the moving-wall branch): walls on
    the low-x side become type 3, the others stay type 2, and both types get a
    translation velocity and an angular velocity about their own centre
    (the Wall2/Wall3 lines of the reference's *.data file, main.cpp:554-555).
    In 2D the rotation axis is z, so walls stay in the plane."""
    typ = g.type.copy()
    pos = g.pos
    wall = typ >= 2
    xmin_fluid = pos[typ < 2, 0].min()
    side = wall & (pos[:, 0] < xmin_fluid - 2.0 * g.spacing)
    typ[side] = 3
    for w, sel in ((2, wall & ~side), (3, side)):
        c = pos[sel].mean(axis=0)
        for d in range(3):
            p.wall_center[w][d] = float(c[d])
    if dim == 2:
        vel = {2: (0.02, -0.01, 0.0), 3: (0.0, 0.03, 0.0)}
        omg = {2: (0.0, 0.0, 1.5), 3: (0.0, 0.0, -0.8)}
    else:
        vel = {2: (0.02, -0.01, 0.005), 3: (0.0, 0.03, -0.01)}
        omg = {2: (0.0, 0.0, 2.0), 3: (0.5, -1.0, 0.7)}
    for w in (2, 3):
        for d in range(3):
            p.wall_velocity[w][d] = vel[w][d]
            p.wall_omega[w][d] = omg[w][d]
    return p, GridData(g.time, g.spacing, g.domain_min, g.domain_max, typ, g.pos, g.vel)


def tiny_2d_moving(steps=3):
    """tiny_2d with translating and rotating walls of both wall types (test scene)."""
    p, g = tiny_2d(steps=steps)
    return moving_walls(p, g, 2)


def tiny_3d_moving(steps=3):
    """tiny_3d with translating and rotating walls of both wall types (test scene)."""
    p, g = tiny_3d(steps=steps)
    return moving_walls(p, g, 3)


def box_scene(seed: int, dim: int, nf_target: int, mu: float, gravity, dt=5.0e-4, steps=20):
    """A fluid square (2D) or cube (3D) of about nf_target particles resting on the
    floor, centred, in an open-top 4-layer box whose interior is 2.5x the fluid
    extent per axis (the batched and serving workloads)."""
    l0 = 0.01
    n = max(2, int(round(nf_target ** (1.0 / dim))))
    inner = int(math.ceil(2.5 * n))
    lo = [(inner - n) / 2.0] * dim
    lo[dim - 1] = 0.0
    hi = [a + n for a in lo]
    fluid = lattice(lo + [0.0] * (3 - dim), hi + [0.0] * (3 - dim), l0, dim)
    walls = box_walls([0] * dim, [inner] * dim, 4, l0, dim, open_top=True)
    pad = 5
    dmin = [-(4 + pad) * l0] * 3
    dmax = [(inner + 4 + pad) * l0] * 3
    dmax[dim - 1] = (inner + n + pad) * l0  # headroom above the open top
    if dim == 2:
        dmin[2], dmax[2] = -0.05, 0.05
    # A domain width that is a whole multiple of MaxRadius gives cells of MaxRadius,
    # where rounding can make ceil(MaxRadius / CellWidth) = 2 (initializeDomain,
    # main.cpp:1078-1096): the reference's neighbour search then refuses the domain
    # (main.cpp:1402).  Widen such a domain by half a spacing.
    max_r = max(3.5 * l0, 2.0 * (1.7 * l0), 3.5 * l0)
    for d in range(dim):
        w = dmax[d] - dmin[d]
        cells = math.floor(w / max_r)
        if math.ceil(max_r / (w / cells)) > 1:
            dmax[d] += 0.5 * l0
    g = assemble(fluid, walls, 0.0, l0, dmin, dmax, seed, dim)
    p = base_params(dim, mu, gravity, dt, steps)
    return p, g


def box_scene_counts(dim: int, nf_target: int) -> tuple[int, int]:
    """(fluid, total) particle counts of box_scene(dim, nf_target) without building it."""
    n = max(2, int(round(nf_target ** (1.0 / dim))))
    inner = int(math.ceil(2.5 * n))
    if dim == 2:
        walls = (inner + 8) * (inner + 4) - inner * inner
    else:
        walls = (inner + 8) * (inner + 8) * (inner + 4) - inner ** 3
    return n ** dim, n ** dim + walls


def batch_sweep_member(s: int, seed0=1000, steps=20, nf_range=(2000, 200000), mu_range=(0.01, 10.0)):
    """Scene s of the batched workload: seed seed0+s, 2D or 3D with probability 1/2,
    a log-uniform fluid count in nf_range and viscosity in mu_range, and gravity 9.8
    along a random unit vector (in-plane in 2D)."""
    u = splitmix64_uniform(seed0 + s + 0x5EED, 6)
    dim = 2 if u[0] < 0.5 else 3
    nf = math.exp(math.log(nf_range[0]) + u[1] * (math.log(nf_range[1]) - math.log(nf_range[0])))
    mu = math.exp(math.log(mu_range[0]) + u[2] * (math.log(mu_range[1]) - math.log(mu_range[0])))
    if dim == 2:
        a = 2.0 * math.pi * u[3]
        g = (9.8 * math.cos(a), 9.8 * math.sin(a), 0.0)
    else:
        z = 2.0 * u[3] - 1.0
        a = 2.0 * math.pi * u[4]
        r = math.sqrt(max(0.0, 1.0 - z * z))
        g = (9.8 * r * math.cos(a), 9.8 * r * math.sin(a), 9.8 * z)
    return box_scene(seed0 + s, dim, int(nf), mu, g, steps=steps)


def batch_sweep(count=64, seed0=1000, steps=20, nf_range=(2000, 200000), mu_range=(0.01, 10.0)):
    """The batched workload: scenes 0..count-1 of batch_sweep_member."""
    return [batch_sweep_member(s, seed0, steps, nf_range, mu_range) for s in range(count)]


def serve_stream(count: int, rate=200.0, seed=7, median=50000, sigma=1.0, clip=(5000, 500000),
                 mu_range=(0.01, 10.0), steps_range=(1, 20)):
    """Requests of the serving workload: Poisson arrivals at `rate` per second,
    1..20 steps each, lognormal fluid counts (median, sigma) clipped to `clip`,
    log-uniform viscosity, 2D or 3D with probability 1/2.  Returns a list of
    (arrival_s, steps, scene_seed, dim, nf, mu) tuples; box_scene builds each."""
    u = splitmix64_uniform(seed, 6 * count).reshape(count, 6)
    t = 0.0
    out = []
    for k in range(count):
        t += -math.log(1.0 - u[k, 0]) / rate
        steps = steps_range[0] + int(u[k, 1] * (steps_range[1] - steps_range[0] + 1))
        # Box-Muller normal from two uniforms
        z = math.sqrt(-2.0 * math.log(1.0 - u[k, 2])) * math.cos(2.0 * math.pi * u[k, 3])
        nf = int(min(clip[1], max(clip[0], median * math.exp(sigma * z))))
        mu = math.exp(math.log(mu_range[0]) + u[k, 4] * (math.log(mu_range[1]) - math.log(mu_range[0])))
        dim = 2 if u[k, 5] < 0.5 else 3
        out.append((t, steps, 20000 + k, dim, nf, mu))
    return out


CONFIGS = {
    "2d_40x40": config_2d,
    "3d_64k": config_3d_drop,
    "3d_1m": config_3d_slab,
    "tiny2d": tiny_2d,
    "tiny3d": tiny_3d,
}
