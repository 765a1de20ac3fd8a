"""Synthetic initial conditions for the dam-break workloads.

All coordinates are rounded to float32 and returned as float64 arrays, so
the FP32 device path and an FP64 checker start from identical states.
"""
import os

import numpy as np

from .params import BOUNDARY_AKINCI2012, BOUNDARY_BENDER2019, VISCOSITY_WEILER2018, default_params


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def lattice_block(dims, spacing, corner, jitter=0.0, seed=0):
    """dims[0] x dims[1] x dims[2] particles on a cubic lattice starting at
    `corner`, with uniform jitter in [-jitter, jitter]^3 from `seed`.
    Order: x slowest, z fastest (the loop order of
    SimulatorBase::createFluidBlocks, SimulatorBase.cpp:1474-1480)."""
    i, j, k = np.meshgrid(np.arange(dims[0]), np.arange(dims[1]), np.arange(dims[2]), indexing="ij")
    x = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1) * spacing + np.asarray(corner, np.float64)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        x = x + rng.uniform(-jitter, jitter, size=x.shape)
    return f32(x)


def fluid_block(box_min, box_max, particle_radius):
    """Lattice of SimulatorBase::createFluidBlocks, dense mode 0
    (SimulatorBase.cpp:1418-1526): spacing 2r, start box_min + 2r,
    round(extent / 2r) - 1 particles per axis."""
    box_min = np.asarray(box_min, np.float64)
    box_max = np.asarray(box_max, np.float64)
    diam = 2.0 * particle_radius
    steps = [int(round(d / diam)) - 1 for d in (box_max - box_min)]
    start = box_min + 2.0 * particle_radius
    return lattice_block(steps, diam, start)


def box_wall_particles(lo, hi, spacing):
    """One layer of Akinci2012 boundary particles on all six faces of the box
    [lo, hi], on a `spacing` grid (each surface grid point once)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    n = np.rint((hi - lo) / spacing).astype(int)
    g = np.meshgrid(*[np.arange(m + 1) for m in n], indexing="ij")
    idx = np.stack([a.ravel() for a in g], axis=1)
    on = np.zeros(len(idx), bool)
    for c in range(3):
        on |= (idx[:, c] == 0) | (idx[:, c] == n[c])
    return f32(lo + idx[on] * spacing)


def dam_break_small(per_side=30, seed=1, boundary="akinci", **kw):
    """27,000-particle dam break: 30^3 lattice at spacing 0.05 (r = 0.025,
    h = 0.1) from corner (0.1, 0.1, 0.1), jitter U[-0.001, 0.001]^3 from
    seed 1, zero velocity, in the box [0,4]x[0,3]x[0,2] (the stated
    [0,2]x[0,1.5]x[0,1] cannot hold a 1.5 m block) with one layer of Akinci
    particles on a 0.05 grid (or a volume-map wall). Cubic kernel, rho0 1000,
    CFL factor 1 (h_max 0.005), maxError 0.01 %, maxErrorV 0.1 %, XSPH 0.01."""
    dims = per_side if hasattr(per_side, "__len__") else (per_side,) * 3
    x = lattice_block(dims, 0.05, (0.1, 0.1, 0.1), jitter=0.001, seed=seed)
    lo, hi = (0.0, 0.0, 0.0), (4.0, 3.0, 2.0)
    params = dict(particle_radius=0.025, density0=1000.0, time_step_size=0.005, cfl_method=1,
                  cfl_factor=1.0, cfl_max_time_step_size=0.005, max_error=0.01, max_error_v=0.1,
                  xsph=0.01)
    params.update(kw)
    scene = dict(x=x, v=np.zeros_like(x), box_lo=lo, box_hi=hi)
    if boundary == "akinci":
        params["boundary_handling_method"] = BOUNDARY_AKINCI2012
        scene["boundary_x"] = box_wall_particles(lo, hi, 0.05)
    else:
        params["boundary_handling_method"] = BOUNDARY_BENDER2019
        scene["vm_res"] = volume_map_resolution(lo, hi, 0.1, 0.05)
    scene["params"] = default_params(**params)
    return scene


def volume_map_resolution(lo, hi, support_radius, cell, thickness=0.0):
    """Cells per axis so that the volume-map domain (box grown by 8h +
    thickness, SimulatorBase.cpp:2596-2600) is cut into cells of about
    `cell` metres."""
    ext = np.asarray(hi, np.float64) - np.asarray(lo, np.float64) + 2 * (8 * support_radius + thickness)
    return np.maximum(np.rint(ext / cell), 1).astype(np.uint32)


def dam_break_1m(per_side=100, seed=1, box_hi=(16.0, 6.0, 6.0), vm_cell=0.05, **kw):
    """1,000,000-particle dam break: 100^3 lattice at spacing 0.05 (r = 0.025)
    from corner (0.1, 0.1, 0.1) with jitter U[-0.001, 0.001]^3, in the box
    [0,16]x[0,6]x[0,6] (the stated [0,8]x[0,6]x[0,3] cannot hold the 5 m
    block), wall as a Bender2019 volume map with 0.05 m cells, Weiler2018
    viscosity nu 0.01 with CG tolerance 1e-4, XSPH 0.01, maxError 0.01 %."""
    dims = per_side if hasattr(per_side, "__len__") else (per_side,) * 3
    x = lattice_block(dims, 0.05, (0.1, 0.1, 0.1), jitter=0.001, seed=seed)
    lo, hi = (0.0, 0.0, 0.0), tuple(box_hi)
    params = dict(particle_radius=0.025, density0=1000.0, time_step_size=0.005, cfl_method=1,
                  cfl_factor=1.0, cfl_max_time_step_size=0.005, max_error=0.01, max_error_v=0.1,
                  xsph=0.01, viscosity_method=VISCOSITY_WEILER2018, viscosity=0.01,
                  viscosity_boundary=0.0, visco_max_iter=100, visco_max_error=1e-4,
                  boundary_handling_method=BOUNDARY_BENDER2019)
    params.update(kw)
    return dict(x=x, v=np.zeros_like(x), box_lo=lo, box_hi=hi,
                vm_res=volume_map_resolution(lo, hi, 0.1, vm_cell), params=default_params(**params))


def uniform_points(n, seed=3):
    rng = np.random.default_rng(seed)
    return f32(rng.uniform(0.0, 1.0, size=(n, 3)))


def clustered_points(n, seed=3, clusters=64, sigma=0.03, background=0.2):
    """Mixture of `clusters` Gaussian clusters (sigma, centres U[0.1,0.9]^3)
    holding (1 - background) of the points plus a uniform background, clipped
    to the unit cube."""
    rng = np.random.default_rng(seed)
    nb = int(round(n * background))
    nc = n - nb
    centres = rng.uniform(0.1, 0.9, size=(clusters, 3))
    lab = rng.integers(0, clusters, size=nc)
    pts = centres[lab] + rng.normal(0.0, sigma, size=(nc, 3))
    pts = np.concatenate([pts, rng.uniform(0.0, 1.0, size=(nb, 3))])
    return f32(np.clip(pts, 0.0, 1.0))


def dam_break_variant(seed, per_side=None, aspect=None, n_spheres=None, spacing=0.05, vm_cell=0.05, **kw):
    """Seeded dam-break variant for the learned warm start: a cubic fluid
    block of `per_side` (default U{60..120}) particles per axis at spacing
    0.05 (r = 0.025) from corner (0.1, 0.1, 0.1), jitter U[-0.001, 0.001]^3;
    box [0, L] x [0, b + 0.4] x [0, b + 0.2] with b the block edge and
    L = aspect * b (aspect U[1.5, 3]); `n_spheres` (default U{0..8}) static
    spheres of radius U[0.1, 0.4] placed in the box clear of the block, each a
    Bender2019 volume map of its analytic SDF. Boundary, solver and viscosity
    settings as in the 1M dam break."""
    rng = np.random.default_rng(7919 * seed + 17)
    n = int(per_side if per_side is not None else rng.integers(60, 121))
    a = float(aspect if aspect is not None else rng.uniform(1.5, 3.0))
    ns = int(n_spheres if n_spheres is not None else rng.integers(0, 9))
    b = n * spacing
    hi = (max(a * b, b + 0.4), b + 0.4, b + 0.2)
    spheres = place_spheres(rng, ns, hi, (0.05, 0.05, 0.05), (b + 0.15, b + 0.15, b + 0.15))
    sc = dam_break_1m(per_side=n, seed=seed, box_hi=hi, vm_cell=vm_cell, **kw)
    sc["spheres"] = spheres
    sc["sphere_res"] = vm_cell
    sc["variant"] = {"seed": seed, "per_side": n, "aspect": a, "spheres": spheres}
    return sc


def place_spheres(rng, count, box_hi, block_lo, block_hi, rmin=0.1, rmax=0.4, gap=0.05, tries=1000):
    """Up to `count` spheres of radius U[rmin, rmax] inside [0, box_hi], at
    least `gap` clear of the block [block_lo, block_hi] and of each other."""
    spheres = []
    block_lo, block_hi = np.asarray(block_lo, float), np.asarray(block_hi, float)
    for _ in range(tries):
        if len(spheres) >= count:
            break
        r = rng.uniform(rmin, rmax)
        if min(box_hi) < 2 * r:
            continue
        c = np.array([rng.uniform(r, box_hi[k] - r) for k in range(3)])
        q = np.clip(c, block_lo, block_hi)
        if np.linalg.norm(c - q) < r + gap:
            continue
        if any(np.linalg.norm(c - np.array(s[:3])) < r + s[3] + gap for s in spheres):
            continue
        spheres.append((float(c[0]), float(c[1]), float(c[2]), float(r)))
    return spheres


def ensemble_scene(seed, nmin=20_000, nmax=500_000, spacing=0.05, vm_cell=0.05, **kw):
    """One scene of the batched ensemble: a fluid block of about N particles,
    N log-uniform in [nmin, nmax], with seeded proportions (each axis x
    U[0.7, 1.4]), in a seeded box (length 1.5-3x the block's, height and
    depth 1.1-1.6x), placed at a seeded position on the floor, with 0-8
    static spheres of radius 0.1-0.4 clear of it; solver settings of the 1M
    dam break."""
    rng = np.random.default_rng(104729 * seed + 3)
    n = float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
    f = rng.uniform(0.7, 1.4, 3)
    f /= np.prod(f) ** (1.0 / 3.0)
    dims = tuple(max(4, int(round(n ** (1.0 / 3.0) * fk))) for fk in f)
    ext = np.array(dims) * spacing
    hi = (float(ext[0] * rng.uniform(1.5, 3.0) + 0.2), float(ext[1] * rng.uniform(1.1, 1.6) + 0.2),
          float(ext[2] * rng.uniform(1.1, 1.6) + 0.2))
    corner = (float(rng.uniform(0.1, hi[0] - ext[0] - 0.1)), 0.1, float(rng.uniform(0.1, hi[2] - ext[2] - 0.1)))
    ns = int(rng.integers(0, 9))
    spheres = place_spheres(rng, ns, hi, np.array(corner) - 0.05, np.array(corner) + ext + 0.05)
    x = lattice_block(dims, spacing, corner, jitter=0.001, seed=seed)
    sc = dam_break_1m(per_side=dims, seed=seed, box_hi=hi, vm_cell=vm_cell, **kw)
    sc["x"] = x
    sc["v"] = np.zeros_like(x)
    sc["spheres"] = spheres
    sc["sphere_res"] = vm_cell
    sc["variant"] = {"seed": seed, "dims": dims, "box": hi, "corner": corner, "spheres": spheres}
    return sc


DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def read_obj(path):
    """Vertices (n, 3) float64 and triangles (m, 3) uint32 (0-based) of an
    OBJ mesh; polygons are fanned into triangles."""
    v, f = [], []
    for line in open(path):
        p = line.split()
        if not p:
            continue
        if p[0] == "v":
            v.append([float(t) for t in p[1:4]])
        elif p[0] == "f":
            idx = [int(t.split("/")[0]) for t in p[1:]]
            idx = [i - 1 if i > 0 else len(v) + i for i in idx]
            for k in range(1, len(idx) - 1):
                f.append([idx[0], idx[k], idx[k + 1]])
    return np.array(v, np.float64), np.array(f, np.uint32)
