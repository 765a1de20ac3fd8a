"""The oracle's explicit (standard) viscosity, Viscosity_Standard.cpp:
240-400, pinned without retyping its formula.

Fluid part: on an interior lattice particle the operator
d nu sum_j m_j/rho_j (v_ij . x_ij) / (|x_ij|^2 + 0.01 h^2) gradW_ij with
d = 10 is a Riemann sum of an integral that has a closed form up to a 1-D
quadrature of the kernel: for v = (c y^2, 0, 0) the odd moments cancel and

    a_x = -d nu c (4 pi / 15) int_0^h W'(r) r^5 / (r^2 + 0.01 h^2) dr,

which for the eps -> 0 limit is nu * Laplacian(v)_x = 2 nu c. On a cubic
lattice of spacing s the computed densities give m_j / rho_j = s^3, so the
sum is a Riemann sum of that integral: at s = h/6 it matches the quadrature
to 0.1 % (at s = h/2, the particle spacing of a scene, the Riemann error is
28 %). A wrong factor, sign, transposed x_ij or a dropped 0.01 h^2 (3.7 %)
all break the 0.3 % bound.

Boundary part: invariants only (zero for a particle at rest, linear in its
velocity, scaled by viscosityBoundary, opposing motion along the wall)."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from oracle import oracle
from splishsplash_amd import scenes
from splishsplash_amd.params import BOUNDARY_AKINCI2012, BOUNDARY_BENDER2019, VISCOSITY_STANDARD, default_params


def cubic_dW(r, h):
    k = 8.0 / (math.pi * h ** 3)
    q = r / h
    if q <= 0.5:
        return k * (18.0 * q * q - 12.0 * q) / h
    if q <= 1.0:
        return -6.0 * k * (1.0 - q) ** 2 / h
    return 0.0


def lattice_sim(nu, c, per_h=6):
    r = 0.025
    n = 4 * per_h + 1  # neighbours of the centre particle see full supports
    x = scenes.lattice_block((n, n, n), 4 * r / per_h, (0.0, 0.0, 0.0), jitter=0.0, seed=0).astype(np.float64)
    v = np.zeros_like(x)
    v[:, 0] = c * x[:, 1] ** 2
    prm = default_params(particle_radius=r, boundary_handling_method=-1, viscosity_method=VISCOSITY_STANDARD,
                         viscosity=nu, gravity=(0.0, 0.0, 0.0), xsph=0.0)
    o = oracle.OracleDFSPH(prm)
    o.set_fluid(x, v)
    o.init()
    o.set_time(0.0, 0.004)
    for st in ("neighborhood_search", "densities", "clear_accel", "viscosity"):
        o.stage(st)
    return x, o.get("acceleration")


def test_standard_viscosity_matches_the_laplacian_quadrature():
    nu, c = 0.01, 3.0
    x, a = lattice_sim(nu, c)
    h = 0.1
    centre = x.mean(axis=0)
    i = int(np.argmin(np.linalg.norm(x - centre, axis=1)))
    I, _ = quad(lambda r: cubic_dW(r, h) * r ** 5 / (r * r + 0.01 * h * h), 0.0, h, points=[0.5 * h], limit=200)
    expect = -10.0 * nu * c * (4.0 * math.pi / 15.0) * I
    assert expect == pytest.approx(2.0 * nu * c, rel=0.05)  # the 0.01 h^2 term costs a few percent
    assert a[i, 0] == pytest.approx(expect, rel=0.003), (a[i], expect)
    assert abs(a[i, 1]) < 1e-6 * abs(expect) and abs(a[i, 2]) < 1e-6 * abs(expect)
    # linear in nu and c
    _, a2 = lattice_sim(2 * nu, -c)
    assert a2[i, 0] == pytest.approx(-2.0 * a[i, 0], rel=1e-12)


@pytest.mark.parametrize("bmode", ["akinci", "bender"])
def test_standard_viscosity_boundary_invariants(bmode):
    sc = scenes.dam_break_small(per_side=10, boundary=bmode)
    prm = sc["params"]
    prm.viscosity_method = VISCOSITY_STANDARD
    prm.viscosity = 0.0  # boundary part only
    prm.xsph = 0.0
    prm.gravity[:] = (0.0, 0.0, 0.0)

    x = sc["x"] - 0.06  # the block's first layers within h of the floor and two walls

    def accel(vb, vel):
        prm.viscosity_boundary = vb
        o = oracle.OracleDFSPH(prm)
        o.set_fluid(x, vel)
        if bmode == "akinci":
            o.add_boundary_particles(sc["boundary_x"])
        else:
            o.set_volume_map_box(sc["box_lo"], sc["box_hi"], sc["vm_res"])
        o.init()
        o.set_time(0.0, 0.004)
        for st in ("neighborhood_search", "volume_maps", "densities", "clear_accel", "viscosity"):
            o.stage(st)
        return o.get("acceleration")

    n = len(x)
    still = accel(0.5, np.zeros((n, 3)))
    assert not still.any()
    vel = np.zeros((n, 3))
    vel[:, 0] = 0.7  # along the floor y = 0 and the back wall z = 0
    a1 = accel(0.5, vel)
    a2 = accel(1.0, vel)
    a3 = accel(0.5, 2 * vel)
    np.testing.assert_allclose(a2, 2 * a1, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(a3, 2 * a1, rtol=1e-12, atol=1e-14)
    near = np.abs(a1[:, 0]) > 1e-9
    assert near.sum() > 50
    # drag: the wall force opposes the tangential motion
    assert (a1[near, 0] < 0).all()
