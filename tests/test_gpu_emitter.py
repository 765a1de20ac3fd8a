"""Particle states and emitters on the device against the FP64 oracle.

The oracle's emitter system is pinned by closed forms on the CPU; here the
device's emitter system (animation, layer placement, reuse, maxEmitter cap,
state release) and the state checks in its sweeps are compared with it:
stepwise from identical states (the oracle receives the device's fields and
particle states before each step; positions and velocities to 1e-4
normwise, particle counts, ids and states exactly), and over free runs of
120 steps (counts and states exactly, the emitter's animated particles to
1e-5 m, the splashing fluid to 5e-2 normwise: FP32 and FP64 free runs drift
apart once a solve's iteration count flips near its tolerance)."""
import math

import numpy as np
import pytest

from oracle import oracle
from splishsplash_amd import capi, scenes
from splishsplash_amd.params import BOUNDARY_AKINCI2012, BOUNDARY_BENDER2019, default_params

pytestmark = pytest.mark.gpu

R = 0.025
FIELDS = ["position", "velocity", "p / rho^2", "p_v / rho^2", "velocity difference"]


def rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


def rotation(axis, angle):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def pair(bmode=-1, vm=None, block=(6, 6, 6), corner=(-0.15, 0.0, -0.15), **kw):
    prm = default_params(particle_radius=R, boundary_handling_method=bmode, cfl_method=0, time_step_size=0.001, **kw)
    x = scenes.lattice_block(block, 2 * R, corner, jitter=0.002, seed=3)
    g = capi.TimeStepDFSPH(prm)
    o = oracle.OracleDFSPH(prm)
    g.set_fluid(x)
    o.set_fluid(x)
    if vm is not None:
        g.set_volume_map_box(*vm)
        o.set_volume_map_box(*vm)
    return g, o


def both(g, o, fn, *a, **kw):
    getattr(g, fn)(*a, **kw)
    getattr(o, fn)(*a, **kw)


def sync(g, o):
    for nm in FIELDS:
        o.set(nm, g.get(nm).astype(np.float64))
    o.set("state", g.get("state").astype(np.float64))


def check_step(g, o, tag):
    assert g.n == o.n, tag
    np.testing.assert_array_equal(g.get("state"), o.get("state").astype(np.uint32), err_msg=str(tag))
    assert g.emitter_counts() == o.emitter_counts(), tag
    assert g.counters()["time"] == o.counters()["time"], tag
    for nm in ("position", "velocity"):
        assert rel(g.get(nm), o.get(nm)) <= 1e-4, (tag, nm, rel(g.get(nm), o.get(nm)))


def run_stepwise(g, o, steps):
    g.init()
    o.init()
    for k in range(steps):
        sync(g, o)
        g.step(1)
        o.step(1)
        check_step(g, o, k)


def test_box_emitter_stepwise():
    # rotated box emitter shooting into a free block under gravity: layers
    # every 2r / v = 0.025 s, animation, release and the pressure solve
    # over fresh particles
    g, o = pair()
    both(g, o, "add_emitter", 5, 4, [0.0, 0.45, 0.0], rotation([0.0, 0.0, 1.0], -1.2), 2.0, type=0)
    run_stepwise(g, o, 90)
    assert g.emitter_counts()[0] == 4 * 20


def test_circle_emitter_in_a_container_stepwise():
    # circle emitter inside a volume-map box: animated particles see no
    # boundary; layers fall on the block, start and end times apply
    vm = ([-0.4, -0.05, -0.4], [0.4, 0.9, 0.4], [16, 19, 16])
    g, o = pair(bmode=BOUNDARY_BENDER2019, vm=vm)
    both(g, o, "add_emitter", 5, 0, [0.0, 0.6, 0.0], rotation([0.0, 0.0, 1.0], -math.pi / 2), 1.5, type=1,
         emit_start_time=0.01, emit_end_time=0.08)
    run_stepwise(g, o, 110)
    e = g.emitter_counts()[0]
    assert e > 0 and e % 21 == 0  # 21 disc points for width 5


def test_reuse_stepwise():
    # particles leaving the reuse box (falling below y = 0.25) are slowed and
    # placed again by the next layers, before any new particle is added
    g, o = pair(block=(4, 4, 4), corner=(-0.1, 0.3, -0.1))
    both(g, o, "enable_reuse_particles", [-1.0, 0.25, -1.0], [1.0, 2.0, 1.0])
    both(g, o, "add_emitter", 4, 4, [0.0, 0.5, 0.0], rotation([0.0, 0.0, 1.0], -math.pi / 2), 2.0)
    run_stepwise(g, o, 120)
    e, r = g.emitter_counts()
    assert r > 0 and e + r == 16 * 5


def test_cap_stepwise():
    # maxEmitterParticles = 50 with 16-particle layers: 16 + 16 + 16 + 2
    g, o = pair(block=(4, 4, 4), corner=(-0.1, 0.0, -0.1))
    both(g, o, "set_max_emitter_particles", 50)
    both(g, o, "add_emitter", 4, 4, [0.0, 0.5, 0.0], rotation([0.0, 0.0, 1.0], -math.pi / 2), 2.0)
    run_stepwise(g, o, 110)
    assert g.emitter_counts() == (50, 0) and g.n == 64 + 50


@pytest.mark.parametrize("bmode", [BOUNDARY_AKINCI2012, BOUNDARY_BENDER2019])
def test_emitter_free_run(bmode):
    # no synchronisation: the device's emission schedule, placement and
    # states stay identical to the oracle's over a free run
    if bmode == BOUNDARY_AKINCI2012:
        sc = scenes.dam_break_small(per_side=8)
        prm = sc["params"]
        prm.cfl_method = 0
        g = capi.TimeStepDFSPH(prm)
        o = oracle.OracleDFSPH(prm)
        for s in (g, o):
            s.set_fluid(sc["x"])
            s.add_boundary_particles(sc["boundary_x"])
        pos = np.float64(sc["x"]).mean(0) + [0.0, 0.3, 0.0]
    else:
        vm = ([-0.4, -0.05, -0.4], [0.4, 0.9, 0.4], [16, 19, 16])
        g, o = pair(bmode=bmode, vm=vm)
        pos = [0.0, 0.5, 0.0]
    both(g, o, "add_emitter", 4, 3, pos, rotation([1.0, 0.0, 1.0], 2.0), 1.0)
    for k in range(4):
        g.step(30)
        o.step(30)
        assert g.n == o.n, k
        st = g.get("state")
        np.testing.assert_array_equal(st, o.get("state").astype(np.uint32))
        assert g.emitter_counts() == o.emitter_counts()
        # the emitter moves its particles on a fixed path: tight; the fluid
        # itself drifts apart as FP32 and FP64 free runs of a splashing
        # fluid do (iteration counts flip near the tolerance)
        xg, xo = g.get("position"), o.get("position")
        anim = st == 1
        assert anim.any()
        assert np.abs(xg[anim] - xo[anim]).max() <= 1e-5, k
        assert rel(xg, xo) <= 5e-2, (k, rel(xg, xo))


def test_fixed_particles_stay_put():
    # state 2 (fixed): no velocity or position update, no pressure
    # acceleration, but they still push their neighbours
    sc = scenes.dam_break_small(per_side=10)
    g = capi.TimeStepDFSPH(sc["params"])
    o = oracle.OracleDFSPH(sc["params"])
    for s in (g, o):
        s.set_fluid(sc["x"])
        s.add_boundary_particles(sc["boundary_x"])
        s.init()
    st = np.zeros(g.n, np.uint32)
    st[::7] = 2
    g.set("state", st)
    o.set("state", st.astype(np.float64))
    x0 = g.get("position")
    for k in range(20):
        sync(g, o)
        g.step(1)
        o.step(1)
        assert rel(g.get("position"), o.get("position")) <= 1e-4, k
    x = g.get("position")
    np.testing.assert_array_equal(x[::7], x0[::7])
    assert not g.get("pressure acceleration")[::7].any()
    assert np.abs(x - x0).max() > 1e-3
    np.testing.assert_array_equal(g.get("state"), st)


def test_emitter_rejects_bad_input():
    g, _ = pair()
    with pytest.raises(capi.SPHError):
        g.add_emitter(4, 3, [0, 0, 0], np.eye(3), 0.0)
    with pytest.raises(capi.SPHError):
        g.add_emitter(4, 3, [0, 0, 0], np.eye(3), 1.0, type=2)
    with pytest.raises(capi.SPHError):
        g.add_emitter(0, 3, [0, 0, 0], np.eye(3), 1.0)
    g.init()
    with pytest.raises(capi.SPHError):
        g.set("state", np.full(g.n, 3, np.uint32))
