"""Gauss-Legendre rule, serendipity cell and volume map of the oracle."""
import itertools

import numpy as np
import pytest

from oracle import oracle


def test_gauss_rule_matches_reference_table(golden):
    g = golden["gauss_p30"]
    x, w = oracle.gauss_legendre(g["n"])
    np.testing.assert_allclose(x, g["x"], rtol=0, atol=2e-16)
    # The reference table's weights differ from the exact Gauss-Legendre
    # weights by up to 1.5e-13 relative (its first weight ends ...758110,
    # the exact one ...754095); the oracle's rule is the exact one.
    np.testing.assert_allclose(w, g["w"], rtol=2e-13, atol=0)


def test_gauss_rule_exact_for_polynomials():
    x, w = oracle.gauss_legendre(16)
    for k in range(0, 32):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert np.dot(w, x ** k) == pytest.approx(exact, abs=2e-15)


def test_gauss_monomials_match_reference_integrate(golden):
    # GaussQuadrature::integrate (GaussQuadrature.cpp:5930-5962) restated in
    # numpy with the oracle's rule, against the compiled reference.
    x, w = oracle.gauss_legendre(16)
    for m in golden["gauss_monomials"]:
        lo, hi = np.array(m["lo"]), np.array(m["hi"])
        c0, c1 = 0.5 * (hi - lo), 0.5 * (hi + lo)
        pts = [c0[a] * x + c1[a] for a in range(3)]
        f = [pts[a] ** m["exp"][a] for a in range(3)]
        val = np.einsum("i,j,k,i,j,k->", w, w, w, f[0], f[1], f[2]) * c0.prod()
        assert val == pytest.approx(m["value"], rel=1e-13, abs=1e-300)
        # and the closed form of the monomial integral (degree <= 31 exact)
        exact = np.prod([(hi[a] ** (m["exp"][a] + 1) - lo[a] ** (m["exp"][a] + 1)) / (m["exp"][a] + 1)
                         for a in range(3)])
        assert val == pytest.approx(exact, rel=1e-12)


def test_serendipity_interpolates_nodes():
    nodes = oracle.serendipity_nodes()
    assert len({tuple(p) for p in nodes}) == 32
    for k in range(32):
        N, _ = oracle.serendipity(nodes[k])
        e = np.zeros(32)
        e[k] = 1
        np.testing.assert_allclose(N, e, atol=1e-14)


def test_serendipity_partition_and_cubic_reproduction():
    rng = np.random.default_rng(2)
    nodes = oracle.serendipity_nodes()
    mons = [m for m in itertools.product(range(4), repeat=3) if sum(m) <= 3]
    coef = rng.normal(size=len(mons))

    def f(p):
        return sum(c * p[0] ** a * p[1] ** b * p[2] ** d for c, (a, b, d) in zip(coef, mons))

    def df(p):
        g = np.zeros(3)
        for c, (a, b, d) in zip(coef, mons):
            if a:
                g[0] += c * a * p[0] ** (a - 1) * p[1] ** b * p[2] ** d
            if b:
                g[1] += c * b * p[0] ** a * p[1] ** (b - 1) * p[2] ** d
            if d:
                g[2] += c * d * p[0] ** a * p[1] ** b * p[2] ** (d - 1)
        return g

    fn = np.array([f(p) for p in nodes])
    for _ in range(20):
        xi = rng.uniform(-1, 1, 3)
        N, dN = oracle.serendipity(xi)
        assert N.sum() == pytest.approx(1.0, abs=1e-14)
        np.testing.assert_allclose(dN.sum(axis=0), 0, atol=1e-13)
        assert N @ fn == pytest.approx(f(xi), abs=1e-12)
        np.testing.assert_allclose(dN.T @ fn, df(xi), atol=1e-11)


def test_serendipity_gradient_finite_difference():
    rng = np.random.default_rng(3)
    for _ in range(5):
        xi = rng.uniform(-0.9, 0.9, 3)
        _, dN = oracle.serendipity(xi)
        for c in range(3):
            e = np.zeros(3)
            e[c] = 1e-6
            Np, _ = oracle.serendipity(xi + e)
            Nm, _ = oracle.serendipity(xi - e)
            np.testing.assert_allclose((Np - Nm) / 2e-6, dN[:, c], atol=1e-8)


def test_box_sdf_closed_form():
    lo, hi = (0, 0, 0), (4, 3, 2)
    pts = np.array([[1, 1, 1], [0.1, 1.5, 1.0], [2, 2.9, 1], [-0.3, 1, 1], [-0.3, -0.4, 1], [5, 4, 3]])
    d = oracle.box_sdf(lo, hi, pts)
    np.testing.assert_allclose(d, [1.0, 0.1, 0.1, -0.3, -0.5, -np.sqrt(3)], atol=1e-15)


def test_volume_node_matches_reference_halfspace(golden):
    # Node at distance d from the face y = 0 of a large box: the box distance
    # is d + xi_y over the whole support, i.e. the half-space integrand the
    # probe integrated with the reference's GaussQuadrature and CubicKernel.
    g = golden["halfspace_volume"]
    lo, hi = (0.0, 0.0, 0.0), (10.0, 10.0, 10.0)
    res = (12, 12, 12)
    pts = np.array([[5.0, d, 5.0] for d in g["d"]])
    val, _, ok = oracle.volume_map_eval(lo, hi, res, g["h"], pts, field=1, interpolate=False)
    assert ok.all()
    np.testing.assert_allclose(val, g["volume"], rtol=1e-12, atol=1e-15)
    # deep in the fluid the volume vanishes; deep in the wall it is 0.8 times
    # the quadrature of the ball (4/3 pi h^3 up to the error of a product
    # rule on a discontinuous integrand)
    assert val[-1] == 0.0
    assert val[0] == pytest.approx(0.8 * 4 / 3 * np.pi * g["h"] ** 3, rel=5e-2)


def test_volume_map_interpolation_exact_where_sdf_linear():
    lo, hi = (0.0, 0.0, 0.0), (2.0, 2.0, 2.0)
    res = (36, 36, 36)
    rng = np.random.default_rng(4)
    pts = np.stack([rng.uniform(0.5, 1.5, 50), rng.uniform(0.01, 0.3, 50), rng.uniform(0.5, 1.5, 50)], 1)
    val, grad, ok = oracle.volume_map_eval(lo, hi, res, 0.1, pts, field=0, interpolate=True)
    assert ok.all()
    np.testing.assert_allclose(val, pts[:, 1], atol=1e-12)
    np.testing.assert_allclose(grad, np.tile([0, 1, 0], (50, 1)), atol=1e-10)
    out = np.array([[-1.0, 1, 1], [3.5, 1, 1]])
    _, _, ok = oracle.volume_map_eval(lo, hi, res, 0.1, out, field=0, interpolate=True)
    assert not ok.any()


def test_sphere_sdf_closed_form():
    c, R = np.array([1.0, 0.5, 0.75]), 0.3
    pts = np.random.default_rng(5).uniform(0.0, 2.0, (40, 3))
    d = oracle.sphere_sdf(c, R, pts)
    np.testing.assert_allclose(d, np.linalg.norm(pts - c, axis=1) - R, atol=1e-15)
    # inverted and with thickness: sign * (d - thickness)
    d2 = oracle.sphere_sdf(c, R, pts, invert=True, thickness=0.05)
    np.testing.assert_allclose(d2, -(np.linalg.norm(pts - c, axis=1) - R - 0.05), atol=1e-15)


def test_sphere_volume_tends_to_halfspace():
    # Near the surface of a sphere of radius 1000 h the volume integrand is
    # the half-space one up to O(h / R): compare with the map of the face
    # y = 0 of a (non-inverted) box, which is exactly the half space there.
    h, Rb = 0.1, 100.0
    pts = np.array([[0.0, d, 0.0] for d in np.linspace(-0.05, 0.12, 8)])
    v, _, ok = oracle.sphere_volume_map_eval((0.0, -Rb, 0.0), Rb, (4, 4, 4), h, pts, field=1, interpolate=False)
    vb, _, _ = oracle.volume_map_eval((-5, -10, -5), (5, 0, 5), (4, 4, 4), h, pts, field=1, interpolate=False,
                                      invert=False)
    assert ok.all()
    np.testing.assert_allclose(v, vb, rtol=2e-3)
    # monotone in the distance; zero once the ball |xi| < h stays at sdf >= h
    assert np.all(np.diff(v) < 0)
    far, _, _ = oracle.sphere_volume_map_eval((0.0, -Rb, 0.0), Rb, (4, 4, 4), h, [[0.0, 0.2001, 0.0]], field=1,
                                              interpolate=False)
    assert far[0] == 0.0


def test_sphere_volume_map_interpolates_sdf():
    # The cubic element reproduces the sphere's distance to O(cell^4): at
    # 0.05 m cells and R = 0.3 m the error is far below the particle radius.
    c, R = np.array([1.0, 1.0, 1.0]), 0.3
    res = [int(round(2 * (R + 0.8) / 0.05))] * 3
    rng = np.random.default_rng(6)
    dirs = rng.normal(size=(50, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    pts = c + dirs * rng.uniform(R - 0.05, R + 0.1, (50, 1))
    val, grad, ok = oracle.sphere_volume_map_eval(c, R, res, 0.1, pts, field=0, interpolate=True)
    assert ok.all()
    np.testing.assert_allclose(val, np.linalg.norm(pts - c, axis=1) - R, atol=1e-4)
    np.testing.assert_allclose(grad, dirs, atol=5e-3)
