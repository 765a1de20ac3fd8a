"""The oracle's signed distance of a closed triangle mesh (the distance
field of a mesh volume map), pinned by an analytic box (a box mesh must give
the box's exact distance), by brute-force sampling of the triangles
(distance) and by a tessellated sphere (sign and value within the chord
error)."""
import os

import numpy as np

from oracle import oracle
from splishsplash_amd.scenes import DATA, read_obj

def box_sdf(x, lo, hi):
    q = np.maximum(np.maximum(lo - x, 0.0), x - hi)
    outside = np.linalg.norm(q, axis=1)
    inside = np.max(np.maximum(lo - x, x - hi), axis=1)  # negative inside
    return np.where(np.all((x >= lo) & (x <= hi), axis=1), inside, outside)


def test_box_mesh_gives_the_box_distance():
    v, f = read_obj(os.path.join(DATA, "unit_box.obj"))
    scale, shift = np.array([2.0, 1.0, 0.5]), np.array([0.3, -0.2, 0.1])
    v = v * scale + shift
    lo, hi = shift - scale / 2, shift + scale / 2
    x = np.random.default_rng(1).uniform(lo - 0.7, hi + 0.7, (4000, 3))
    np.testing.assert_allclose(oracle.mesh_sdf(v, f, x), box_sdf(x, lo, hi), rtol=0, atol=1e-12)


def test_distance_matches_dense_surface_sampling():
    v, f = read_obj(os.path.join(DATA, "unit_sphere.obj"))
    rng = np.random.default_rng(2)
    x = rng.uniform(-1.6, 1.6, (60, 3))
    # dense barycentric samples of every triangle
    u = np.linspace(0, 1, 41)
    a, b = np.meshgrid(u, u)
    keep = a + b <= 1
    a, b = a[keep], b[keep]
    tri = v[f]  # (m, 3, 3)
    pts = (tri[:, 0, None, :] * (1 - a - b)[None, :, None] + tri[:, 1, None, :] * a[None, :, None]
           + tri[:, 2, None, :] * b[None, :, None]).reshape(-1, 3)
    brute = np.array([np.sqrt(((pts - p) ** 2).sum(axis=1)).min() for p in x])
    d = np.abs(oracle.mesh_sdf(v, f, x))
    assert (d <= brute + 1e-12).all()  # the exact distance is below any sample's
    edge = np.linalg.norm(tri[:, 1] - tri[:, 0], axis=1).max()
    assert (brute - d <= edge / 40 + 1e-12).all()  # and within the sampling spacing


def test_tessellated_sphere_sign_and_value():
    v, f = read_obj(os.path.join(DATA, "unit_sphere.obj"))
    rng = np.random.default_rng(3)
    x = rng.normal(size=(3000, 3))
    r = rng.uniform(0.0, 2.0, 3000)
    x = x / np.linalg.norm(x, axis=1, keepdims=True) * r[:, None]
    s = oracle.mesh_sdf(v, f, x)
    far = np.abs(r - 1.0) > 0.05
    assert (np.sign(s[far]) == np.sign(r[far] - 1.0)).all()
    # the flat facets sit inside the sphere by at most the chord sagitta
    assert np.abs(s - (r - 1.0)).max() < 0.02
