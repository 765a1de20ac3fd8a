"""FP64 restatement of the reference's fluid-block sampler and initial
velocity, SimulatorBase::createFluidBlocks (Simulator/SimulatorBase.cpp:
1418-1516) and SimulatorBase::setInitialVelocity (:1079-1093), 3D case.

TEST INFRASTRUCTURE: the oracle is the checker for the library's scene
front-end. Only tests/ import this module.

Pinned in tests/test_scene.py by geometry that the lattice definitions fix
(nearest-neighbour distance 2r in every mode, twelve touching neighbours per
interior particle in the dense mode, the first particle 2r inside the
block's corner, all particles inside the block), not by restating the
formulas."""
import math

import numpy as np


def _round(x):
    """C's round(): halves away from zero."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def fluid_block(r, mode, start, end, translation=(0, 0, 0), scale=(1, 1, 1), v0=(0, 0, 0), w0=(0, 0, 0)):
    """Positions (FP64, sampling order) and velocities of one fluid block."""
    diam = 2.0 * r
    eps = 1.0e-9
    xshift, yshift = diam, diam
    if mode == 1:
        yshift = math.sqrt(3.0) * r + eps
    elif mode == 2:
        xshift = math.sqrt(6.0) * diam / 3.0 + eps
        yshift = math.sqrt(3.0) * r + eps
    mn = [scale[c] * start[c] + translation[c] for c in range(3)]
    mx = [scale[c] * end[c] + translation[c] for c in range(3)]
    diff = [mx[c] - mn[c] for c in range(3)]
    if mode == 1:
        diff[0] -= diam
        diff[2] -= diam
    elif mode == 2:
        diff[0] -= xshift
        diff[2] -= diam
    sx = int(_round(diff[0] / xshift)) - 1
    sy = int(_round(diff[1] / yshift)) - 1
    sz = int(_round(diff[2] / diam)) - 1
    if sx <= 1 or sy <= 1 or sz <= 1:
        return np.zeros((0, 3)), np.zeros((0, 3))
    st = [mn[c] + diam for c in range(3)]
    out = []
    for j in range(sx):
        for k in range(sy):
            for l in range(sz):
                p = [j * xshift + st[0], k * yshift + st[1], l * diam + st[2]]
                if mode == 1:
                    if k % 2 == 0:
                        p[2] += r
                    else:
                        p[0] += r
                elif mode == 2:
                    p[2] += r
                    s = [0.0, 0.0, 0.0]
                    if j % 2:
                        s[2] += diam / (2.0 * (-1 if k % 2 else 1))
                    if k % 2 == 0:
                        s[0] += xshift / 2.0
                    p = [p[c] + s[c] for c in range(3)]
                out.append(p)
    x = np.array(out, np.float64)
    com = x.sum(axis=0) / len(x)
    d = x - com
    v = np.asarray(v0, np.float64) + np.cross(np.asarray(w0, np.float64), d)
    return x, v
