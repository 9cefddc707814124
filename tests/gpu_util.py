"""Helpers shared by the GPU parity tests."""
import math

import numpy as np


def random_rays(m, n, seed, on_wall_frac=0.3):
    """Rays from random positions inside random cells of model m, with the
    start cell computed from the position (as testgrid would), plus rays that
    start exactly on a radial wall with on_wall = 1."""
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 3))
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    ijk = np.zeros((n, 3), np.int32)
    onw = np.zeros(n, np.uint8)
    for k in range(n):
        ir = int(rng.integers(1, m.nrg - 1))
        it = int(rng.integers(1, m.ntg))
        ip = int(rng.integers(1, m.npg))
        wall = rng.uniform() < on_wall_frac
        r = m.rarr[ir - 1] if wall else math.exp(rng.uniform(math.log(m.rarr[ir - 1]), math.log(m.rarr[ir])))
        th = rng.uniform(m.thetarr[it - 1], m.thetarr[it])
        ph = rng.uniform(m.phiarr[ip - 1], m.phiarr[ip])
        pos[k] = [r * math.sin(th) * math.cos(ph), r * math.sin(th) * math.sin(ph), r * math.cos(th)]
        if wall:
            # exact on the shell: r*r must reproduce r2arr exactly only approximately; flag on_wall
            onw[k] = 1
            # point inward/outward at random
        ijk[k] = [ir, it, ip]
    return pos, u, ijk, onw


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b)
    s = np.maximum(np.abs(a), np.abs(b))
    return np.where(s > 0, d / np.where(s > 0, s, 1.0), 0.0)
