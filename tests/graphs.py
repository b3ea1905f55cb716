"""Seeded test graphs shared by the CPU and GPU tests (inputs only)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def karate():
    d = json.load(open(os.path.join(GOLDEN, "karate.json")))
    return np.array(d["src"], np.int32), np.array(d["dst"], np.int32), d["n"], d["triangles"]


def complete(n):
    i, j = np.tril_indices(n, -1)
    return i.astype(np.int32), j.astype(np.int32)


def star(n):
    return np.zeros(n - 1, np.int32), np.arange(1, n, dtype=np.int32)


def bipartite(a, b, seed=0):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, a, 4 * (a + b))
    v = rng.integers(a, a + b, 4 * (a + b))
    return u.astype(np.int32), v.astype(np.int32)


def random_tree(n, seed=0):
    rng = np.random.default_rng(seed)
    parent = np.array([rng.integers(0, i) for i in range(1, n)], np.int32)
    return np.arange(1, n, dtype=np.int32), parent


def messy(n, m, seed=0):
    """Random pairs with duplicates, both orientations and self-loops."""
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m).astype(np.int32)
    v = rng.integers(0, n, m).astype(np.int32)
    k = m // 10
    u = np.concatenate([u, v[:k], u[:k], u[:k]])
    v = np.concatenate([v, u[:k], v[:k], u[:k]])  # reversed, duplicated and self-loop copies
    return u, v


def stencil27_triangles(a, b, c):
    """Triangles of the 27-point stencil on an a x b x c grid, counted by the
    bounding box of each triangle: three mutually adjacent points span 1 or 2
    values per axis, a box shape (x, y, z) fits (a-x+1)(b-y+1)(c-z+1) ways,
    and f(x, y, z) = triples inside an x*y*z box that span all of it."""
    from itertools import combinations, product

    def f(x, y, z):
        pts = list(product(range(x), range(y), range(z)))
        cnt = 0
        for tri in combinations(pts, 3):
            if all(len({p[d] for p in tri}) == (x, y, z)[d] for d in range(3)):
                cnt += 1
        return cnt

    total = 0
    for x, y, z in product((1, 2), repeat=3):
        total += f(x, y, z) * max(a - x + 1, 0) * max(b - y + 1, 0) * max(c - z + 1, 0)
    return total


def stencil9_triangles(a, b):
    """9-point stencil: each unit square is a K4 (4 triangles), nothing else."""
    return 4 * (a - 1) * (b - 1)
