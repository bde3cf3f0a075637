"""Oracle pins for the binned-SAH builder (bvh/build.rs:45-480).

- find_best_plane: an independent numpy statement of the binned SAH (32 bins, cost
  area(L)*nL + area(R)*nR, first strict minimum over dimension-major plane order) must
  choose the same (dimension, split) as the oracle, on soups with many ties.
- partition (the `partition` 0.1.2 crate, build.rs:145): the swap closed form that the
  GPU uses equals a direct simulation of the crate's l/r scan loop.
- build invariants on random soups: 2N-1 nodes, one primitive per leaf, every id once,
  children of the k-th split are 2k+1 / 2k+2, parent links, leaf box = triangle box,
  internal box = union of the range padded by 1e-4, power = (0 + left) + right.
- hand-derived trees for 1, 2 and 3 triangles and a golden 4^3-grid node array.
"""
import os

import numpy as np
import pytest

from conftest import random_soup

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
F32 = np.float32


def np_best_plane(tri):
    """find_best_plane restated with numpy (float32 arithmetic, no FMA)."""
    v = tri.reshape(-1, 3, 3).astype(F32) + F32(0.0)
    cen = ((v[:, 0] + v[:, 1]) + v[:, 2]) / F32(3.0)
    bmin_box, bmax_box = v.min(1), v.max(1)
    best = (F32(3.4028234663852886e38), 0, F32(0.0))
    for d in range(3):
        lo, hi = cen[:, d].min(), cen[:, d].max()
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            scale = F32(32.0) / (hi - lo)
            f = (cen[:, d] - lo) * scale
        b = np.where(np.isnan(f), 0, np.minimum(31, np.floor(np.clip(f, 0, 1e9)))).astype(int)
        for pl in range(31):
            L, R = b <= pl, b > pl
            def area(m):
                if not m.any():
                    return F32(0.0)
                dd = bmax_box[m].max(0) - bmin_box[m].min(0)
                return F32(2.0) * ((dd[0] * dd[1] + dd[0] * dd[2]) + dd[1] * dd[2])
            cost = area(L) * F32(L.sum()) + area(R) * F32(R.sum())
            if cost < best[0]:
                best = (cost, d, lo + (F32(pl) + F32(1.0)) * ((hi - lo) / F32(32.0)))
    return best[1], F32(best[2])


@pytest.mark.parametrize("n,seed,grid", [(3, 0, False), (7, 1, True), (12, 2, True), (40, 3, False),
                                         (200, 4, True), (1000, 5, False)])
def test_best_plane_vs_numpy(n, seed, grid):
    import oracle
    tri = random_soup(n, seed, spread=10.0, grid=grid)
    d, s = oracle.find_best_plane(tri)
    nd, ns = np_best_plane(tri)
    assert (d, F32(s)) == (nd, ns)


def crate_partition(flags):
    """The l/r scan of the `partition` crate, simulated on a list of booleans."""
    d = list(range(len(flags)))
    p = lambda i: flags[i]
    n = len(d)
    if n == 0:
        return d, 0
    l, r = 0, n - 1
    while True:
        while l < n and p(d[l]):
            l += 1
        while r > 0 and not p(d[r]):
            r -= 1
        if l >= r:
            return d, l
        d[l], d[r] = d[r], d[l]


def closed_form_partition(flags):
    flags = np.asarray(flags, bool)
    T = int(flags.sum())
    d = np.arange(len(flags))
    lf = [i for i in range(T) if not flags[i]]
    rt = [i for i in range(len(flags) - 1, T - 1, -1) if flags[i]]
    for a, b in zip(lf, rt):
        d[a], d[b] = d[b], d[a]
    return list(d), T


def test_partition_closed_form():
    rng = np.random.default_rng(0)
    for n in list(range(0, 9)) + [31, 64, 257]:
        for _ in range(50):
            f = list(rng.random(n) < rng.random())
            assert crate_partition(f) == closed_form_partition(f)


def check_invariants(tri, lum, ids, nodes, order):
    n = len(tri)
    assert len(nodes) == 2 * n - 1
    leaf = nodes["left"] == 0xFFFFFFFF
    assert leaf.sum() == n
    assert sorted(nodes["right"][leaf]) == sorted(ids)
    assert sorted(order) == list(range(n))
    inner = np.nonzero(~leaf)[0]
    L, R = nodes["left"][inner], nodes["right"][inner]
    # the k-th internal node in DFS preorder of splits owns 2k+1, 2k+2
    assert np.array_equal(np.sort(L), 2 * np.arange(len(inner)) + 1)
    assert np.array_equal(R, L + 1)
    assert np.all(nodes["parent"][L] == inner) and np.all(nodes["parent"][R] == inner)
    assert nodes["parent"][0] == 0xFFFFFFFF
    v = tri.reshape(-1, 3, 3) + F32(0.0)
    id2prim = {int(i): p for p, i in enumerate(ids)}
    for k in np.nonzero(leaf)[0]:
        p = id2prim[int(nodes["right"][k])]
        assert nodes["min"][k].tobytes() == v[p].min(0).tobytes()
        assert nodes["max"][k].tobytes() == v[p].max(0).tobytes()
        assert nodes["power"][k] == lum[p]
    # internal: union of the subtree's leaves padded by 1e-4, power (0 + l) + r
    def walk(k):
        if nodes["left"][k] == 0xFFFFFFFF:
            p = id2prim[int(nodes["right"][k])]
            return v[p].min(0), v[p].max(0)
        a = walk(nodes["left"][k])
        b = walk(nodes["right"][k])
        mn, mx = np.minimum(a[0], b[0]), np.maximum(a[1], b[1])
        assert nodes["min"][k].tobytes() == (mn - F32(0.0001)).astype(F32).tobytes()
        assert nodes["max"][k].tobytes() == (mx + F32(0.0001)).astype(F32).tobytes()
        want = (F32(0.0) + nodes["power"][nodes["left"][k]]) + nodes["power"][nodes["right"][k]]
        assert nodes["power"][k] == want
        return mn, mx
    if n <= 2000:
        walk(0)


@pytest.mark.parametrize("n,seed,grid", [(1, 0, False), (2, 1, False), (3, 2, True), (17, 3, True),
                                         (300, 4, False), (1500, 5, True)])
def test_build_invariants(n, seed, grid):
    import oracle
    tri = random_soup(n, seed, grid=grid)
    rng = np.random.default_rng(seed)
    lum = rng.uniform(0, 5, n).astype(F32)
    ids = rng.permutation(n).astype(np.uint32)
    nodes, order, box, pw, depth = oracle.build_bl_bvh(tri, lum, ids)
    check_invariants(tri, lum, ids, nodes, order)


def test_hand_trees():
    import oracle
    # one triangle: the root is the leaf; returned box / power are the primitive's
    t1 = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], F32)
    nodes, order, box, pw, _ = oracle.build_bl_bvh(t1, [2.5], [9])
    assert len(nodes) == 1 and nodes[0]["left"] == 0xFFFFFFFF and nodes[0]["right"] == 9
    assert list(box) == [0, 0, 0, 1, 1, 0] and pw == F32(2.5)
    # two triangles: no SAH, left = first, right = second (build.rs:203-230)
    t2 = np.array([[5, 0, 0, 6, 0, 0, 5, 1, 0], [0, 0, 0, 1, 0, 0, 0, 1, 0]], F32)
    nodes, order, box, pw, _ = oracle.build_bl_bvh(t2, [1.0, 2.0], [10, 11])
    assert list(nodes["left"]) == [1, 0xFFFFFFFF, 0xFFFFFFFF]
    assert list(nodes["right"]) == [2, 10, 11]
    assert list(nodes["parent"]) == [0xFFFFFFFF, 0, 0]
    assert nodes[0]["power"] == F32(3.0)
    assert box[0] == F32(0.0) - F32(0.0001) and box[3] == F32(6.0) + F32(0.0001)
    # three unit triangles along x at 0, 10, 11: the SAH separates the far-left one
    t3 = np.array([[11, 0, 0, 12, 0, 0, 11, 1, 0], [0, 0, 0, 1, 0, 0, 0, 1, 0],
                   [10, 0, 0, 11, 0, 0, 10, 1, 0]], F32)
    nodes, order, *_ = oracle.build_bl_bvh(t3, [1, 1, 1], [0, 1, 2])
    # partition of [0,1,2] with c.x < split: only prim 1 is true -> swap(0, 1): [1, 0, 2]
    assert list(order) == [1, 0, 2]
    assert list(nodes["left"]) == [1, 0xFFFFFFFF, 3, 0xFFFFFFFF, 0xFFFFFFFF]
    assert list(nodes["right"]) == [2, 1, 4, 0, 2]


def test_golden_grid():
    """Node array of the 2-triangle faces of a 4^3 block of voxels (regression pin written
    by tests/golden/make_bvh_golden.py from the oracle)."""
    import oracle
    g = np.load(os.path.join(GOLDEN, "bvh_grid4.npz"))
    nodes, order, *_ = oracle.build_bl_bvh(g["tri"], g["lum"], g["ids"])
    assert nodes.tobytes() == g["nodes"].tobytes()
    assert np.array_equal(order, g["order"])
