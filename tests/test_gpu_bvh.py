"""GPU builder parity: libvrgi's batched binned-SAH build vs the oracle's restatement of
build_bl_bvh / build_tl_bvh (bvh/build.rs:249-480). Bit-exact node arrays (40-byte
BvhNode records) and primitive order."""
import numpy as np
import pytest

from conftest import random_soup, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import vrgi
    mats = np.zeros(6, vrgi.MATERIAL_DTYPE)
    lum = np.zeros(6, np.float32)
    c = vrgi.Context(vrgi.config(16, 16), mats, lum)
    yield c
    c.close()


def _same_nodes(a, b):
    assert a.shape == b.shape
    assert a.tobytes() == b.tobytes(), _first_diff(a, b)


def _first_diff(a, b):
    i = np.nonzero(a.view(np.uint8).reshape(len(a), 40) != b.view(np.uint8).reshape(len(b), 40))[0]
    if len(i) == 0:
        return "equal"
    k = i[0]
    return f"first differing node {k} of {len(a)}: gpu {a[k]} oracle {b[k]}"


@pytest.mark.parametrize("n,seed,grid", [(1, 0, False), (2, 1, False), (3, 2, False), (5, 3, True),
                                         (64, 4, False), (257, 5, True), (1000, 6, False),
                                         (20000, 7, True), (150000, 8, False), (300000, 9, True)])
def test_bl_bvh_bit_exact(ctx, n, seed, grid):
    import oracle
    tri = random_soup(n, seed, spread=20.0 if grid else 50.0, grid=grid)
    rng = np.random.default_rng(seed)
    lum = rng.uniform(0.0, 10.0, n).astype(np.float32)
    ids = rng.permutation(n).astype(np.uint32)
    say(f"build_bl_bvh n={n} seed={seed} grid={grid}")
    g_nodes, g_order, g_box, g_pw = ctx.build_bl_bvh(tri, lum, ids)
    o_nodes, o_order, o_box, o_pw, _ = oracle.build_bl_bvh(tri, lum, ids)
    _same_nodes(g_nodes, o_nodes)
    assert np.array_equal(g_order, o_order)
    assert g_box.tobytes() == o_box.tobytes() and np.float32(g_pw) == np.float32(o_pw)


def test_bl_bvh_same_size_other_depth(ctx):
    """Builds of the same size whose trees differ in depth, back to back in one context:
    the builder queues as many top levels as the previous same-sized build had, so it
    sees a build with more levels (further level groups) and one with fewer (empty
    levels); every tree stays bit-exact."""
    import oracle
    n = 40000
    shallow = random_soup(n, 11, spread=50.0)
    rng = np.random.default_rng(12)
    c = np.geomspace(1.0, 1e4, n)[:, None] * np.array([1.0, 0.01, 0.02])[None, :]  # geometric spacing
    deep = (c[:, None, :] + rng.uniform(-0.5, 0.5, (n, 3, 3))).reshape(n, 9).astype(np.float32)
    lum = rng.uniform(0.0, 10.0, n).astype(np.float32)
    ids = np.arange(n, dtype=np.uint32)
    depths = []
    for k, tri in enumerate([shallow, deep, shallow, deep]):
        say(f"build_bl_bvh same size, build {k}")
        g_nodes, g_order, _, _ = ctx.build_bl_bvh(tri, lum, ids)
        o_nodes, o_order, _, _, _ = oracle.build_bl_bvh(tri, lum, ids)
        _same_nodes(g_nodes, o_nodes)
        assert np.array_equal(g_order, o_order)
        depths.append(_depth(g_nodes))
    assert depths[1] > depths[0] + 4, depths


def _depth(nodes):
    """Tree depth from the node records (left child, right child / primitive)."""
    raw = nodes.view(np.uint32).reshape(len(nodes), 10)
    left = raw[:, 0]
    depth = np.zeros(len(nodes), np.int64)
    for i in range(len(nodes)):  # parents precede children in index order
        if left[i] != 0xFFFFFFFF:
            depth[left[i]] = depth[i] + 1
            depth[raw[i, 1]] = depth[i] + 1
    return int(depth.max())


def test_bl_bvh_degenerate_centroids(ctx):
    """Every centroid equal: scale = 32/0 = inf, NaN bins -> 0, the partition is
    empty and the half split fallback (build.rs:152-157) runs at every level."""
    import oracle
    n = 777
    tri = np.tile(np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32), (n, 1))
    lum = np.ones(n, np.float32)
    ids = np.arange(n, dtype=np.uint32)
    say("build_bl_bvh degenerate centroids")
    g = ctx.build_bl_bvh(tri, lum, ids)
    o = oracle.build_bl_bvh(tri, lum, ids)
    _same_nodes(g[0], o[0])
    assert np.array_equal(g[1], o[1])


def test_tl_bvh_bit_exact(ctx):
    import oracle
    rng = np.random.default_rng(11)
    m = 700
    lo = rng.uniform(0, 200, (m, 3)).astype(np.float32)
    aabbs = np.concatenate([lo, lo + rng.uniform(0.5, 30, (m, 3)).astype(np.float32)], 1)
    xf = np.zeros((m, 12), np.float32)
    xf[:, [0, 4, 8]] = 1.0
    xf[:, 9:] = rng.integers(-64, 64, (m, 3)) * 32
    pw = rng.uniform(0, 100, m).astype(np.float32)
    ids = np.arange(m, dtype=np.uint32)
    say("build_tl_bvh m=700")
    _same_nodes(ctx.build_tl_bvh(xf, aabbs, pw, ids), oracle.build_tl_bvh(xf, aabbs, pw, ids))


def test_scene_bvhs_bit_exact(world0):
    """Geometry BVH over every exposed-face triangle, the light-triangle BVH, the light
    TL-BVH and every object's light BL-BVH of the 64^3 world equal the oracle's."""
    import oracle
    import vrgi
    say("scene BVHs 64^3 world")
    with vrgi.Context(vrgi.config(32, 32), world0["mats"], world0["lum"]) as c:
        objs = world0["objs"]
        handles = []
        for key, xf, v in objs:
            h = c.upload(v)
            handles.append((key, h, v))
            c.add(key, h, xf)
        c.trace_closest(np.zeros((1, 3), np.float32), np.ones((1, 3), np.float32))  # forces the build
        s = oracle.Scene(world0["mats"], world0["lum"])
        s.add_objects(objs)
        s.build()
        g = c.read(vrgi.BUF_GEOM_BVH, vrgi.NODE_DTYPE)
        _same_nodes(g, s.geom_bvh())
        assert np.array_equal(c.read(vrgi.BUF_GEOM_PRIM_ORDER, np.uint32)[: s.count(0)], s.geom_order())
        _same_nodes(c.read(vrgi.BUF_LIGHT_TL_BVH, vrgi.NODE_DTYPE), s.tl_bvh())
        _same_nodes(c.read(vrgi.BUF_LIGHT_GEOM_BVH, vrgi.NODE_DTYPE), s.light_geom_bvh())
        import ctypes
        for key, h, v in handles:
            if not h:
                continue
            n = ctypes.c_size_t()
            c.L.vrgi_object_light_bvh(h, None, 0, ctypes.byref(n))
            if n.value == 0:
                continue
            nodes = np.zeros(n.value, vrgi.NODE_DTYPE)
            c.L.vrgi_object_light_bvh(h, vrgi.ptr(nodes), n.value, ctypes.byref(n))
            _same_nodes(nodes, s.light_bvh(key))
        for key, h, v in handles:
            if h:
                c.L.vrgi_object_release(h)


def test_full_size_invariants():
    """Full-size workload (256x128x256 world, ~2.7M triangles): structural invariants of
    the GPU geometry BVH (2N-1 nodes, one primitive per leaf, every primitive once,
    parent / child links agree, the k-th split node owns children 2k+1 and 2k+2,
    boxes contain their children, leaf boxes equal the triangle boxes)."""
    import vrgi
    blocks, nt = vrgi.world(1, 0)
    mats, lum = vrgi.world_materials(1, 0)
    tr = vrgi.world_translucent(1, nt)
    objs = vrgi.chunk_objects(blocks, nt, tr)
    say("full-size geometry BVH invariants")
    with vrgi.Context(vrgi.config(32, 32), mats, lum) as c:
        c.add_objects(objs)
        c.trace_closest(np.zeros((1, 3), np.float32), np.ones((1, 3), np.float32))
        nodes = c.read(vrgi.BUF_GEOM_BVH, vrgi.NODE_DTYPE)
        depth = c.stats().bvh_depth
    # the depth the traversal stacks are sized by is the deepest leaf's (root 0)
    level, front = 0, np.array([0])
    while True:
        inner_f = front[nodes["left"][front] != 0xFFFFFFFF]
        if not len(inner_f):
            break
        front = np.concatenate([nodes["left"][inner_f], nodes["right"][inner_f]]).astype(np.int64)
        level += 1
    assert depth == level, (depth, level)
    n = (len(nodes) + 1) // 2
    tris = np.concatenate([v["position"].reshape(-1, 9) + np.tile(xf[9:], 3) for _, xf, v in objs if len(v)])
    assert len(tris) == n
    leaf = nodes["left"] == 0xFFFFFFFF
    assert leaf.sum() == n
    assert np.array_equal(np.sort(nodes["right"][leaf]), np.arange(n, dtype=np.uint32))
    inner = np.nonzero(~leaf)[0]
    L, R = nodes["left"][inner], nodes["right"][inner]
    assert np.array_equal(R, L + 1) and np.all(L % 2 == 1)
    assert np.all(nodes["parent"][L] == inner) and np.all(nodes["parent"][R] == inner)
    assert nodes["parent"][0] == 0xFFFFFFFF
    for side in (L, R):
        assert np.all(nodes["min"][inner] <= nodes["min"][side])
        assert np.all(nodes["max"][inner] >= nodes["max"][side])
    t = tris[nodes["right"][leaf]].reshape(-1, 3, 3)
    assert nodes["min"][leaf].tobytes() == t.min(1).astype(np.float32).tobytes()
    assert nodes["max"][leaf].tobytes() == t.max(1).astype(np.float32).tobytes()
