"""The benchmark workloads at their real scene sizes against the oracle:

- the 1080p workload's world (256x128x256 terrain + 200 boxes, ~2.75M triangles): the
  geometry, light-geometry and light TL BVHs byte for byte against the oracle's build,
  and three frames of its configuration (4 bounces, Russian roulette from bounce 2, NEE
  on every bounce, ReSTIR GI with 5 neighbours, BVH rebuild every frame, the seeded
  camera spline) at 240x136, every per-bounce buffer bit-exact;
- the large dynamic workload's world (512x256x512, ~11M triangles at load): one frame of
  its 1% random-walk toggles through the device voxel world, whose BVHs equal those of
  the per-chunk CPU path byte for byte, with the geometry BVH's structural invariants
  (2N-1 nodes, one primitive per leaf, each primitive once, children 2k+1 / 2k+2).
"""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world1():
    import oracle
    import vrgi
    blocks, nt = vrgi.world(1, 0)
    mats, lum = vrgi.world_materials(1, 0)
    tr = vrgi.world_translucent(1, nt)
    objs = vrgi.chunk_objects(blocks, nt, tr)
    say("oracle build of the 256x128x256 world")
    s = oracle.Scene(mats, lum)
    s.add_objects(objs)
    s.build()
    return dict(objs=objs, mats=mats, lum=lum, scene=s)


def test_world1_bvhs_bit_exact(world1):
    import vrgi
    s = world1["scene"]
    say("full-size BVHs of the 256x128x256 world vs the oracle")
    with vrgi.Context(vrgi.config(32, 32), world1["mats"], world1["lum"]) as c:
        c.add_objects(world1["objs"])
        c.trace_closest(np.zeros((1, 3), np.float32), np.ones((1, 3), np.float32))
        g = c.read(vrgi.BUF_GEOM_BVH, np.uint8)
        lg = c.read(vrgi.BUF_LIGHT_GEOM_BVH, np.uint8)
        tl = c.read(vrgi.BUF_LIGHT_TL_BVH, np.uint8)
        order = c.read(vrgi.BUF_GEOM_PRIM_ORDER, np.uint32)
    ref = s.geom_bvh()
    assert len(ref) > 5_000_000
    assert g.tobytes() == ref.tobytes()
    assert order[: s.count(0)].tobytes() == s.geom_order().tobytes()
    assert lg.tobytes() == s.light_geom_bvh().tobytes()
    assert tl.tobytes() == s.tl_bvh().tobytes()


def test_world1_frames_bit_exact(world1):
    import oracle
    import vrgi
    from test_gpu_render import _check_frame
    s = world1["scene"]
    W, H, B, RR = 240, 136, 4, 2
    r = oracle.Renderer(W, H, B, rr_start=RR, rr_min=0.1)
    cfg = vrgi.config(W, H, B, rr_start=RR, rr_min=0.1,
                      flags=vrgi.FLAG_KEEP_BOUNCES | vrgi.FLAG_REBUILD_EVERY_FRAME)
    with vrgi.Context(cfg, world1["mats"], world1["lum"]) as c:
        c.add_objects(world1["objs"])
        for f in range(3):
            cam = vrgi.camera_spline(1, f, 16)
            say(f"1080p workload's world at {W}x{H}, spline frame {f}")
            c.render(cam, nee_type=1, spatial=5)
            c.sync()
            ref = r.render(s, cam, f, nee_type=1, spatial=5)
            plen = _check_frame(c, ref, W * H, B, 1)
            assert plen.max() == B and (plen < B).any()


def test_large_world_edits_bvhs():
    import vrgi
    blocks, nt = vrgi.world(4, 0)
    mats, lum = vrgi.world_materials(4, 0)
    tr = vrgi.world_translucent(4, nt)
    dx, dy, dz = blocks.shape
    rng = np.random.default_rng(4)
    walkers = np.stack([rng.integers(0, dx, 64), rng.integers(dy // 4, dy // 2, 64), rng.integers(0, dz, 64)],
                       1).astype(np.uint32)
    steps = blocks.size // 100
    cfg = vrgi.config(64, 48, 4, rr_start=2, mode=1, radius=30.0)
    cam = vrgi.camera_look([0.1 * dx, 0.9 * dy, 0.1 * dz], [0.6 * dx, 0.15 * dy, 0.6 * dz])
    with vrgi.Context(cfg, mats, lum) as g, vrgi.Context(cfg, mats, lum) as r:
        say("large dynamic world: voxel load")
        g.voxels_load(blocks, nt, tr)
        xyz, vals = vrgi.world_walk(blocks, nt, 9000, walkers, steps)
        say(f"large dynamic world: {len(vals)} voxel writes, re-mesh")
        g.voxels_set(xyz, vals)
        g.voxels_commit()
        say("large dynamic world: the CPU chunk path")
        r.add_objects(vrgi.chunk_objects(blocks, nt, tr))
        for c in (g, r):
            c.render(cam, nee_type=1, spatial=5)
            c.sync()
        for b in (vrgi.BUF_GEOM_BVH, vrgi.BUF_LIGHT_GEOM_BVH, vrgi.BUF_LIGHT_TL_BVH, vrgi.BUF_RADIANCE,
                  vrgi.BUF_RESTIR, vrgi.BUF_DI_LIGHT):
            assert g.read(b, np.uint8).tobytes() == r.read(b, np.uint8).tobytes(), b
        nodes = g.read(vrgi.BUF_GEOM_BVH, vrgi.NODE_DTYPE)
    n = (len(nodes) + 1) // 2
    assert n > 10_000_000
    leaf = nodes["left"] == 0xFFFFFFFF
    assert leaf.sum() == n
    assert np.array_equal(np.sort(nodes["right"][leaf]), np.arange(n, dtype=np.uint32))
    inner = np.nonzero(~leaf)[0]
    L, R = nodes["left"][inner], nodes["right"][inner]
    assert np.array_equal(R, L + 1) and np.array_equal(np.sort(L), 2 * np.arange(len(inner), dtype=np.uint32) + 1)
    assert np.all(nodes["parent"][L] == inner) and np.all(nodes["parent"][R] == inner)
