"""Device-resident voxel world: load, edits, re-mesh of the dirty chunks.

The reference path is the CPU one: chunk::gen_mesh per chunk (the CPU mesher) and
upload_object per chunk (chunk_objects + add_objects), whose frames are the oracle's
(the frame parity tests). A loaded or edited voxel world must give the same canonical
BVHs (geometry, light geometry, light TL) and the same frames, bit for bit."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _cam():
    import vrgi
    return vrgi.camera_look([6.0, 62.0, 8.0], [40.0, 0.0, 44.0])


def _bufs(c):
    import vrgi
    return {k: c.read(b, np.uint8).tobytes() for k, b in
            [("bvh", vrgi.BUF_GEOM_BVH), ("light_bvh", vrgi.BUF_LIGHT_GEOM_BVH), ("tl", vrgi.BUF_LIGHT_TL_BVH),
             ("radiance", vrgi.BUF_RADIANCE), ("restir", vrgi.BUF_RESTIR), ("temporal", vrgi.BUF_TEMPORAL_RESERVOIR),
             ("bgra", vrgi.BUF_BGRA8)]}


def _replace_all(c, blocks, world0):
    """The CPU chunk path: every chunk re-meshed and re-uploaded under its key."""
    import vrgi
    c.add_objects(vrgi.chunk_objects(blocks, world0["n_types"], world0["tr"]))


@pytest.mark.parametrize("mode", [0, 1])
def test_voxels_load_and_edits_match_cpu_chunks(world0, mode):
    import vrgi
    W, H, B = 96, 64, 3
    cfg = vrgi.config(W, H, B, mode=mode, radius=30.0 if mode else 15.0)
    blocks = world0["blocks"].copy()
    nt = world0["n_types"]
    walkers = np.array([[10, 40, 10], [31, 20, 33], [50, 30, 5], [63, 63, 63]], np.uint32)
    with vrgi.Context(cfg, world0["mats"], world0["lum"]) as g, \
            vrgi.Context(cfg, world0["mats"], world0["lum"]) as r:
        say(f"voxel world load, mode {mode}")
        g.voxels_load(blocks, nt, world0["tr"])
        _replace_all(r, blocks, world0)
        for f in range(4):
            if f > 0:  # random-walk toggles, applied to both
                xyz, vals = vrgi.world_walk(blocks, nt, 1000 + f, walkers, 3000)
                say(f"frame {f}: {len(vals)} voxel writes")
                g.voxels_set(xyz, vals)
                _replace_all(r, blocks, world0)
            g.render(_cam(), nee_type=1, spatial=5)
            r.render(_cam(), nee_type=1, spatial=5)
            g.sync()
            r.sync()
            a, b = _bufs(g), _bufs(r)
            for k in a:
                assert a[k] == b[k], f"frame {f} {k}"
        grid, commits, remeshed = g.voxels_read()
        assert np.array_equal(grid, blocks)
        assert commits == 4 and remeshed > 8


def test_voxels_dirty_chunks(world0):
    """A write inside a chunk re-meshes that chunk; one on a chunk corner also the three
    neighbours across its faces; a chunk emptied of blocks becomes an empty handle."""
    import vrgi
    blocks = world0["blocks"].copy()
    nt = world0["n_types"]
    with vrgi.Context(vrgi.config(32, 32, 3), world0["mats"], world0["lum"]) as g:
        g.voxels_load(blocks, nt, world0["tr"])
        _, c0, m0 = g.voxels_read()
        assert c0 == 1 and m0 == 8  # 64^3 = 2 x 2 x 2 chunks
        g.voxels_set([[10, 10, 10]], [nt])
        g.voxels_commit()
        _, c1, m1 = g.voxels_read()
        assert (c1, m1 - m0) == (2, 1)
        g.voxels_set([[31, 31, 31]], [2])
        g.voxels_commit()
        _, c2, m2 = g.voxels_read()
        assert m2 - m1 == 4
        # empty the chunk (1, 1, 1) entirely; frames still render and match the CPU path
        xs = np.stack(np.meshgrid(np.arange(32, 64), np.arange(32, 64), np.arange(32, 64), indexing="ij"), -1)
        xyz = xs.reshape(-1, 3).astype(np.uint32)
        g.voxels_set(xyz, np.full(len(xyz), nt, np.uint16))
        blocks[10, 10, 10] = nt
        blocks[31, 31, 31] = 2
        blocks[32:, 32:, 32:] = nt
        with vrgi.Context(vrgi.config(32, 32, 3), world0["mats"], world0["lum"]) as r:
            _replace_all(r, blocks, world0)
            g.render(_cam(), nee_type=1, spatial=5)
            r.render(_cam(), nee_type=1, spatial=5)
            g.sync()
            r.sync()
            a, b = _bufs(g), _bufs(r)
            for k in a:
                assert a[k] == b[k], k


def test_voxels_in_batch(world0):
    """Voxel edits of a scene in a batch are committed by the batched frame."""
    import vrgi
    blocks = world0["blocks"].copy()
    nt = world0["n_types"]
    walkers = np.array([[20, 30, 20]], np.uint32)
    cfg = vrgi.config(64, 48, 3)
    with vrgi.Context(cfg, world0["mats"], world0["lum"]) as g, \
            vrgi.Context(cfg, world0["mats"], world0["lum"]) as r:
        g.voxels_load(blocks, nt, world0["tr"])
        _replace_all(r, blocks, world0)
        with vrgi.Batch([g], spp=1, rays_per_block=64, pixels_per_block=64) as bt:
            for f in range(3):
                xyz, vals = vrgi.world_walk(blocks, nt, 7 + f, walkers, 500)
                g.voxels_set(xyz, vals)
                _replace_all(r, blocks, world0)
                bt.render([_cam()], nee_type=1, spatial=5)
                r.render(_cam(), nee_type=1, spatial=5)
            bt.sync()
            r.sync()
            a, b = _bufs(g), _bufs(r)
            for k in ("bvh", "light_bvh", "tl", "radiance", "restir", "temporal"):
                assert a[k] == b[k], k


def _oracle_chunks(blocks, nt, tr):
    """chunk_objects with the oracle's mesher: (key, transform, vertices) per 32^3 chunk."""
    import oracle
    nx, ny, nz = (s // 32 for s in blocks.shape)
    out = []
    for cx in range(nx):
        for cy in range(ny):
            for cz in range(nz):
                v = oracle.gen_mesh_chunk(blocks, nt, tr, cx, cy, cz)
                xf = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 32 * cx, 32 * cy, 32 * cz], np.float32)
                out.append(((cx * ny + cy) * nz + cz, xf, v))
    return out


@pytest.mark.parametrize("mode,cache", [(0, False), (1, False), (1, True)])
def test_voxel_edits_vs_oracle(world0, mode, cache):
    """The dynamic workload's path on the 64^3 world: random-walk toggles through the
    device voxel world (re-mesh of the dirty chunks, light gather, rebuild), 4 bounces with
    Russian roulette from bounce 2, NEE, ReSTIR GI or DI, with or without the learned
    cache; every frame equals the oracle's render of the edited world (its chunks meshed by
    the oracle's mesher), given the GPU's cache radiance in cache mode."""
    import oracle
    import vrgi
    W, H, B, RR = 96, 64, 4, 2
    frac, seed = 0.25, 3
    blocks = world0["blocks"].copy()
    nt = world0["n_types"]
    walkers = np.array([[10, 40, 10], [31, 20, 33], [50, 30, 5], [40, 50, 60]], np.uint32)
    cfg = vrgi.config(W, H, B, rr_start=RR, mode=mode, radius=30.0 if mode else 15.0)
    r = oracle.Renderer(W, H, B, rr_start=RR, mode=mode, radius=30.0 if mode else 15.0)
    cam = _cam()
    with vrgi.Context(cfg, world0["mats"], world0["lum"]) as g:
        g.voxels_load(blocks, nt, world0["tr"])
        if cache:
            g.cache_enable(box_lo=(0, 0, 0), box_hi=(0, 0, 0), train_fraction=frac, train_bounces=B, batch=2048,
                           steps_per_frame=2, seed=seed, query_vertex=0)
        for f in range(4):
            if f > 0:
                xyz, vals = vrgi.world_walk(blocks, nt, 2000 + f, walkers, 3000)
                g.voxels_set(xyz, vals)
            say(f"edited world vs oracle: mode {mode}, cache {cache}, frame {f}")
            g.render(cam, nee_type=1, spatial=5)
            g.sync()
            s = oracle.Scene(world0["mats"], world0["lum"])
            s.add_objects(_oracle_chunks(blocks, nt, world0["tr"]))
            s.build()
            if cache:
                crad = g.read(vrgi.BUF_CACHE_RADIANCE).reshape(W * H, 3)
                ref = r.render_cache(s, cam, f, crad, frac, oracle.frame_seed(seed, f, 5), nee_type=1, spatial=5,
                                     query_vertex=0)
            else:
                ref = r.render(s, cam, f, nee_type=1, spatial=5)
            assert g.read(vrgi.BUF_GEOM_BVH, np.uint8).tobytes() == s.geom_bvh().tobytes()
            assert g.read(vrgi.BUF_RADIANCE).tobytes() == ref["l_out"][0].tobytes()
            assert g.read(vrgi.BUF_RESTIR).tobytes() == ref["restir"].tobytes()
            assert g.read(vrgi.BUF_BGRA8, np.uint8).tobytes() == ref["bgra8"].tobytes()
            if mode == 1:
                assert g.read(vrgi.BUF_DI_LIGHT, np.uint32).tobytes() == ref["di_light"].tobytes()
                assert g.read(vrgi.BUF_DI_WEIGHTS).tobytes() == ref["di_weights"].tobytes()
            else:
                assert g.read(vrgi.BUF_TEMPORAL_RESERVOIR).tobytes() == ref["temporal"].tobytes()
