"""Batched scenes: many scenes of uneven size per launch (one multi-tree build, one launch
per stage, graph-replayed passes, tuned tile sizes).

Each pass of a batched frame is one frame of every scene, so a scene's buffers after the
batch equal, bit for bit, those of the same scene rendered alone (whose frames are the
oracle's: the frame parity tests). The 64^3 world is also checked against the oracle
directly inside a batch. The spp mean is (r_1 + r_2) * 0.5 in binary32."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu

SIZES = [(64, 48), (96, 64), (40, 40), (128, 96), (72, 56)]


def _scenes(world0, mode=0, flags=0, B=3):
    """The 64^3 world plus batch scenes 1..4, at small uneven extents."""
    import vrgi
    out = [(world0["objs"], world0["mats"], world0["lum"], SIZES[0],
            vrgi.camera_look([6.0, 62.0, 8.0], [40.0, 0.0, 44.0]))]
    for i in range(1, len(SIZES)):
        objs, mats, lum, _, cam = vrgi.batch_scene(i)
        out.append((objs, mats, lum, SIZES[i], cam))
    ctxs = []
    for objs, mats, lum, (W, H), cam in out:
        c = vrgi.Context(vrgi.config(W, H, B, flags=flags, mode=mode, radius=30.0 if mode else 15.0), mats, lum)
        c.add_objects(objs)
        ctxs.append(c)
    return ctxs, [o[4] for o in out]


def _close(ctxs):
    for c in ctxs:
        c.close()


def _state(c, bvh=False):
    import vrgi
    bufs = [("restir", vrgi.BUF_RESTIR), ("radiance", vrgi.BUF_RADIANCE), ("bgra", vrgi.BUF_BGRA8),
            ("temporal", vrgi.BUF_TEMPORAL_RESERVOIR), ("path_len", vrgi.BUF_PATH_LEN)]
    if bvh:
        bufs += [("bvh", vrgi.BUF_GEOM_BVH), ("light_bvh", vrgi.BUF_LIGHT_GEOM_BVH)]
    return {k: c.read(b, np.uint8).tobytes() for k, b in bufs}


def _alone(world0, mode, frames, flags=0, nee=1, spatial=5):
    """Each scene rendered alone: its state after `frames` frames (with its canonical
    BVHs) and its ReSTIR output of every frame."""
    import vrgi
    ctxs, cams = _scenes(world0, mode=mode, flags=flags)
    res = []
    try:
        for c, cam in zip(ctxs, cams):
            outs = []
            for f in range(frames):
                c.render(cam, nee_type=nee, spatial=spatial)
                c.sync()
                outs.append(c.read(vrgi.BUF_RESTIR).reshape(-1, 3).copy())
            st = _state(c)
            st["bvh"] = c.read(vrgi.BUF_GEOM_BVH, np.uint8).tobytes()
            st["light_bvh"] = c.read(vrgi.BUF_LIGHT_GEOM_BVH, np.uint8).tobytes()
            res.append((st, outs))
    finally:
        _close(ctxs)
    return res


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_matches_scenes_alone(world0, mode):
    """Two batched frames of 2 spp = 4 frames of every scene: every scene's buffers equal
    those of the scene rendered alone, and the mean is (r_3 + r_4) * 0.5."""
    import vrgi
    say(f"batch vs alone, restir mode {mode}")
    ref = _alone(world0, mode, 4)
    ctxs, cams = _scenes(world0, mode=mode)
    try:
        with vrgi.Batch(ctxs, spp=2, rays_per_block=128, pixels_per_block=128) as bt:
            for f in range(2):
                bt.render(cams, nee_type=1, spatial=5)
            bt.sync()
            for k, c in enumerate(ctxs):
                st, outs = ref[k]
                got = _state(c, bvh=True)
                for key in st:
                    assert got[key] == st[key], f"scene {k} {key}"
                want = (outs[2] + outs[3]) * np.float32(0.5)
                assert bt.mean(k).tobytes() == want.tobytes(), f"scene {k} mean"
            info = bt.info()
            assert info.n_scenes == len(ctxs) and info.graph == 2  # static scenes: each batched frame one graph of its passes
            assert info.rays > 0
    finally:
        _close(ctxs)


def test_batch_scene0_against_oracle(world0, scenes):
    """The 64^3 world inside a batch of five uneven scenes: its frames are the oracle's."""
    import oracle
    import vrgi
    s, _ = scenes
    W, H = SIZES[0]
    r = oracle.Renderer(W, H, 3)
    ctxs, cams = _scenes(world0)
    try:
        with vrgi.Batch(ctxs, spp=1, rays_per_block=64, pixels_per_block=256) as bt:
            for f in range(3):
                say(f"batched frame {f} against the oracle")
                bt.render(cams, nee_type=1, spatial=5)
                bt.sync()
                ref = r.render(s, cams[0], f, nee_type=1, spatial=5)
                c = ctxs[0]
                assert c.read(vrgi.BUF_RADIANCE).tobytes() == ref["l_out"][0].tobytes()
                assert c.read(vrgi.BUF_TEMPORAL_RESERVOIR).tobytes() == ref["temporal"].tobytes()
                assert c.read(vrgi.BUF_RESTIR).tobytes() == ref["restir"].tobytes()
                assert c.read(vrgi.BUF_BGRA8, np.uint8).tobytes() == ref["bgra8"].tobytes()
    finally:
        _close(ctxs)


def test_batch_tiles_graph_and_tuning_agree(world0):
    """Fixed tiles of 64 and 256, direct launches instead of the graph, and the autotuner's
    own choice all give the same bits."""
    import vrgi
    runs = []
    for kw in [dict(rays_per_block=64, pixels_per_block=64), dict(rays_per_block=256, pixels_per_block=256),
               dict(rays_per_block=128, pixels_per_block=64, flags=vrgi.BATCH_NO_GRAPH), dict()]:
        say(f"batch tiles {kw}")
        ctxs, cams = _scenes(world0)
        try:
            with vrgi.Batch(ctxs, spp=1, **kw) as bt:
                for f in range(7):  # the autotuner times six batched frames
                    bt.render(cams, nee_type=1, spatial=5)
                bt.sync()
                info = bt.info()
                if not kw:
                    assert info.tuned == 1 and info.rays_per_block in (64, 128, 256)
                    assert info.pixels_per_block in (64, 128, 256)
                runs.append([_state(c) for c in ctxs])
        finally:
            _close(ctxs)
    for r in runs[1:]:
        assert r == runs[0]


@pytest.mark.parametrize("spp,frames", [(1, 3), (2, 4)])
def test_batch_rebuild_every_frame(world0, spp, frames):
    """Scenes that rebuild every frame: the batched multi-tree build gives each scene the
    canonical BVHs its own build gives, and the frames still agree. The first batched frame
    builds with readbacks; the later ones replay that build inside one graph with every
    pass of the frame (info.graph 2)."""
    import vrgi
    flags = vrgi.FLAG_REBUILD_EVERY_FRAME
    say(f"batched rebuilds, spp {spp}, {frames} batched frames")
    ref = _alone(world0, 0, spp * frames, flags=flags)
    ctxs, cams = _scenes(world0, flags=flags)
    try:
        with vrgi.Batch(ctxs, spp=spp, rays_per_block=64, pixels_per_block=128) as bt:
            for f in range(frames):
                bt.render(cams, nee_type=1, spatial=5)
            bt.sync()
            info = bt.info()
            assert info.graph == 2 and info.build_ms < 0
            for k, c in enumerate(ctxs):
                got = _state(c, bvh=True)
                for key in got:
                    assert got[key] == ref[k][0][key], f"scene {k} {key}"
    finally:
        _close(ctxs)


def test_batch_frame_graph_after_edit(world0):
    """A scene's geometry changes between batched frames (an instance moves): that frame
    builds with readbacks again, the next ones replay the new build in a new frame graph,
    and every scene still equals the same scene rendered alone through the same edit."""
    import vrgi
    flags = vrgi.FLAG_REBUILD_EVERY_FRAME
    say("batched frame graph across a scene edit")
    iso = [1, 0, 0, 0, 1, 0, 0, 0, 1, 3.0, 0.0, -2.0]
    key = vrgi.batch_scene(1)[0][0][0]  # an object of scene 1
    states = []
    for batched in (False, True):
        ctxs, cams = _scenes(world0, flags=flags)
        try:
            if batched:
                with vrgi.Batch(ctxs, spp=1, rays_per_block=64, pixels_per_block=128) as bt:
                    for f in range(5):
                        if f == 2:
                            ctxs[1].update(key, iso)
                        bt.render(cams, nee_type=1, spatial=5)
                    bt.sync()
                    assert bt.info().graph == 2
                    states.append([_state(c, bvh=True) for c in ctxs])
            else:
                for f in range(5):
                    if f == 2:
                        ctxs[1].update(key, iso)
                    for c, cam in zip(ctxs, cams):
                        c.render(cam, nee_type=1, spatial=5)
                for c in ctxs:
                    c.sync()
                states.append([_state(c, bvh=True) for c in ctxs])
        finally:
            _close(ctxs)
    for k in range(len(states[0])):
        for key in states[0][k]:
            assert states[1][k][key] == states[0][k][key], f"scene {k} {key}"


def test_batch_rejects_unsupported(world0):
    """Scenes with the learned cache, ray sorting or per-bounce buffers cannot join a
    batch; a scene in a batch cannot be resized or destroyed."""
    import vrgi
    ctxs, cams = _scenes(world0)
    extra = vrgi.Context(vrgi.config(32, 32, 3, sort_bits=3), world0["mats"], world0["lum"])
    try:
        with pytest.raises(vrgi.VrgiError):
            vrgi.Batch(ctxs + [extra])
        with vrgi.Batch(ctxs) as bt:
            with pytest.raises(vrgi.VrgiError):
                check_resize = vrgi.lib().vrgi_resize(ctxs[0].h, 32, 32)
                vrgi.check(check_resize)
            with pytest.raises(vrgi.VrgiError):
                vrgi.Batch(ctxs[:2])  # already in a batch
            bt.render(cams, nee_type=1, spatial=5)
        # out of the batch: single renders work again (the scene rebuilds its own BVH)
        ctxs[1].render(cams[1], nee_type=1, spatial=5)
        ctxs[1].sync()
    finally:
        extra.close()
        _close(ctxs)
