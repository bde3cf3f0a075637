"""Texture atlas (vrgi_set_textures): shading reads each slot's reflectivity,
emissivity and metallicity images at the hit's uv (raytrace.rs:437-481) through the
reference's default sampler. Frames of the 64^3 world with a random atlas are bit-exact
against the oracle, with and without per-frame rebuilds and the learned cache's albedo;
clearing the atlas restores the constant materials; the luminance helper equals
utils.rs:223-235 as the oracle states it; ReSTIR DI refuses an atlas, and a textured
scene inside a batch renders the oracle's frames."""
import numpy as np
import pytest

from conftest import say
from test_gpu_render import _camera0, _check_frame

pytestmark = pytest.mark.gpu


def _atlas(n_slots, seed, lum):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 256, (n_slots, 3, 16, 16, 4), dtype=np.uint8)
    a[:, 0, ..., 3] = np.where(rng.random((n_slots, 16, 16)) < 0.9, 255, rng.integers(0, 256, (n_slots, 16, 16)))
    a[:, 2, ..., 0] = np.where(rng.random((n_slots, 16, 16)) < 0.85, 0, a[:, 2, ..., 0])
    a[:, 1, ..., :3] = np.where(lum[:, None, None, None] > 0, a[:, 1, ..., :3], 0)  # emitters stay the world's
    return a


@pytest.mark.parametrize("rebuild", [False, True])
def test_textured_frames_parity(world0, rebuild):
    import oracle
    import vrgi
    W, H, B = 96, 64, 3
    atlas = _atlas(len(world0["mats"]), 7, world0["lum"])
    s = oracle.Scene(world0["mats"], world0["lum"])
    s.add_objects(world0["objs"])
    s.set_textures(atlas)
    s.build()
    r = oracle.Renderer(W, H, B)
    cam = _camera0()
    flags = vrgi.FLAG_KEEP_BOUNCES | (vrgi.FLAG_REBUILD_EVERY_FRAME if rebuild else 0)
    with vrgi.Context(vrgi.config(W, H, B, flags=flags), world0["mats"], world0["lum"]) as c:
        c.add_objects(world0["objs"])
        c.set_textures(atlas)
        for f in range(3):
            say(f"textured frame {f} rebuild={rebuild}")
            c.render(cam, nee_type=1, spatial=5)
            c.sync()
            ref = r.render(s, cam, f, nee_type=1, spatial=5)
            _check_frame(c, ref, W * H, B, 1)
        # back to the constant materials: a fresh oracle scene without the atlas
        c.set_textures(None)
        s.set_textures(None)
        s.build()
        say("atlas cleared")
        c.render(cam, nee_type=1, spatial=5)
        c.sync()
        _check_frame(c, r.render(s, cam, 3, nee_type=1, spatial=5), W * H, B, 1)


def test_textured_cache_albedo(world0):
    """Learned-cache frames read the albedo of the textured hit: the frame's records (and
    its per-bounce buffers) equal the oracle's with the atlas, given the GPU's cache
    radiance."""
    import oracle
    import vrgi
    W, H, TB = 96, 64, 5
    atlas = _atlas(len(world0["mats"]), 8, world0["lum"])
    s = oracle.Scene(world0["mats"], world0["lum"])
    s.add_objects(world0["objs"])
    s.set_textures(atlas)
    s.build()
    cam = _camera0()
    with vrgi.Context(vrgi.config(W, H, 3, flags=vrgi.FLAG_KEEP_BOUNCES), world0["mats"], world0["lum"]) as c:
        c.add_objects(world0["objs"])
        c.set_textures(atlas)
        c.cache_enable(log2_table=14, box_lo=(0, 0, 0), box_hi=(0, 0, 0), train_fraction=0.25, train_bounces=TB,
                       batch=4096, steps_per_frame=1, seed=3)
        say("textured cache frame")
        c.render(cam, nee_type=1, spatial=2)
        c.sync()
        crad = c.read(vrgi.BUF_CACHE_RADIANCE).reshape(W * H, 3)
        ref = oracle.Renderer(W, H, TB).render_cache(s, cam, 0, crad, 0.25, oracle.frame_seed(3, 0, 5), nee_type=1,
                                                     spatial=2)
        alb = c.read(vrgi.BUF_BOUNCE_ALBEDO).reshape(TB, W * H, 3)
        plen = c.read(vrgi.BUF_PATH_LEN, np.uint32)
        live = np.arange(TB)[:, None] < plen[None, :]
        assert alb[live].tobytes() == ref["albedo"].reshape(TB, W * H, 3)[live].tobytes()
        assert c.read(vrgi.BUF_RADIANCE).tobytes() == ref["l_out"][0].tobytes()


def test_texture_luminances_and_refusals(world0):
    import oracle
    import vrgi
    L = vrgi.lib()
    atlas = _atlas(len(world0["mats"]), 9, np.ones(len(world0["mats"])))
    out = np.zeros(len(atlas), np.float32)
    vrgi.check(L.vrgi_texture_luminances(len(atlas), 16, 16, vrgi.ptr(atlas), vrgi.ptr(out)))
    assert out.tobytes() == oracle.texture_luminances(atlas).tobytes()
    # ReSTIR DI reads the constant materials: no atlas there
    with vrgi.Context(vrgi.config(32, 32, 3, mode=1), world0["mats"], world0["lum"]) as c:
        with pytest.raises(RuntimeError):
            c.set_textures(atlas)
    c1 = vrgi.Context(vrgi.config(32, 32, 3), world0["mats"], world0["lum"])
    try:
        with pytest.raises(RuntimeError):  # too few slots for the material table
            c1.set_textures(atlas[:2])
        with vrgi.Batch([c1], spp=1):
            with pytest.raises(RuntimeError):  # not while the scene is in a batch
                c1.set_textures(atlas)
    finally:
        c1.close()


def test_textured_scene_in_batch(world0):
    """The textured 64^3 world in a batch beside an untextured scene: three batched frames
    (one graph each once the geometry replays) equal the oracle's textured frames."""
    import oracle
    import vrgi
    W, H = 96, 64
    atlas = _atlas(len(world0["mats"]), 11, world0["lum"])
    s = oracle.Scene(world0["mats"], world0["lum"])
    s.add_objects(world0["objs"])
    s.set_textures(atlas)
    s.build()
    r = oracle.Renderer(W, H, 3)
    cam = _camera0()
    c0 = vrgi.Context(vrgi.config(W, H, 3), world0["mats"], world0["lum"])
    c1 = vrgi.Context(vrgi.config(48, 32, 3), world0["mats"], world0["lum"])
    try:
        c0.add_objects(world0["objs"])
        c0.set_textures(atlas)
        c1.add_objects(world0["objs"][: len(world0["objs"]) // 2])
        with vrgi.Batch([c0, c1], spp=1, rays_per_block=64, pixels_per_block=256) as bt:
            for f in range(3):
                say(f"textured scene in a batch, frame {f}")
                bt.render([cam, cam], nee_type=1, spatial=5)
                bt.sync()
                ref = r.render(s, cam, f, nee_type=1, spatial=5)
                assert c0.read(vrgi.BUF_RADIANCE).tobytes() == ref["l_out"][0].tobytes()
                assert c0.read(vrgi.BUF_RESTIR).tobytes() == ref["restir"].tobytes()
                assert c0.read(vrgi.BUF_BGRA8, np.uint8).tobytes() == ref["bgra8"].tobytes()
    finally:
        c0.close()
        c1.close()
