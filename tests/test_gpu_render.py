"""GPU parity of the traversal and of whole frames against the oracle.

Closest hit (raytrace.rs:354-388 replaced by software traversal): primitive id exact,
t within 1e-6 relative (the tests require bit equality, which is stronger). Whole
frames of the 64^3 world: path lengths, hit ids and every per-bounce buffer of the live
bounces bit-exact, radiance / ReSTIR reservoirs / final image bit-exact, over three
frames so that the 3-slot temporal history is exercised."""
import ctypes

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _rays(n, seed, lo=-4.0, hi=68.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def test_closest_hit_parity(world0, scenes):
    import vrgi
    s, objs = scenes
    o, d = _rays(200000, 1)
    say("closest hit 200k random rays, 64^3 world")
    with vrgi.Context(vrgi.config(16, 16), world0["mats"], world0["lum"]) as c:
        c.add_objects(objs)
        gh, gt, guv = c.trace_closest(o, d)
    oh, ot, ouv = s.trace_closest(o, d)
    assert np.array_equal(gh, oh)
    hit = oh != 0xFFFFFFFF
    assert hit.mean() > 0.3
    np.testing.assert_allclose(gt[hit], ot[hit], rtol=1e-6, atol=0)
    assert gt[hit].tobytes() == ot[hit].tobytes()
    assert guv[hit].tobytes() == ouv[hit].tobytes()
    # the oracle BVH walk itself agrees with brute force on a subset
    bh, bt, _ = s.trace_closest(o[:2000], d[:2000], brute=True)
    assert np.array_equal(bh, oh[:2000])


def test_watertight_edges(world0, scenes):
    """Rays aimed exactly at shared voxel edges and corners: no ray leaks through."""
    import vrgi
    s, objs = scenes
    rng = np.random.default_rng(5)
    n = 20000
    # targets on integer lattice edges of the heightmap top surfaces, eye above the world
    tx = rng.integers(1, 63, n).astype(np.float32)
    tz = rng.integers(1, 63, n).astype(np.float32) + rng.choice([0.0, 0.5], n).astype(np.float32)
    eye = np.stack([rng.uniform(0, 64, n), np.full(n, 90.0), rng.uniform(0, 64, n)], 1).astype(np.float32)
    tgt = np.stack([tx, np.full(n, 8.0, np.float32), tz], 1)
    d = tgt - eye
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    say("watertight edge rays")
    with vrgi.Context(vrgi.config(16, 16), world0["mats"], world0["lum"]) as c:
        c.add_objects(objs)
        gh, gt, _ = c.trace_closest(eye, d.astype(np.float32))
    oh, ot, _ = s.trace_closest(eye, d.astype(np.float32))
    assert np.array_equal(gh, oh)
    # every ray aimed below the minimum terrain height (8) hits something
    assert np.all(gh != 0xFFFFFFFF)


def _camera0():
    import vrgi
    return vrgi.camera_look([6.0, 62.0, 8.0], [40.0, 0.0, 44.0])


NO_RR = 0xFFFFFFFF


def _check_frame(c, ref, P, B, nee_type, full_nee=False):
    """Every per-bounce buffer of the live bounces, the radiance, the reservoirs and the
    image of the frame just rendered by context c equal the oracle's frame ref bit for bit.
    The NEE pdf buffer: bit for bit with VRGI_FLAG_FULL_NEE_PDF; by default (the lazy NEE
    pdf) bit for bit on every ray whose next vertex returns radiance, and 0 or bit for bit
    on the others, whose pdf cannot reach any output."""
    import vrgi
    plen = c.read(vrgi.BUF_PATH_LEN, np.uint32)
    assert np.array_equal(plen, ref["path_len"])
    live = np.arange(B)[:, None] < plen[None, :]  # [B, P]
    hit = c.read(vrgi.BUF_HIT_PRIM, np.uint32).reshape(B, P)
    assert np.array_equal(hit[live], ref["hit_prim"][live])
    for buf, key, k in [(vrgi.BUF_BOUNCE_ORIGIN, "origin", 3), (vrgi.BUF_BOUNCE_DIRECTION, "direction", 3)]:
        g = c.read(buf).reshape(B, P, k)
        o = ref[key][1:]
        assert g[live].tobytes() == o[live].tobytes(), key
    for buf, key, k in [(vrgi.BUF_BOUNCE_NORMAL, "normal", 3), (vrgi.BUF_BOUNCE_EMISSIVITY, "emissivity", 3),
                        (vrgi.BUF_BOUNCE_REFLECTIVITY, "reflectivity", 3),
                        (vrgi.BUF_BOUNCE_MIS_WEIGHT, "mis_weight", 1), (vrgi.BUF_BOUNCE_BSDF_PDF, "bsdf_pdf", 1)]:
        g = c.read(buf).reshape(B, P, k)
        o = ref[key].reshape(B, P, k)
        assert g[live].tobytes() == o[live].tobytes(), key
    if nee_type:
        cont = np.arange(B)[:, None] + 1 < plen[None, :]
        g = c.read(vrgi.BUF_BOUNCE_NEE_PDF).reshape(B, P)
        same = g.view(np.uint32) == ref["nee_pdf"].view(np.uint32)
        if full_nee:
            assert np.all(same[cont])
        else:
            lz = np.zeros((B, P), bool)
            lz[:-1] = np.all(ref["l_out"][1:] == 0, axis=2)  # radiance of the next vertex is 0
            assert np.all(same[cont & ~lz])
            assert np.all((same | (g == 0))[cont & lz])
            assert (~same[cont]).any() or not (ref["nee_pdf"][cont & lz] != 0).any()
        assert (ref["nee_pdf"] > 0).any() or nee_type == 2
    assert c.read(vrgi.BUF_RADIANCE).tobytes() == ref["l_out"][0].tobytes()
    assert c.read(vrgi.BUF_TEMPORAL_RESERVOIR).tobytes() == ref["temporal"].tobytes()
    assert c.read(vrgi.BUF_SPATIAL_RESERVOIR).tobytes() == ref["spatial"].tobytes()
    assert c.read(vrgi.BUF_RESTIR).tobytes() == ref["restir"].tobytes()
    assert c.read(vrgi.BUF_BGRA8, np.uint8).tobytes() == ref["bgra8"].tobytes()
    return plen


@pytest.mark.parametrize("nee_type,spatial,rr_start,sort_bits,serial,full", [
    (0, 0, NO_RR, 0, False, False), (1, 5, NO_RR, 0, False, False), (2, 3, NO_RR, 0, False, False),
    # the NEE pdf of every ray: beside the next bounce's trace, or after all bounces on the
    # frame stream
    (1, 5, NO_RR, 0, False, True), (1, 5, NO_RR, 0, True, True), (1, 2, 1, 0, True, True),
    (1, 2, 1, 0, False, False),
    # Morton-ordered ray queues (with and without Russian roulette) give the same bits
    (1, 5, NO_RR, 4, False, False), (1, 2, 1, 3, False, True), (0, 1, 1, 5, False, False)])
def test_frames_parity(world0, scenes, nee_type, spatial, rr_start, sort_bits, serial, full):
    import oracle
    import vrgi
    s, objs = scenes
    W, H, B = 96, 64, 3
    cam = _camera0()
    r = oracle.Renderer(W, H, B, rr_start=rr_start)
    cfg = vrgi.config(W, H, B, rr_start=rr_start, sort_bits=sort_bits,
                      flags=vrgi.FLAG_KEEP_BOUNCES | (vrgi.FLAG_SERIAL_NEE if serial else 0)
                      | (vrgi.FLAG_FULL_NEE_PDF if full else 0))
    with vrgi.Context(cfg, world0["mats"], world0["lum"]) as c:
        c.add_objects(objs)
        for f in range(4):
            say(f"frame {f} nee={nee_type} spatial={spatial} rr={rr_start} sort={sort_bits} serial={serial} "
                f"full={full}")
            c.render(cam, nee_type=nee_type, spatial=spatial)
            c.sync()
            ref = r.render(s, cam, f, nee_type=nee_type, spatial=spatial)
            _check_frame(c, ref, W * H, B, nee_type, full_nee=full)


def _cubes_crossed(o, d, lo, tmin=1e-4, tmax=1000.0):
    """Number of unit cubes [lo, lo + 1] each ray crosses within [tmin, tmax] (slab test)."""
    n = np.zeros(len(o), np.int64)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d.astype(np.float64)
        for k in range(0, len(lo), 512):
            c = lo[None, k:k + 512, :].astype(np.float64)
            t0 = (c - o[:, None, :]) * inv[:, None, :]
            t1 = (c + 1.0 - o[:, None, :]) * inv[:, None, :]
            tn = np.maximum(np.nanmax(np.minimum(t0, t1), axis=2), tmin)
            tf = np.minimum(np.nanmin(np.maximum(t0, t1), axis=2), tmax)
            n += (tn < tf).sum(axis=1)
    return n


def test_frames_parity_many_light_crossings(world0):
    """Four layers of separated emissive blocks above the terrain: the NEE pdf rays cross
    many light triangles each (more than a ray keeps for the walk pass, so some rays take
    the overflow pass), and the frames stay bit-exact against the oracle, whose sum runs
    over the crossings in the order they are met."""
    import oracle
    import vrgi
    blocks = world0["blocks"].copy()
    nt = world0["n_types"]
    lit = 4  # first emissive block type of the 64^3 world's materials
    assert world0["lum"][6 * lit] > 0
    cubes = []
    for y in (50, 53, 56, 59):
        for x in range(2, 62, 2):
            for z in range(2, 62, 2):
                blocks[x, y, z] = lit
                cubes.append((x, y, z))
    objs = vrgi.chunk_objects(blocks, nt, world0["tr"])
    s = oracle.Scene(world0["mats"], world0["lum"])
    s.add_objects(objs)
    s.build()
    W, H, B = 96, 64, 3
    cam = vrgi.camera_look([6.0, 48.0, 8.0], [40.0, 0.0, 44.0])
    r = oracle.Renderer(W, H, B)
    cfg = vrgi.config(W, H, B, flags=vrgi.FLAG_KEEP_BOUNCES | vrgi.FLAG_FULL_NEE_PDF)
    with vrgi.Context(cfg, world0["mats"], world0["lum"]) as c:
        c.add_objects(objs)
        for f in range(2):
            say(f"many light crossings: frame {f}")
            c.render(cam, nee_type=1, spatial=5)
            c.sync()
            ref = r.render(s, cam, f, nee_type=1, spatial=5)
            plen = _check_frame(c, ref, W * H, B, 1, full_nee=True)
    # the scene does what it is for: many pdf rays cross 3 or more emissive cubes, i.e.
    # at least 6 light triangles, and some 5 or more (10 triangles, above the 8 a ray
    # keeps for the walk pass)
    P = W * H
    cont = (1 < plen) & (ref["mis_weight"][0] != 0)
    o = ref["origin"][1][cont]
    d = ref["direction"][1][cont]
    n = _cubes_crossed(o[:4000], d[:4000], np.array(cubes, np.float64))
    assert (n >= 3).sum() >= 50, np.bincount(n)
    assert (n >= 5).sum() >= 10, np.bincount(n)


@pytest.mark.parametrize("serial_build", [False, True])
def test_back_to_back_rebuild_frames(world0, scenes, serial_build):
    """Frames issued without waiting, each rebuilding the BVHs. By default frame f + 1
    builds into the other buffer set on a second stream while frame f's rays still read
    the first. The FP64 running mean of the frames equals the mean of the oracle's
    frames bit for bit (non-finite samples count as 0), and the last frame's reservoirs
    and image are bit-exact."""
    import oracle
    import vrgi
    s, objs = scenes
    W, H, B, n = 96, 64, 3, 5
    cam = _camera0()
    r = oracle.Renderer(W, H, B)
    flags = vrgi.FLAG_REBUILD_EVERY_FRAME | (vrgi.FLAG_SERIAL_BUILD if serial_build else 0)
    say(f"{n} back-to-back frames with a rebuild each, serial build {serial_build}")
    with vrgi.Context(vrgi.config(W, H, B, flags=flags), world0["mats"], world0["lum"]) as c:
        c.add_objects(objs)
        c.accum_enable(True)
        for f in range(n):
            c.render(cam, nee_type=1, spatial=5)
        c.sync()
        mean, frames = c.accum_read()
        restir = c.read(vrgi.BUF_RESTIR)
        bgra = c.read(vrgi.BUF_BGRA8, np.uint8)
    assert frames == n
    acc = np.zeros((W * H, 3), np.float64)
    for f in range(n):
        ref = r.render(s, cam, f, nee_type=1, spatial=5)
        lo = ref["l_out"][0].reshape(-1, 3)
        acc += np.where(np.isfinite(lo), lo, 0.0).astype(np.float64)
    want = acc * (1.0 / n)
    assert mean.tobytes() == want.tobytes()
    assert restir.tobytes() == ref["restir"].tobytes()
    assert bgra.tobytes() == ref["bgra8"].tobytes()


def test_restir_di_parity(world0, scenes):
    """ReSTIR DI mode (32 candidates, 1 temporal, 5 spatial neighbours within 30 px):
    reservoir light ids exact and weights (w_sum, W, M, p_hat) within 1e-5 relative
    (the test requires bit equality), final DI radiance bit-exact, over 3 frames."""
    import oracle
    import vrgi
    s, objs = scenes
    W, H, B = 96, 64, 3
    cam = _camera0()
    r = oracle.Renderer(W, H, B, mode=1, radius=30.0)
    cfg = vrgi.config(W, H, B, mode=1, radius=30.0)
    with vrgi.Context(cfg, world0["mats"], world0["lum"]) as c:
        c.add_objects(objs)
        for f in range(3):
            say(f"DI frame {f}")
            c.render(cam, nee_type=0, spatial=5)
            c.sync()
            ref = r.render(s, cam, f, nee_type=0, spatial=5)
            gl = c.read(vrgi.BUF_DI_LIGHT, np.uint32)
            assert np.array_equal(gl, ref["di_light"])
            gw = c.read(vrgi.BUF_DI_WEIGHTS).reshape(-1, 4)
            np.testing.assert_allclose(gw, ref["di_weights"], rtol=1e-5, atol=0)
            assert gw.tobytes() == ref["di_weights"].tobytes()
            assert c.read(vrgi.BUF_RESTIR).tobytes() == ref["restir"].tobytes()
            assert (gl != 0xFFFFFFFF).mean() > 0.1


@pytest.mark.parametrize("view", [1, 2])
def test_debug_views_parity(world0, scenes, view):
    """The postprocess views (postprocess.rs:39-88): 1 = ReSTIR radiance, 2 = the debug
    buffer (spatial W / 10), as BGRA8 bit-exact against the oracle over 3 frames."""
    import oracle
    import vrgi
    s, objs = scenes
    W, H, B = 96, 64, 3
    cam = _camera0()
    r = oracle.Renderer(W, H, B)
    with vrgi.Context(vrgi.config(W, H, B), world0["mats"], world0["lum"]) as c:
        c.add_objects(objs)
        for f in range(3):
            say(f"debug view {view} frame {f}")
            c.render(cam, nee_type=1, debug_view=view, spatial=5)
            c.sync()
            ref = r.render(s, cam, f, nee_type=1, debug_view=view, spatial=5)
            img = c.read(vrgi.BUF_BGRA8, np.uint8)
            assert img.tobytes() == ref["bgra8"].tobytes()
            assert len(np.unique(img.reshape(-1, 4)[:, :3])) > 4


def test_tonemap_and_accumulate_views(world0):
    """View 3 (ReSTIR radiance, Reinhard per channel) and view 4 (the path tracer's
    running mean since the camera, the scene or the view last changed, tonemapped):
    BGRA8 bit-exact against the oracle through restarts by a view change, a camera move
    and a scene edit."""
    import oracle
    import vrgi
    objs = world0["objs"]
    s = oracle.Scene(world0["mats"], world0["lum"])
    s.add_objects(objs)
    s.build()
    W, H, B = 64, 48, 3
    cam0 = _camera0()
    cam1 = vrgi.camera_look([10.0, 60.0, 6.0], [36.0, 0.0, 40.0])
    r = oracle.Renderer(W, H, B)
    lit_key = next(k for k, xf, v in objs if len(v))
    steps = [(4, cam0, None), (4, cam0, None), (4, cam0, None), (3, cam0, None), (4, cam0, None), (4, cam0, None),
             (4, cam1, None), (4, cam1, "edit"), (4, cam1, None)]
    with vrgi.Context(vrgi.config(W, H, B), world0["mats"], world0["lum"]) as c:
        c.add_objects(objs)
        imgs = []
        for f, (view, cam, edit) in enumerate(steps):
            if edit:
                c.remove(lit_key)
                s.remove(lit_key)
                s.build()
            say(f"view {view} frame {f}{' after an edit' if edit else ''}")
            c.render(cam, nee_type=1, debug_view=view, spatial=5)
            c.sync()
            ref = r.render(s, cam, f, nee_type=1, debug_view=view, spatial=5)
            img = c.read(vrgi.BUF_BGRA8, np.uint8)
            assert img.tobytes() == ref["bgra8"].tobytes(), f
            imgs.append(img)
        assert imgs[2].tobytes() != imgs[0].tobytes()  # the mean moved on
        bad = vrgi.Prefs(1, 5, 0)
        assert vrgi.lib().vrgi_render(c.h, ctypes.byref(cam0), ctypes.byref(bad), None) == vrgi.VRGI_E_INVALID
