"""The scene-mutation half of the C ABI against the oracle: Scene::update_object with
rotated isometries, remove_object, dispose_old_objects (scene.rs:150-231),
Renderer::rebuild between frames (interactive_rendering.rs:818) including the
minimised (zero) extent, concurrent SceneUploader::upload_object from 15 threads (the
reference's chunk-worker pool, chunk_manager.rs:214-258), objects released after their
context, and objects of one context refused by another."""
import ctypes as C
import math
import threading

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu

LIT = 4  # an emissive block type of the 64^3 world's materials


def rot(axis, deg, t):
    """Isometry as InstanceData's column-major mat4x3: rotation about a unit axis, then t."""
    a = np.asarray(axis, np.float64)
    a /= np.linalg.norm(a)
    th = math.radians(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    R = np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * (K @ K)
    return np.concatenate([R[:, 0], R[:, 1], R[:, 2], np.asarray(t, np.float64)]).astype(np.float32)


def prop_objects(world0, n=4, seed=3):
    """Small meshes (a few lamp and stone voxels each, chunk-local coordinates 0..8)."""
    import vrgi
    rng = np.random.default_rng(seed)
    nt = world0["n_types"]
    out = []
    for k in range(n):
        b = np.full((32, 32, 32), nt, np.uint16)
        for _ in range(10):
            x, y, z = rng.integers(0, 8, 3)
            b[x, y, z] = LIT if rng.random() < 0.5 else 1
        out.append(vrgi.gen_mesh_chunk(b, nt, world0["tr"], 0, 0, 0))
    return out


def _same_frame(c, ref):
    import vrgi
    assert c.read(vrgi.BUF_RADIANCE).tobytes() == ref["l_out"][0].tobytes()
    assert c.read(vrgi.BUF_RESTIR).tobytes() == ref["restir"].tobytes()
    assert c.read(vrgi.BUF_BGRA8, np.uint8).tobytes() == ref["bgra8"].tobytes()


def test_update_rotated_remove_dispose(world0):
    """Props under rotated isometries are added, moved (vrgi_scene_update with new
    rotations), removed and disposed of between frames; every frame and the light TL-BVH
    equal the oracle's scene under the same edits."""
    import oracle
    import vrgi
    objs = world0["objs"]
    props = prop_objects(world0)
    base_key = 1000
    xfs = [rot((0, 1, 0), 30, (10, 52, 12)), rot((1, 0.3, 0), 45, (30, 50, 40)),
           rot((0.2, 0.5, 1), 75, (44, 54, 20)), rot((1, 1, 1), 120, (20, 56, 30))]
    moves = [rot((0, 1, 0), -20, (14, 50, 16)), rot((1, 0, 0.4), 10, (28, 52, 36)),
             rot((0.5, 0.5, 0.1), 200, (40, 55, 25)), rot((0, 0, 1), 33, (22, 58, 34))]
    s = oracle.Scene(world0["mats"], world0["lum"])
    s.add_objects(objs)
    for k, v in enumerate(props):
        s.add(base_key + k, v, xfs[k])
    s.build()
    W, H, B = 96, 64, 3
    cam = vrgi.camera_look([6.0, 62.0, 8.0], [40.0, 0.0, 44.0])
    r = oracle.Renderer(W, H, B)
    with vrgi.Context(vrgi.config(W, H, B), world0["mats"], world0["lum"]) as c:
        c.add_objects(objs)
        handles = []
        for k, v in enumerate(props):
            o = c.upload(v)
            c.add(base_key + k, o, xfs[k])
            handles.append(o)
        frame = 0

        def step(tag):
            nonlocal frame
            say(f"scene edits: {tag}, frame {frame}")
            c.render(cam, nee_type=1, spatial=5)
            c.sync()
            ref = r.render(s, cam, frame, nee_type=1, spatial=5)
            _same_frame(c, ref)
            tl = c.read(vrgi.BUF_LIGHT_TL_BVH, np.uint8)
            assert tl.tobytes() == s.tl_bvh().tobytes(), tag
            frame += 1

        step("rotated props added")
        for k in range(len(props)):
            c.update(base_key + k, moves[k])
            s.remove(base_key + k)
            s.add(base_key + k, props[k], moves[k])
        s.build()
        step("props moved")
        # remove two props and one terrain chunk with lights; the caller still holds the
        # props' handles, which survive the disposal and can be added again
        lit_key = next(k for k, xf, v in objs if len(v) and (v["t"] // 6 == LIT).any())
        for key in (base_key + 1, base_key + 2, lit_key):
            c.remove(key)
            s.remove(key)
        s.build()
        step("props and a lit chunk removed")
        check = vrgi.check
        check(c.L.vrgi_dispose_old(c.h))
        step("after dispose_old")
        c.add(base_key + 7, handles[2], xfs[0])
        s.add(base_key + 7, props[2], xfs[0])
        s.build()
        step("a removed prop added again under a new key")
        # an unknown key is an error (the reference panics, scene.rs:215)
        assert c.L.vrgi_scene_update(c.h, C.c_uint32(987654), vrgi.ptr(xfs[0])) == vrgi.VRGI_E_INVALID
        for o in handles:
            c.L.vrgi_object_release(o)


def test_objects_outlive_their_context(world0):
    """An object released after vrgi_destroy (a handle dropped late) frees its memory
    without touching the destroyed context; an object of one context is refused by
    another."""
    import vrgi
    props = prop_objects(world0, n=2)
    ident = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32)
    say("object released after its context")
    a = vrgi.Context(vrgi.config(16, 16), world0["mats"], world0["lum"])
    b = vrgi.Context(vrgi.config(16, 16), world0["mats"], world0["lum"])
    try:
        oa = a.upload(props[0])
        ob = b.upload(props[1])
        assert a.L.vrgi_scene_add(a.h, 1, ob, vrgi.ptr(ident)) == vrgi.VRGI_E_INVALID
        a.add(1, oa, ident)
        a.render(vrgi.camera_look([0, 10, 0], [4, 0, 4]), nee_type=1)
        a.sync()
    finally:
        a.close()
    # `oa` is still held here: releasing it after the context is gone is valid
    assert a.L.vrgi_object_release(oa) == 0
    b.L.vrgi_object_release(ob)
    b.close()


def test_resize_between_frames(world0, scenes):
    """Frames before and after Renderer::rebuild to a larger, then a smaller extent, with
    FP64 accumulation on: every frame equals the oracle's renderer of that extent (a fresh
    history at each extent; the frame counter continues), and the accumulated mean restarts
    at each resize. A zero extent renders nothing until the next non-zero one."""
    import oracle
    import vrgi
    s, objs = scenes
    B = 3
    cam = vrgi.camera_look([6.0, 62.0, 8.0], [40.0, 0.0, 44.0])
    with vrgi.Context(vrgi.config(64, 48, B), world0["mats"], world0["lum"]) as c:
        c.add_objects(objs)
        c.accum_enable(True)
        frame = 0
        for W, H, n in [(64, 48, 2), (112, 80, 3), (48, 40, 2)]:
            if frame:
                say(f"resize to {W}x{H}")
                vrgi.check(c.L.vrgi_resize(c.h, W, H))
                c.cfg.width, c.cfg.height = W, H
            r = oracle.Renderer(W, H, B)
            acc = np.zeros((W * H, 3), np.float64)
            for _ in range(n):
                say(f"{W}x{H} frame {frame}")
                out = c.render(cam, nee_type=1, spatial=5)
                assert out.frame_index == frame
                c.sync()
                ref = r.render(s, cam, frame, nee_type=1, spatial=5)
                _same_frame(c, ref)
                lo = ref["l_out"][0].reshape(-1, 3)
                acc += np.where(np.isfinite(lo), lo, 0.0).astype(np.float64)
                frame += 1
            mean, nf = c.accum_read()
            assert nf == n
            assert mean.tobytes() == (acc * (1.0 / n)).tobytes()
        # minimised: no frame, zeroed output, counter unchanged
        vrgi.check(c.L.vrgi_resize(c.h, 0, 40))
        out = c.render(cam, nee_type=1, spatial=5)
        assert not out.radiance and not out.restir and not out.bgra8
        vrgi.check(c.L.vrgi_resize(c.h, 48, 40))
        out = c.render(cam, nee_type=1, spatial=5)
        assert out.frame_index == frame
        c.sync()
        ref = oracle.Renderer(48, 40, B).render(s, cam, frame, nee_type=1, spatial=5)
        _same_frame(c, ref)


def test_concurrent_uploads(world0, scenes):
    """upload_object from 15 threads at once (ctypes releases the GIL, so the uploads
    overlap): every object's light BL-BVH equals the serial upload's byte for byte, and the
    scene built from them renders the oracle's frame."""
    import oracle
    import vrgi
    s, objs = scenes
    objs = [o for o in objs if len(o[2])]
    W, H, B = 64, 48, 3
    cam = vrgi.camera_look([6.0, 62.0, 8.0], [40.0, 0.0, 44.0])

    def light_bvh(c, o):
        n = C.c_size_t()
        vrgi.check(c.L.vrgi_object_light_bvh(o, None, C.c_size_t(0), C.byref(n)))
        out = np.zeros(n.value, vrgi.NODE_DTYPE)
        if n.value:
            vrgi.check(c.L.vrgi_object_light_bvh(o, vrgi.ptr(out), C.c_size_t(n.value), C.byref(n)))
        return out.tobytes()

    with vrgi.Context(vrgi.config(W, H, B), world0["mats"], world0["lum"]) as c:
        serial = [c.upload(v) for _, _, v in objs]
        want = [light_bvh(c, o) for o in serial]
        for o in serial:
            c.L.vrgi_object_release(o)
        got = [None] * len(objs)
        errors = []
        nxt = iter(range(len(objs)))
        lock = threading.Lock()

        def worker():
            while True:
                with lock:
                    i = next(nxt, None)
                if i is None:
                    return
                try:
                    got[i] = c.upload(objs[i][2])
                except Exception as e:  # reported below
                    errors.append(e)
                    return

        say(f"{len(objs)} uploads from 15 threads")
        th = [threading.Thread(target=worker) for _ in range(15)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors, errors
        for i, (key, xf, v) in enumerate(objs):
            assert light_bvh(c, got[i]) == want[i], key
            c.add(key, got[i], xf)
            c.L.vrgi_object_release(got[i])
        say("frame of the concurrently uploaded scene")
        c.render(cam, nee_type=1, spatial=5)
        c.sync()
        ref = oracle.Renderer(W, H, B).render(s, cam, 0, nee_type=1, spatial=5)
        _same_frame(c, ref)
