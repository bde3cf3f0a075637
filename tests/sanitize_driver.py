"""Driver of the sanitizer test (run in a subprocess with the sanitizer runtime
preloaded; not collected by pytest). `oracle`: the oracle's scene builds, frames in both
ReSTIR modes, a learned-cache frame whose training records overflow their buffer, the
NEE pdf, the standalone builders and the mesher. `host`: the library's host side
without a GPU -- world generation, the CPU mesher, the random-walk edits, configuration,
and every entry point's argument checks."""
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "vulkan-voxel-restir-gi_amd")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402


def run_oracle():
    import oracle
    import vrgi
    from conftest import random_soup
    # standalone builders and the brute-force plane search on tie-heavy soups
    for n, grid in ((1, False), (2, True), (3, False), (700, True), (5000, False)):
        tri = random_soup(n, n, grid=grid)
        lum = np.linspace(0, 3, n).astype(np.float32)
        nodes, order, box, pw, depth = oracle.build_bl_bvh(tri, lum, np.arange(n, dtype=np.uint32))
        assert len(nodes) == 2 * n - 1
        oracle.find_best_plane(tri)
    xf = np.tile(np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32), (9, 1))
    xf[:, 9:] = np.arange(27, dtype=np.float32).reshape(9, 3)
    boxes = np.tile(np.array([0, 0, 0, 1, 1, 1], np.float32), (9, 1))
    oracle.build_tl_bvh(xf, boxes, np.arange(1, 10, dtype=np.float32), np.arange(9, dtype=np.uint32))
    # mesher on a random world with a translucent type
    rng = np.random.default_rng(3)
    b = rng.integers(0, 6, (64, 32, 32)).astype(np.uint16)
    tr = np.array([0, 0, 0, 0, 1], np.uint8)
    for cx in range(2):
        oracle.gen_mesh_chunk(b, 5, tr, cx, 0, 0)
    # the 64^3 world: GI and DI frames, NEE pdf, a cache frame with a tiny record buffer
    blocks, nt, mats, lum, trw, objs = _world0()
    s = oracle.Scene(mats, lum)
    s.add_objects(objs)
    s.build()
    cam = vrgi.camera_look([6.0, 62.0, 8.0], [40.0, 0.0, 44.0])
    for mode in (0, 1):
        r = oracle.Renderer(48, 32, 3, rr_start=1, mode=mode, radius=30.0 if mode else 15.0, di_candidates=8)
        for f in range(3):
            r.render(s, cam, f, nee_type=1, spatial=5, debug_view=f % 3)
    r = oracle.Renderer(48, 32, 4)
    out = r.render_cache(s, cam, 0, np.ones((48 * 32, 3), np.float32), 0.5, 7, nee_type=1, spatial=2,
                         trec_cap=16)
    assert out["n_trec"] > 16
    o = np.tile(np.array([32.0, 70.0, 32.0], np.float32), (256, 1))
    d = rng.normal(size=(256, 3)).astype(np.float32)
    d[:, 1] = -np.abs(d[:, 1])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    s.nee_pdf(o, np.array([0, -1, 0], np.float32), d)
    s.trace_closest(o, d)
    s.trace_closest(o[:32], d[:32], brute=True)
    s.remove(objs[0][0])
    s.build()
    print("oracle ok")


def _world0():
    import vrgi
    blocks, nt = vrgi.world(0, 0)
    mats, lum = vrgi.world_materials(0, 0)
    tr = vrgi.world_translucent(0, nt)
    return blocks, nt, mats, lum, tr, vrgi.chunk_objects(blocks, nt, tr)


def run_host():
    import vrgi
    L = vrgi.lib()
    assert L.vrgi_abi_version() == 2
    cfg = vrgi.Config()
    L.vrgi_config_reference(C.byref(cfg), 64, 48)
    # worlds, materials, translucency, the CPU mesher, edits
    blocks, nt, mats, lum, tr, objs = _world0()
    assert sum(len(v) for _, _, v in objs) > 0
    for kind in (1, 4):
        dims = np.zeros(3, np.uint32)
        n = C.c_uint32()
        vrgi.check(L.vrgi_world_dims(kind, vrgi.ptr(dims), C.byref(n)))
    b1, nt1 = vrgi.world(1, 0)
    tr1 = vrgi.world_translucent(1, nt1)
    for c in ((0, 0, 0), (3, 1, 5), (7, 3, 7)):
        vrgi.gen_mesh_chunk(b1, nt1, tr1, *c)
    walkers = np.array([[10, 40, 10], [250, 100, 3], [0, 0, 0]], np.uint32)
    vrgi.world_walk(b1, nt1, 5, walkers, 200000)
    # a destination too small is an error, not an overrun
    tr_small = np.zeros(1, np.uint8)
    assert L.vrgi_world_translucent(0, vrgi.ptr(tr_small), 1) != 0 or nt <= 1
    # argument checks of the entry points (no device is touched)
    ctx = C.c_void_p()
    assert L.vrgi_create(None, None, None, 0, 0, C.byref(ctx)) == vrgi.VRGI_E_INVALID
    assert L.vrgi_create(C.byref(cfg), vrgi.ptr(mats), vrgi.ptr(lum), len(lum), 0, C.byref(ctx)) != 0
    for fn, args in [("vrgi_destroy", (None,)), ("vrgi_upload_object", (None, None, 0, None)),
                     ("vrgi_object_release", (None,)), ("vrgi_scene_add", (None, 0, None, None)),
                     ("vrgi_scene_update", (None, 0, None)), ("vrgi_scene_remove", (None, 0)),
                     ("vrgi_dispose_old", (None,)), ("vrgi_resize", (None, 1, 1)),
                     ("vrgi_render", (None, None, None, None)), ("vrgi_sync", (None,)),
                     ("vrgi_build_bl_bvh", (None, None, None, None, 0, None, None, None, None)),
                     ("vrgi_build_bl_bvh_noalloc", (None, None, None, None, 0, None, None, None, None)),
                     ("vrgi_build_tl_bvh", (None, None, None, None, None, 0, None)),
                     ("vrgi_voxels_set", (None, 0, None, None)), ("vrgi_accum_read", (None, None, None))]:
        assert getattr(L, fn)(*args) == vrgi.VRGI_E_INVALID, fn
        assert L.vrgi_last_error()
    print("host ok")


if __name__ == "__main__":
    {"oracle": run_oracle, "host": run_host}[sys.argv[1]]()
