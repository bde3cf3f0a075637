"""CPU checks of the library's host side: the exposed-face mesher against the oracle's
restatement of chunk.rs:158-292 and against geometry (every triangle's normal points
into the air neighbour), and the C ABI surface (every symbol of include/vrgi.h is
exported; argument errors are reported without touching a device)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _random_world(seed, dims=(64, 40, 48), n_types=5):
    rng = np.random.default_rng(seed)
    air = n_types
    b = rng.integers(0, n_types + 1, dims).astype(np.uint16)
    b[rng.random(dims) < 0.45] = air
    tr = np.zeros(n_types, np.uint8)
    tr[n_types - 1] = 1  # one translucent block type (glass-like)
    return b, n_types, tr


@pytest.mark.parametrize("seed", [0, 1])
def test_mesher_matches_oracle(seed):
    import oracle
    import vrgi
    b, nt, tr = _random_world(seed)
    for cx in range(2):
        for cy in range(2):
            for cz in range(2):
                g = vrgi.gen_mesh_chunk(b, nt, tr, cx, cy, cz)
                o = oracle.gen_mesh_chunk(b, nt, tr, cx, cy, cz)
                assert g.tobytes() == o.tobytes()


def test_mesher_world0_matches_oracle(world0):
    import oracle
    import vrgi
    b, nt, tr = world0["blocks"], world0["n_types"], world0["tr"]
    for key, xf, v in world0["objs"]:
        cx, cy, cz = (int(x) // 32 for x in xf[9:])
        assert v.tobytes() == oracle.gen_mesh_chunk(b, nt, tr, cx, cy, cz).tobytes()


def test_mesher_normals_face_air():
    """Geometric pin independent of the face tables: each triangle's winding normal
    cross(v1 - v0, v2 - v0) is axis-aligned, points out of its voxel, and the voxel on
    that side is air or translucent; two triangles per face, 6 vertices."""
    import vrgi
    b, nt, tr = _random_world(3, dims=(32, 32, 32))
    v = vrgi.gen_mesh_chunk(b, nt, tr, 0, 0, 0)
    assert len(v) % 6 == 0
    p = v["position"].reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    assert np.all(np.count_nonzero(n, axis=1) == 1)
    c = p.mean(1)
    blk = v["t"][::3] // 6
    face = v["t"][::3] % 6
    axis = np.abs(n).argmax(1)
    sign = np.sign(n[np.arange(len(n)), axis])
    assert np.array_equal(axis, face // 2) and np.array_equal(sign, np.where(face % 2, 1, -1))
    vox = np.floor(c - 0.5 * n / np.abs(n).max(1, keepdims=True) * 0.5).astype(int)
    nb = vox + (n / np.abs(n).max(1, keepdims=True)).astype(int)
    assert np.array_equal(b[vox[:, 0], vox[:, 1], vox[:, 2]], blk)
    inside = np.all((nb >= 0) & (nb < 32), 1)
    nbb = np.full(len(nb), nt)
    nbb[inside] = b[nb[inside, 0], nb[inside, 1], nb[inside, 2]]
    assert np.all((nbb == nt) | (tr[np.minimum(nbb, nt - 1)] == 1))


def test_abi_exports_every_declared_symbol():
    import vrgi
    hdr = open(os.path.join(ROOT, "include", "vrgi.h")).read()
    names = sorted(set(re.findall(r"\b(vrgi_[a-z0-9_]+)\s*\(", hdr)))
    assert len(names) >= 27
    L = vrgi.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_argument_errors_without_device():
    import vrgi
    L = vrgi.lib()
    cfg = vrgi.Config()
    L.vrgi_config_reference(ctypes.byref(cfg), 1920, 1080)
    assert (cfg.num_bounces, cfg.restir_spatial_radius, cfg.rr_start_bounce) == (3, 15.0, 0xFFFFFFFF)
    assert L.vrgi_abi_version() == 2
    h = ctypes.c_void_p()
    assert L.vrgi_create(None, None, None, 0, 0, ctypes.byref(h)) == -1
    assert b"null" in L.vrgi_last_error()
    bad = vrgi.Config()
    L.vrgi_config_reference(ctypes.byref(bad), 0, 10)
    assert L.vrgi_create(ctypes.byref(bad), None, None, 0, 0, ctypes.byref(h)) == -1
    assert L.vrgi_upload_object(None, None, 0, ctypes.byref(h)) == -1
    assert L.vrgi_render(None, None, None, None) == -1
    assert L.vrgi_scene_update(None, 0, None) == -1
