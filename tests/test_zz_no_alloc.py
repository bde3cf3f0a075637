"""The allocation-free build_bl_bvh entry point (vrgi_build_bl_bvh_noalloc) against the
allocating one (vrgi_build_bl_bvh), byte for byte: node arrays, primitive order, root box
and power, on calls in a row of decreasing size on the one per-GPU scratch array (the
second call sees what the first left there), up to the documented maximum, and the
limit's error above it (checked on the host, before any device work)."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_soup, say

pytestmark = pytest.mark.gpu


def _build(c, fn, tri, lum, ids):
    import vrgi
    n = len(tri)
    nodes = np.zeros(2 * n - 1, vrgi.NODE_DTYPE)
    order = np.zeros(n, np.uint32)
    box = np.zeros(6, np.float32)
    pw = C.c_float()
    code = fn(c.h, vrgi.ptr(tri), vrgi.ptr(lum), vrgi.ptr(ids), C.c_uint64(n), vrgi.ptr(nodes), vrgi.ptr(order),
              vrgi.ptr(box), C.byref(pw))
    return code, nodes, order, box, pw.value


def test_noalloc_build_matches_allocating_build(world0):
    import vrgi
    L = vrgi.lib()
    maxn = 65536
    with vrgi.Context(vrgi.config(16, 16), world0["mats"], world0["lum"]) as c:
        for k, (n, grid) in enumerate([(50000, True), (3000, False), (maxn, True), (1, False), (700, True)]):
            tri = random_soup(n, 100 + k, spread=80.0, grid=grid)
            rng = np.random.default_rng(k)
            lum = np.where(rng.random(n) < 0.5, rng.uniform(0.1, 5.0, n), 0.0).astype(np.float32)
            ids = rng.permutation(n).astype(np.uint32)
            say(f"allocation-free build call {k}: {n} primitives")
            a = _build(c, L.vrgi_build_bl_bvh_noalloc, tri, lum, ids)
            b = _build(c, L.vrgi_build_bl_bvh, tri, lum, ids)
            assert a[0] == 0 and b[0] == 0, vrgi.lib().vrgi_last_error()
            for x, y in zip(a[1:4], b[1:4]):
                assert x.tobytes() == y.tobytes()
            assert np.float32(a[4]).tobytes() == np.float32(b[4]).tobytes()
        tri = random_soup(maxn + 1, 7)
        z = np.zeros(maxn + 1, np.float32)
        code = _build(c, L.vrgi_build_bl_bvh_noalloc, tri, z, z.view(np.uint32))[0]
        assert code == vrgi.VRGI_E_LIMIT
