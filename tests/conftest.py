import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vulkan-voxel-restir-gi_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libvrgi.so")


_GPU = None


def gpu_available():
    """True when the HIP runtime reports at least one device."""
    global _GPU
    if _GPU is None:
        import ctypes
        try:
            hip = ctypes.CDLL("libamdhip64.so")
            n = ctypes.c_int(0)
            _GPU = hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
        except OSError:
            _GPU = False
    return _GPU


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


def say(msg):
    """Names the GPU case about to run (flushed before the device call)."""
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="session")
def world0():
    """64^3 benchmark world (seed 0): blocks, n_types, materials, luminances, chunk objects."""
    import vrgi
    blocks, nt = vrgi.world(0, 0)
    mats, lum = vrgi.world_materials(0, 0)
    tr = vrgi.world_translucent(0, nt)
    objs = vrgi.chunk_objects(blocks, nt, tr)
    return dict(blocks=blocks, n_types=nt, mats=mats, lum=lum, tr=tr, objs=objs)


@pytest.fixture(scope="session")
def scenes(world0):
    """The oracle's scene of the 64^3 world (built once) and its chunk objects."""
    import oracle
    objs = world0["objs"]
    s = oracle.Scene(world0["mats"], world0["lum"])
    s.add_objects(objs)
    s.build()
    return s, objs


def random_soup(n, seed, spread=50.0, size=2.0, grid=False):
    """Random triangles; grid=True snaps vertices to integers (voxel-like, many ties)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, spread, (n, 1, 3))
    v = c + rng.uniform(-size, size, (n, 3, 3))
    if grid:
        v = np.round(v)
    return v.reshape(n, 9).astype(np.float32)
