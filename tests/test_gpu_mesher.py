"""GPU exposed-face mesher (chunk.rs:158-292) against the oracle's mesher, vertex for vertex
(the primitive order every BVH is built over), on both benchmark worlds and on a world
with translucent blocks and edits."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _check_world(blocks, nt, tr, chunks):
    """GPU mesher output == the oracle's restatement of chunk::gen_mesh, and == the
    library's own CPU mesher (the caller-side chunk path), chunk by chunk."""
    import oracle
    import vrgi
    with vrgi.Context(vrgi.config(16, 16), *vrgi.world_materials(0, 0)) as c:
        got = c.mesh_chunks(blocks, nt, tr, chunks)
    for (cx, cy, cz), g in zip(chunks, got):
        ref = oracle.gen_mesh_chunk(blocks, nt, tr, cx, cy, cz)
        assert g.tobytes() == ref.tobytes(), (cx, cy, cz)
        assert vrgi.gen_mesh_chunk(blocks, nt, tr, cx, cy, cz).tobytes() == ref.tobytes(), (cx, cy, cz)
    return sum(len(g) for g in got)


@pytest.mark.parametrize("kind", [0, 1])
def test_gpu_mesher_matches_cpu(kind):
    import vrgi
    blocks, nt = vrgi.world(kind, 0)
    tr = vrgi.world_translucent(kind, nt)
    nx, ny, nz = (s // 32 for s in blocks.shape)
    chunks = [(x, y, z) for x in range(nx) for y in range(ny) for z in range(nz)]
    say(f"GPU mesher, world kind {kind}, {len(chunks)} chunks")
    n = _check_world(blocks, nt, tr, chunks)
    assert n > 0


def test_gpu_mesher_edits_and_translucency():
    """Random voxel toggles (the dynamic workload's edits) and translucent types."""
    import vrgi
    rng = np.random.default_rng(7)
    blocks, nt = vrgi.world(0, 0)
    blocks = blocks.copy()
    idx = rng.integers(0, blocks.size, blocks.size // 100)
    flat = blocks.reshape(-1)
    flat[idx] = np.where(flat[idx] == nt, rng.integers(0, nt, len(idx)), nt).astype(np.uint16)
    tr = (rng.uniform(size=nt) < 0.3).astype(np.uint8)
    chunks = [(1, 0, 1), (0, 1, 0), (1, 1, 1), (0, 0, 0)]
    say("GPU mesher after 1% toggles, translucent types")
    _check_world(blocks, nt, tr, chunks)
