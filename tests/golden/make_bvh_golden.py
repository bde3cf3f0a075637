"""Writes bvh_grid4.npz: the oracle's build_bl_bvh node array for the exposed faces of a
4x4x4 block of voxels (two triangles per face, mesher order). A regression pin: the
structural checks that do not depend on this file are in tests/test_oracle_bvh.py.
Run: python tests/golden/make_bvh_golden.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))


def grid4_triangles():
    tris = []
    faces = {  # (axis, side) -> 4 corners (unit quad) as offsets
        (0, 0): [(0, 0, 1), (0, 1, 0), (0, 0, 0), (0, 1, 1)],
        (0, 1): [(1, 1, 0), (1, 0, 1), (1, 0, 0), (1, 1, 1)],
        (1, 0): [(0, 0, 0), (1, 0, 0), (0, 0, 1), (1, 0, 1)],
        (1, 1): [(0, 1, 1), (1, 1, 0), (0, 1, 0), (1, 1, 1)],
        (2, 0): [(0, 1, 0), (1, 0, 0), (0, 0, 0), (1, 1, 0)],
        (2, 1): [(0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)],
    }
    for x in range(4):
        for y in range(4):
            for z in range(4):
                for (ax, side), c in faces.items():
                    nb = [x, y, z]
                    nb[ax] += 1 if side else -1
                    if all(0 <= q < 4 for q in nb):
                        continue
                    p = [np.add((x, y, z), o) for o in c]
                    tris.append(np.concatenate([p[0], p[1], p[2]]))
                    tris.append(np.concatenate([p[3], p[1], p[0]]))
    return np.array(tris, np.float32)


def main():
    import oracle
    tri = grid4_triangles()
    n = len(tri)
    lum = np.linspace(0.5, 3.0, n).astype(np.float32)
    ids = np.arange(n, dtype=np.uint32)
    nodes, order, *_ = oracle.build_bl_bvh(tri, lum, ids)
    out = os.path.join(HERE, "bvh_grid4.npz")
    np.savez(out, tri=tri, lum=lum, ids=ids, nodes=nodes, order=order)
    print(out, n)


if __name__ == "__main__":
    main()
