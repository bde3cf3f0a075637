"""Oracle pins for the per-pixel kernels (raygen / raytrace / nee_pdf /
outgoing_radiance / ReSTIR) by closed forms and invariants that do not restate the
oracle's code:

- sky only: radiance = 50 where the camera ray has d.y > 0.9 (raytrace.rs:425), else 0;
- a Lambert floor under the sky: E[L] = (100 a / pi) * int_0.9^1 c dc (uniform hemisphere
  sampling, bsdf_pdf = cos/pi, q = 1/(2 pi): the reference's estimator, including its
  extra 1/pi), checked statistically;
- a mirror floor: L = albedo * 50 * 2 pi exactly where the reflected ray has y > 0.9
  (bsdf_pdf 1, q = 1/(2 pi), raytrace.rs:482-489, outgoing_radiance.rs:82-88);
- light BVH: the reverse-walk probabilities of all light primitives sum to 1, and the
  forward walk's pick frequencies follow them (chi-square);
- ReSTIR GI reservoirs: after the first frame on an empty history, M = 1, z = the
  initial sample and W = 1 / p_omega; the 3-slot history gives M in {1, 2} three frames
  later with the reset probability 0.4 (restir_temporal_resampling.rs:267-272).
"""
import math

import numpy as np
import pytest

F32 = np.float32
VERT = np.dtype([("position", "<f4", 3), ("t", "<u4"), ("uv", "<f4", 2)])
MAT = np.dtype([("reflectivity", "<f4", 4), ("emissivity", "<f4", 4), ("metallicity", "<f4", 4)])
IDENT = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], F32)


def mats_of(albedo=0.5, metal=0.0, alpha=1.0, emis=0.0):
    m = np.zeros(6, MAT)
    m["reflectivity"] = [albedo, albedo, albedo, alpha]
    m["emissivity"] = [emis, emis, emis, 1.0]
    m["metallicity"] = [metal, 0, 0, 1]
    lum = np.full(6, 255.0 * 3 * emis, F32)
    return m, lum


def floor_verts(size=4000.0, y=0.0):
    s = size / 2
    p = [(-s, y, -s), (-s, y, s), (s, y, -s), (s, y, -s), (-s, y, s), (s, y, s)]  # normals +y
    v = np.zeros(6, VERT)
    v["position"] = p
    v["t"] = 3  # the "up" face slot of block 0
    return v


def camera(eye, target):
    import vrgi
    return vrgi.camera_look(eye, target)


def test_sky_only():
    import oracle
    m, lum = mats_of()
    s = oracle.Scene(m, lum)
    s.build()
    W, H = 64, 48
    cam = camera([0, 0, 0], [0.3, 1.0, 0.2])
    out = oracle.Renderer(W, H, 3).render(s, cam, 0)
    d = out["direction"][0]
    rad = out["l_out"][0][:, 0]
    far = np.abs(d[:, 1] - 0.9) > 1e-5
    assert np.array_equal(rad[far], np.where(d[far, 1] > 0.9, 50.0, 0.0).astype(F32))
    assert np.all(out["path_len"] == 1)
    assert (d[:, 1] > 0.9).any() and (d[:, 1] < 0.9).any()


def test_lambert_floor_expectation():
    import oracle
    a = 0.6
    m, lum = mats_of(albedo=a)
    s = oracle.Scene(m, lum)
    s.add(0, floor_verts(), IDENT)
    s.build()
    W, H = 160, 120
    cam = camera([0, 10, 0], [0.01, 0.0, 0.02])
    L = oracle.Renderer(W, H, 2).render(s, cam, 0)["l_out"][0][:, 0].astype(np.float64)
    k = 100.0 * a / math.pi
    mean = k * (1 - 0.81) / 2
    var = k * k * (1 - 0.729) / 3 - mean * mean
    assert abs(L.mean() - mean) < 4 * math.sqrt(var / L.size)


def test_mirror_floor_closed_form():
    import oracle
    alb = 0.7
    m, lum = mats_of(albedo=alb, metal=1.0)
    s = oracle.Scene(m, lum)
    s.add(0, floor_verts(), IDENT)
    s.build()
    W, H = 48, 48
    cam = camera([0, 10, 0], [0.2, 0.0, 0.1])
    out = oracle.Renderer(W, H, 2).render(s, cam, 0)
    d1 = out["direction"][1]
    rad = out["l_out"][0][:, 0]
    assert np.all(out["hit_prim"][0] != 0xFFFFFFFF)
    np.testing.assert_allclose(d1[:, 1], -out["direction"][0][:, 1], rtol=1e-6)
    far = np.abs(d1[:, 1] - 0.9) > 1e-5
    want = np.where(d1[far, 1] > 0.9, alb * 50.0 * 2 * math.pi, 0.0)
    np.testing.assert_allclose(rad[far], want, rtol=2e-6)
    assert (d1[:, 1] > 0.9).any()


@pytest.fixture(scope="module")
def lit_scene(world0):
    import oracle
    s = oracle.Scene(world0["mats"], world0["lum"])
    s.add_objects(world0["objs"])
    s.build()
    return s


def _light_leaves(s):
    pairs = s.read(7, np.uint32, 2 * s.count(1)).reshape(-1, 2)
    out = []
    for inst, j in pairs:
        # the light vertex of prim j names its BL leaf; recover it from the light-BVH leaves
        out.append((int(inst), int(j)))
    return pairs


def test_reverse_probabilities_sum_to_one(world0, lit_scene):
    s = lit_scene
    keys = [k for k, xf, v in world0["objs"] if len(v)]
    pairs = s.read(7, np.uint32, 2 * s.count(1)).reshape(-1, 2)
    # a point high above the world looking down sees every light box corner with positive importance
    p = np.array([32.0, 400.0, 32.0], F32)
    n = np.array([0.0, -1.0, 0.0], F32)
    total = 0.0
    per_inst = {}
    for inst in np.unique(pairs[:, 0]):
        key = keys[inst]
        bl = s.light_bvh(key)
        leaves = np.nonzero(bl["left"] == 0xFFFFFFFF)[0]
        for leaf in leaves:
            pr = s.reverse_traverse_light_bvh(p, n, int(inst), int(leaf))
            total += pr
            per_inst[(int(inst), int(bl["right"][leaf]))] = pr
    assert abs(total - 1.0) < 1e-4
    # forward picks follow the same distribution (chi-square over 20000 walks)
    rng_seeds = (np.arange(20000, dtype=np.uint64) * np.uint64(2654435761)) % np.uint64(2 ** 32)
    counts = {}
    for sd in rng_seeds:
        ok, inst, prim, prob, imp = s.traverse_light_bvh(p, n, int(sd))
        assert ok
        counts[(inst, prim)] = counts.get((inst, prim), 0) + 1
        np.testing.assert_allclose(prob, per_inst[(inst, prim)], rtol=1e-4)
    exp = np.array([per_inst[k] * len(rng_seeds) for k in per_inst])
    obs = np.array([counts.get(k, 0) for k in per_inst])
    big = exp > 20
    chi2 = (((obs[big] - exp[big]) ** 2) / exp[big]).sum()
    dof = big.sum() - 1
    assert chi2 < dof + 6 * math.sqrt(2 * dof)


def test_reservoir_invariants(world0, lit_scene):
    import oracle
    s = lit_scene
    W, H = 64, 48
    r = oracle.Renderer(W, H, 3)
    cam = camera([6.0, 62.0, 8.0], [40.0, 0.0, 44.0])
    f0 = r.render(s, cam, 0, spatial=0)
    T = f0["temporal"]
    S_lo = f0["l_out"][1]
    ph = 0.2126 * S_lo[:, 0] + 0.7152 * S_lo[:, 1] + 0.0722 * S_lo[:, 2]
    m = T[:, 18].view(np.uint32)
    pos = ph > 0
    assert np.all(m == 1)
    np.testing.assert_array_equal(T[pos, 12:15], S_lo[pos])
    np.testing.assert_allclose(T[pos, 17], 1.0 / f0["q_omega"][0][pos], rtol=1e-5)
    # spatial with no neighbours reproduces the temporal reservoir's W
    np.testing.assert_array_equal(f0["spatial"][pos, 17], T[pos, 17])
    r.render(s, cam, 1)
    r.render(s, cam, 2)
    f3 = r.render(s, cam, 3)
    m3 = f3["temporal"][:, 18].view(np.uint32)
    assert set(np.unique(m3)) <= {1, 2}
    frac = (m3 == 2).mean()
    assert abs(frac - 0.6) < 4 * math.sqrt(0.24 / m3.size)
