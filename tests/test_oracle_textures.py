"""Oracle pins for the texture path (raytrace.rs:437-481 reading the atlas of
interactive_rendering.rs:416,744; utils.rs:223-235), by values the standards and the
geometry fix rather than by the oracle's own formulas:

- sRGB decode: the published IEC 61966-2-1 values (0, 1, the linear segment, 128 ->
  0.2158605, 188 -> 0.5028865) and monotonicity;
- nearest, clamp-to-edge texel selection at and around texel edges;
- a textured quad whose uv is its (x, z) position / 16: every pixel's reflectivity,
  emissivity, scatter kind and transmission come from the texel under the hit point
  (found from the hit's world position, not from the barycentrics), image by image;
- an atlas of uniform images equals the constant materials it encodes, bit for bit, on
  the 64^3 world;
- the luminance of uniform emissivity images is r + g + b.
"""
import math

import numpy as np

F32 = np.float32
VERT = np.dtype([("position", "<f4", 3), ("t", "<u4"), ("uv", "<f4", 2)])
MAT = np.dtype([("reflectivity", "<f4", 4), ("emissivity", "<f4", 4), ("metallicity", "<f4", 4)])
IDENT = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], F32)


def test_srgb_decode_known_values():
    import oracle
    assert oracle.srgb_decode(0) == 0.0
    assert oracle.srgb_decode(255) == 1.0
    # linear segment below 0.04045 (v <= 10): v / 255 / 12.92
    for v in (1, 5, 10):
        assert abs(oracle.srgb_decode(v) - v / 255.0 / 12.92) <= 1e-9
    assert abs(oracle.srgb_decode(128) - 0.2158605) < 2e-7
    assert abs(oracle.srgb_decode(188) - 0.5028865) < 2e-7
    vals = [oracle.srgb_decode(v) for v in range(256)]
    assert all(b > a for a, b in zip(vals, vals[1:]))
    # the two segments meet within a float ulp at the breakpoint
    assert abs(vals[11] - vals[10] - (vals[12] - vals[11])) < 1e-4


def test_texel_index_nearest_clamp():
    import oracle
    n = 16
    assert oracle.texel_index(0.0, n) == 0
    assert oracle.texel_index(1.0 / 16 - 1e-6, n) == 0
    assert oracle.texel_index(1.0 / 16, n) == 1
    assert oracle.texel_index(0.5, n) == 8
    assert oracle.texel_index(15.5 / 16, n) == 15
    assert oracle.texel_index(1.0, n) == 15   # clamp to edge
    assert oracle.texel_index(7.3, n) == 15
    assert oracle.texel_index(-0.01, n) == 0
    assert oracle.texel_index(float("nan"), n) == 0


def _quad_scene(atlas, slot=3):
    """A 16 x 16 floor in the y = 0 plane, normal +y, whose uv is (x, z) / 16."""
    import oracle
    m = np.zeros(6, MAT)
    m["reflectivity"] = [0.5, 0.5, 0.5, 1.0]
    lum = np.zeros(6, F32)
    p = [(0, 0, 0), (0, 0, 16), (16, 0, 0), (16, 0, 0), (0, 0, 16), (16, 0, 16)]
    v = np.zeros(6, VERT)
    v["position"] = p
    v["t"] = slot
    v["uv"] = [(x / 16.0, z / 16.0) for x, _, z in p]
    s = oracle.Scene(m, lum)
    s.add(0, v, IDENT)
    s.set_textures(atlas)
    s.build()
    return s


def _decode(b):
    c = b / 255.0
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4).astype(F32)


def test_textured_quad_texels():
    """Per pixel, the texel under the hit point decides the material, image by image:
    image 0 rgb = reflectivity (Lambert: / pi), alpha = 1 - transmit probability; image 1
    rgb = emissivity (x 1000 x -cos); image 2 r = metallicity."""
    import oracle
    import vrgi
    atlas = np.zeros((6, 3, 16, 16, 4), np.uint8)
    j, i = np.mgrid[0:16, 0:16]
    a0 = atlas[3, 0]
    a0[..., 0] = 16 * i + 3
    a0[..., 1] = 16 * j + 5
    a0[..., 2] = 200
    a0[..., 3] = 255
    a0[(i + j) % 5 == 0, 3] = 0          # transparent texels: always transmit
    atlas[3, 1, ..., 0] = 7 * i          # emissivity
    atlas[3, 1, ..., 1] = 11 * j
    atlas[3, 1, ..., 2] = 3
    atlas[3, 2, ..., 0] = np.where((i * 3 + j) % 4 == 1, 255, 0)  # metallic texels: mirror
    atlas[3, 2, ..., 1:3] = 99           # channels the reference does not read
    s = _quad_scene(atlas)
    W, H = 128, 96
    cam = vrgi.camera_look([8.0, 6.0, 4.0], [8.0, 0.0, 8.0])
    out = oracle.Renderer(W, H, 2).render(s, cam, 0)
    hit = out["hit_prim"][0] != 0xFFFFFFFF
    assert hit.sum() > 1000
    x = out["origin"][1][:, 0]
    z = out["origin"][1][:, 2]
    ti = np.floor(x).astype(int)
    tj = np.floor(z).astype(int)
    # away from texel edges, where the interpolated uv's rounding could tip floor()
    inner = hit & (np.abs(x - np.round(x)) > 1e-3) & (np.abs(z - np.round(z)) > 1e-3) & \
        (ti >= 0) & (ti < 16) & (tj >= 0) & (tj < 16)
    assert inner.sum() > 1000
    ti, tj = ti[inner], tj[inner]
    refl = out["reflectivity"][0][inner]
    pdf = out["bsdf_pdf"][0][inner]
    d0 = out["direction"][0][inner]
    d1 = out["direction"][1][inner]
    metal = atlas[3, 2, tj, ti, 0] == 255
    clear = (atlas[3, 0, tj, ti, 3] == 0) & ~metal
    lamb = ~metal & ~clear
    assert metal.sum() > 50 and clear.sum() > 50 and lamb.sum() > 50
    # mirror: the texel's reflectivity, pdf 1, d reflected about +y
    want = _decode(atlas[3, 0, tj, ti, :3].astype(np.float64))
    assert np.array_equal(refl[metal], want[metal])
    assert np.all(pdf[metal] == 1.0)
    assert np.array_equal(d1[metal], d0[metal] * np.array([1, -1, 1], F32))
    # transmit: reflectivity 1, pdf 1, d unchanged
    assert np.all(refl[clear] == 1.0) and np.all(pdf[clear] == 1.0)
    assert np.array_equal(d1[clear], d0[clear])
    # Lambert: the texel's reflectivity / pi
    assert np.array_equal(refl[lamb], want[lamb] / F32(math.pi))
    assert np.all(pdf[lamb] < 1.0)
    # emission: 1000 * texel * -cos(d, +y)
    e = _decode(atlas[3, 1, tj, ti, :3].astype(np.float64))
    got = out["emissivity"][0][inner]
    assert np.array_equal(got, (F32(1000.0) * e) * (-d0[:, 1:2]))


def test_uniform_atlas_equals_constant_materials(world0):
    """An atlas whose images are uniform renders the 64^3 world exactly as the constant
    materials those texels decode to (every per-bounce buffer, the radiance, ReSTIR)."""
    import oracle
    import vrgi
    mats, lum = world0["mats"], world0["lum"]
    n = len(mats)
    rng = np.random.default_rng(5)
    bytes_ = rng.integers(0, 256, (n, 3, 4), dtype=np.uint8)
    bytes_[:, 0, 3] = rng.choice([255, 255, 40], n)   # mostly opaque
    bytes_[:, 2, 0] = rng.choice([0, 0, 0, 128], n)   # some metal
    bytes_[:, 1, :3] = np.where(lum[:, None] > 0, bytes_[:, 1, :3], 0)  # emitters stay the world's
    atlas = np.broadcast_to(bytes_[:, :, None, None, :], (n, 3, 16, 16, 4)).copy()
    dec = lambda b: np.array([oracle.srgb_decode(int(v)) for v in b], F32)
    m2 = np.zeros(n, MAT)
    for t in range(n):
        m2["reflectivity"][t] = list(dec(bytes_[t, 0, :3])) + [F32(bytes_[t, 0, 3] / 255.0)]
        m2["emissivity"][t][:3] = dec(bytes_[t, 1, :3])
        m2["metallicity"][t][0] = dec(bytes_[t, 2, :1])[0]
    cam = vrgi.camera_look([6.0, 62.0, 8.0], [40.0, 0.0, 44.0])
    outs = []
    for textured in (False, True):
        s = oracle.Scene(m2, lum)
        s.add_objects(world0["objs"])
        if textured:
            s.set_textures(atlas)
        s.build()
        outs.append(oracle.Renderer(48, 32, 3).render(s, cam, 0, nee_type=1, spatial=2))
    for k in outs[0]:
        assert outs[0][k].tobytes() == outs[1][k].tobytes(), k
    assert (outs[0]["reflectivity"] > 0).any() and (outs[0]["bsdf_pdf"] == 1.0).any()


def test_texture_luminances_uniform():
    import oracle
    atlas = np.zeros((3, 3, 4, 5, 4), np.uint8)
    atlas[0, 1, ..., :3] = (10, 20, 30)
    atlas[1, 1, ..., :3] = (255, 255, 255)
    atlas[2, 0, ..., :3] = 200  # reflectivity does not count
    lum = oracle.texture_luminances(atlas)
    assert lum.tolist() == [60.0, 765.0, 0.0]
    atlas[1, 1, 0, 0, :3] = 0  # one black texel of 20: 765 * 19 / 20
    assert oracle.texture_luminances(atlas)[1] == np.float32(765.0 * 19) / np.float32(20)
