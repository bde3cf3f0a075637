"""Pins of the oracle's NEE pdf, ReSTIR GI and ReSTIR DI by what the mathematics fixes,
each shown to fail when a plausible mistake is planted in the restatement
(`oracle.mutation`, CPU only).

- computeNeePdf (nee_pdf.rs:240-279): with every light parallel to the shading plane
  (light cosine = the shading cosine the reference uses), the pdf is a solid-angle
  density whose integral over the hemisphere is the total pick probability, 1. The
  integral is taken by the change of variables dw = cos_l dA / d^2 over uniform points
  of each light. The barycentric order of nee_pdf.rs:230-232 (Vulkan barycentrics
  (u, v) of the hit (1-u-v) v0 + u v1 + v v2, read as weights (u, v, 1-u-v)) is pinned
  at hits of known barycentrics computed from the geometry in float64. Path-traced
  means with and without NEE agree (the mixture density q = w nee_pdf + (1-w)/(2 pi)
  of outgoing_radiance.rs:82-88 is then the true one).
- ReSTIR GI (restir_temporal_resampling.rs, restir_spatial_resampling.rs,
  restir_finalize.rs) on a Lambert floor under the sky: every sample is a sky
  direction (n_s = 0, Jacobian 1) and every visible point has the same normal, so the
  resampled estimate is unbiased once every reservoir holds a non-zero sample, and its
  mean is the closed form (100 a / pi) (1 - 0.9^2) / 2 with 0 and with 5 spatial
  neighbours. In the first frame every reservoir has M = 1 and W = 1/p_w or 0, and the
  reference's Z (restir_spatial_resampling.rs:391-398: M of the reservoirs whose own
  sample has p_hat > 0) makes every merged W exactly 1/p_w. The Jacobian
  (:299-311) has hand-derived known answers.
- ReSTIR DI (the benchmark mode): over emissive triangles above a Lambert floor the
  frame-averaged DI radiance equals (a / pi) Le * integral of cos_x cos_y / d^2 dA,
  evaluated by Lambert's closed form for polygons in float64.
"""
import math

import numpy as np
import pytest

F32 = np.float32
VERT = np.dtype([("position", "<f4", 3), ("t", "<u4"), ("uv", "<f4", 2)])
MAT = np.dtype([("reflectivity", "<f4", 4), ("emissivity", "<f4", 4), ("metallicity", "<f4", 4)])
IDENT = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], F32)
FLOOR_T = 3


def materials(albedo, emissions):
    """Slot FLOOR_T: Lambert floor of grey albedo; slots 6k + 2: emitters (black)."""
    n = 6 * (len(emissions) + 1)
    m = np.zeros(n, MAT)
    m["reflectivity"][:, 3] = 1.0
    m["emissivity"][:, 3] = 1.0
    m["metallicity"][:, 3] = 1.0
    m["reflectivity"][FLOOR_T, :3] = albedo
    lum = np.zeros(n, F32)
    for k, e in enumerate(emissions):
        t = 6 * (k + 1) + 2
        m["emissivity"][t, :3] = e
        lum[t] = 255.0 * float(np.sum(np.asarray(e, F32)))
    return m, lum


def light_slot(k):
    return 6 * (k + 1) + 2


def floor_verts(size=4000.0):
    s = size / 2
    v = np.zeros(6, VERT)
    v["position"] = [(-s, 0, -s), (-s, 0, s), (s, 0, -s), (s, 0, -s), (-s, 0, s), (s, 0, s)]  # normal +y
    v["t"] = FLOOR_T
    return v


def down_triangle(c, a, b, t):
    """Horizontal triangle facing -y with corners c, c + a, c + b (a, b horizontal)."""
    c, a, b = (np.asarray(x, np.float64) for x in (c, a, b))
    p = np.array([c, c + a, c + b])
    if np.cross(p[1] - p[0], p[2] - p[0])[1] > 0:
        p = p[[0, 2, 1]]
    v = np.zeros(3, VERT)
    v["position"] = p.astype(F32)
    v["t"] = t
    return v


def scene_of(mats, lum, objs):
    import oracle
    s = oracle.Scene(mats, lum)
    for key, v in enumerate(objs):
        s.add(key, v, IDENT)
    s.build()
    return s


def camera(eye, target):
    import vrgi
    return vrgi.camera_look(eye, target)


def tri_points(tri):
    return tri["position"].astype(np.float64)


# ------------------------------------------------------------------ computeNeePdf
# lights far compared with their size (side / height <= 0.25), off the normal (light
# cosines 0.6-0.8), projections disjoint as seen from the shading point
FAR_LIGHTS = [((3.0, 5.0, 1.0), (1.2, 0, 0), (0, 0, 1.2), (1.0, 0.9, 0.8)),
              ((-4.0, 6.0, -2.0), (1.5, 0, 0.3), (0.2, 0, 1.4), (0.3, 0.6, 0.9)),
              ((1.0, 4.5, -4.5), (0.9, 0, 0), (0, 0, 1.1), (0.9, 0.2, 0.2))]


def far_light_scene(albedo=0.6, lights=FAR_LIGHTS, floor=True):
    mats, lum = materials(albedo, [l[3] for l in lights])
    tris = [down_triangle(c, a, b, light_slot(k)) for k, (c, a, b, _) in enumerate(lights)]
    objs = ([floor_verts()] if floor else []) + tris
    return scene_of(mats, lum, objs), tris


def nee_integral(s, tris, p, n, samples=4000, seed=0):
    """Sum over lights of  E_y[ nee_pdf(p, n, w(y)) * A cos_l / d^2 * d^2 / d_Q^2 ],  y
    uniform on the light (kept 2 % away from its edges so that every ray crosses it): the
    integral of the pdf over the hemisphere when no two lights overlap as seen from p.
    d_Q = |Q(y) - p| is the distance the reference measures, to the point Q whose weights
    are y's Vulkan barycentrics (u, v) taken as (u, v, 1-u-v) (nee_pdf.rs:230-232): the
    factor d^2 / d_Q^2 undoes exactly that, so the sum is the total pick probability, 1."""
    rng = np.random.default_rng(seed)
    total = 0.0
    for tri in tris:
        v = tri_points(tri)
        u1, u2 = rng.random(samples), rng.random(samples)
        m = u1 + u2 > 1
        u1[m], u2[m] = 1 - u1[m], 1 - u2[m]
        b = np.stack([1 - u1 - u2, u1, u2], 1) * 0.96 + 0.04 / 3
        y = b @ v
        q = b[:, [1, 2, 0]] @ v
        dq2 = ((q - p) ** 2).sum(1)
        w = y - p
        d2 = (w * w).sum(1)
        w /= np.sqrt(d2)[:, None]
        nl = np.cross(v[1] - v[0], v[2] - v[0])
        area = 0.5 * np.linalg.norm(nl)
        nl /= np.linalg.norm(nl)
        cos_l = np.abs(w @ nl)
        pdf = s.nee_pdf(np.tile(p, (samples, 1)), n, w.astype(F32)).astype(np.float64)
        assert np.all(pdf > 0)
        total += np.mean(pdf * area * cos_l / dq2)
    return total


def _nee_integral_ok():
    s, tris = far_light_scene()
    p = np.array([0.0, 0.0015, 0.0])
    n = np.array([0.0, 1.0, 0.0], F32)
    got = nee_integral(s, tris, p, n)
    return abs(got - 1.0) < 2e-4, got


def test_nee_pdf_integrates_to_one():
    ok, got = _nee_integral_ok()
    assert ok, got


@pytest.mark.parametrize("mut", ["nee_area", "nee_dist", "nee_cos", "nee_bary"])
def test_nee_pdf_integral_catches(mut):
    import oracle
    with oracle.mutation(mut):
        ok, got = _nee_integral_ok()
    assert not ok, (mut, got)


def _nee_bary_ok():
    """One light close to the point (side = height): the pdf at hits of chosen Vulkan
    barycentrics (u, v) is |Q - p|^2 / (cos A) at Q = u v0 + v v1 + (1-u-v) v2 (the pick
    probability is 1 with one light)."""
    lights = [((-0.3, 1.0, -0.2), (1.0, 0, 0), (0, 0, 1.0), (1.0, 1.0, 1.0))]
    s, tris = far_light_scene(lights=lights)
    v = tri_points(tris[0])
    p = np.array([0.0, 0.0015, 0.0])
    n = np.array([0.0, 1.0, 0.0])
    area = 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))
    errs = []
    for u, w in [(0.1, 0.2), (0.6, 0.3), (0.25, 0.05), (0.05, 0.7)]:
        hit = (1 - u - w) * v[0] + u * v[1] + w * v[2]
        d = (hit - p) / np.linalg.norm(hit - p)
        q = u * v[0] + w * v[1] + (1 - u - w) * v[2]
        want = np.sum((q - p) ** 2) / (float(d @ n) * area)
        got = float(s.nee_pdf(p.astype(F32)[None], n.astype(F32), d.astype(F32)[None])[0])
        errs.append(abs(got - want) / want)
    return max(errs) < 1e-4, errs


def test_nee_pdf_barycentric_order():
    ok, errs = _nee_bary_ok()
    assert ok, errs


@pytest.mark.parametrize("mut", ["nee_bary", "nee_area", "nee_dist"])
def test_nee_pdf_barycentric_order_catches(mut):
    import oracle
    with oracle.mutation(mut):
        ok, errs = _nee_bary_ok()
    assert not ok, (mut, errs)


def _nee_on_off_ok(W=256, H=256, frames=3):
    """Path-traced radiance (2 bounces) with NEE on every bounce vs without: equal means
    within 4 standard errors (independent pixel streams and frames)."""
    import oracle
    s, _ = far_light_scene(albedo=0.7)
    cam = camera([0.3, 2.0, 0.2], [0.31, 0.0, 0.21])
    r = oracle.Renderer(W, H, 2)
    means, vars_ = [], []
    for nee in (0, 1):
        xs = []
        for f in range(frames):
            L = r.render(s, cam, 7 + f, nee_type=nee)["l_out"][0].astype(np.float64)
            xs.append(L @ [0.2126, 0.7152, 0.0722])
        x = np.concatenate(xs)
        means.append(x.mean())
        vars_.append(x.var() / x.size)
    z = abs(means[1] - means[0]) / math.sqrt(vars_[0] + vars_[1])
    return z < 4.0, (means, z)


def test_nee_on_off_agree():
    ok, info = _nee_on_off_ok()
    assert ok, info


@pytest.mark.parametrize("mut", ["nee_area", "nee_cos"])
def test_nee_on_off_catches(mut):
    import oracle
    with oracle.mutation(mut):
        ok, info = _nee_on_off_ok()
    assert not ok, (mut, info)


# ------------------------------------------------------------------ ReSTIR GI
SKY_ALBEDO = 0.6
SKY_MEAN = 100.0 * SKY_ALBEDO / math.pi * (1 - 0.81) / 2


def sky_floor():
    mats, lum = materials(SKY_ALBEDO, [])
    return scene_of(mats, lum, [floor_verts()])


def _gi_run(W, H, spatial, frames, warmup, seed=0):
    """Frame-averaged ReSTIR GI output on the floor under the sky; returns the image
    mean and its standard error from the means of 16 x 16 blocks (spatial reuse reaches
    7.5 px, so blocks are close to independent)."""
    import oracle
    s = sky_floor()
    cam = camera([0, 10, 0], [0.01, 0.0, 0.02])
    r = oracle.Renderer(W, H, 2, seed=seed)
    acc = np.zeros(W * H)
    for f in range(warmup + frames):
        out = r.render(s, cam, f, spatial=spatial)
        if f >= warmup:
            acc += out["restir"].astype(np.float64) @ [0.2126, 0.7152, 0.0722]
    img = (acc / frames).reshape(H, W)
    blocks = img.reshape(H // 16, 16, W // 16, 16).mean(axis=(1, 3)).ravel()
    return img.mean(), blocks.std(ddof=1) / math.sqrt(blocks.size)


def _gi_temporal_ok():
    m, se = _gi_run(128, 96, 0, 48, 0)
    return abs(m - SKY_MEAN) < 4 * se, (m, se, SKY_MEAN)


def _gi_spatial_ok():
    m, se = _gi_run(128, 96, 5, 40, 150)
    return abs(m - SKY_MEAN) < 4 * se, (m, se, SKY_MEAN)


def test_restir_gi_temporal_converges_to_closed_form():
    ok, info = _gi_temporal_ok()
    assert ok, info


def test_restir_gi_spatial_converges_to_closed_form():
    ok, info = _gi_spatial_ok()
    assert ok, info


@pytest.mark.parametrize("mut", ["gi_final_cos"])
def test_restir_gi_temporal_catches(mut):
    import oracle
    with oracle.mutation(mut):
        ok, info = _gi_temporal_ok()
    assert not ok, (mut, info)


@pytest.mark.parametrize("mut", ["gi_merge_no_m", "gi_final_cos", "gi_ucw_no_m"])
def test_restir_gi_spatial_catches(mut):
    import oracle
    with oracle.mutation(mut):
        ok, info = _gi_spatial_ok()
    assert not ok, (mut, info)


def _gi_first_frame_ok():
    """Frame 0 (zero-filled history): a sky sample gives M = 1, W = 1/p_w; a zero-radiance
    candidate never replaces the history's zero sample (update with w = 0), whose n_v = 0
    fails geometricallySimilar (restir_spatial_resampling.rs:277-288), so such a pixel
    merges nothing (M = 1, W = 0) and is merged by nobody. A pixel holding a sky sample
    merges exactly its neighbours that hold one, and its merged W is again 1/p_w (all
    p_hat equal, Z = the number of merged reservoirs)."""
    import oracle
    s = sky_floor()
    cam = camera([0, 10, 0], [0.01, 0.0, 0.02])
    r = oracle.Renderer(64, 48, 2)
    out = r.render(s, cam, 0, spatial=5)
    inv_pw = 1.0 / out["q_omega"][0].astype(np.float64)
    Wt = out["temporal"][:, 17].astype(np.float64)
    W = out["spatial"][:, 17].astype(np.float64)
    M = out["spatial"][:, 18].view(np.uint32)
    own = Wt != 0
    ok = (0.05 < own.mean() < 0.2 and np.all(np.abs(Wt[own] / inv_pw[own] - 1) < 1e-6)
          and np.all(W[~own] == 0) and np.all(M[~own] == 1)
          and np.all(np.abs(W[own] / inv_pw[own] - 1) < 1e-5) and np.all((M[own] >= 1) & (M[own] <= 6))
          and (M[own] > 1).any())
    return ok, (own.mean(), np.bincount(M[own]))


def test_restir_gi_first_frame_weights():
    ok, info = _gi_first_frame_ok()
    assert ok, info


def _gi_back_facing_ok():
    """Almost every bounce-1 ray hits the back of an up-facing emitter ceiling, so its sample
    has negative radiance (1000 e (-d . n) < 0, raytrace.rs:477) and p_hat < 0. Frame 0:
    each reservoir takes its sample (w / w_sum = 1), all 6 are merged, and the reference's
    Z counts only reservoirs with p_hat > 0 (restir_spatial_resampling.rs:391-398), so
    Z = 0 and W = w_sum / (0 * p_hat) = +inf (w_sum < 0, 0 * p_hat = -0) for every pixel."""
    import oracle
    mats, lum = materials(SKY_ALBEDO, [(0.5, 0.5, 0.5)])
    ceil = floor_verts(size=400.0)
    ceil["position"][:, 1] = 3.0
    ceil["t"] = light_slot(0)
    s = scene_of(mats, lum, [floor_verts(), ceil])
    cam = camera([0, 1.5, 0], [0.01, 0.0, 0.02])
    r = oracle.Renderer(32, 24, 2)
    out = r.render(s, cam, 0, spatial=5)
    lo = out["temporal"][:, 12:15].astype(np.float64) @ [0.2126, 0.7152, 0.0722]
    W = out["spatial"][:, 17]
    M = out["spatial"][:, 18].view(np.uint32)
    neg = lo < 0  # the rest (grazing rays past the ceiling's edge) hold the zero sample
    ok = neg.mean() > 0.9 and np.all(np.isposinf(W[neg])) and np.all(W[~neg] == 0) and (M[neg] == 6).mean() > 0.5
    return bool(ok), (neg.mean(), np.unique(M), W[:4])


def test_restir_gi_z_filter_back_facing():
    ok, info = _gi_back_facing_ok()
    assert ok, info


@pytest.mark.parametrize("mut", ["gi_z_all"])
def test_restir_gi_z_filter_catches(mut):
    import oracle
    with oracle.mutation(mut):
        ok, info = _gi_back_facing_ok()
    assert not ok, (mut, info)


def reservoir(xv, nv, xs, ns, lo=(1, 1, 1), pw=1.0, ucw=1.0, m=1):
    r = np.zeros(20, F32)
    r[0:3], r[3:6], r[6:9], r[9:12], r[12:15] = xv, nv, xs, ns, lo
    r[15] = pw
    r[17] = ucw
    r[18] = np.array([m], np.uint32).view(F32)[0]
    return r


# hand-derived from restir_spatial_resampling.rs:299-311, (cos_r / cos_q) * (|q_v - x_s| / |r_v - x_s|):
#  A: x_s = (0,2,0), n_s = -y, q_v = 0 (cos 1, distance 2), r_v = (2,0,0) (cos 1/sqrt 2,
#     distance sqrt 8): J = (1/sqrt 2) * (2 / sqrt 8) = 1/2
#  B: x_s = (0,1,0), n_s = -y, q_v = (1,0,0) (cos 1/sqrt 2, distance sqrt 2), r_v = 0 (cos 1,
#     distance 1): J = sqrt 2 * sqrt 2 = 2
#  sky (n_s = 0): 1
JAC_CASES = [((0, 0, 0), (2, 0, 0), (0, 2, 0), (0, -1, 0), 0.5),
             ((1, 0, 0), (0, 0, 0), (0, 1, 0), (0, -1, 0), 2.0),
             ((1, 0, 0), (0, 0, 5), (3, 7, 1), (0, 0, 0), 1.0)]


def _jacobian_ok():
    import oracle
    got = []
    for qv, rv, xs, ns, want in JAC_CASES:
        q = reservoir(qv, (0, 1, 0), xs, ns)
        r = reservoir(rv, (0, 1, 0), (9, 9, 9), (0, 1, 0))
        got.append((oracle.gi_jacobian(q, r), want))
    return all(abs(g - w) <= 1e-6 * w for g, w in got), got


def test_restir_gi_jacobian_known_answers():
    ok, got = _jacobian_ok()
    assert ok, got


@pytest.mark.parametrize("mut", ["gi_jac_sq", "gi_jac_swap"])
def test_restir_gi_jacobian_catches(mut):
    import oracle
    with oracle.mutation(mut):
        ok, got = _jacobian_ok()
    assert not ok, (mut, got)


# ------------------------------------------------------------------ ReSTIR DI
def lambert_polygon(x, n, poly):
    """integral over the polygon of cos_x cos_y / d^2 dA seen from points x (rows) with
    normals n: (1/2) |sum over edges of beta_i (n . gamma_i)| (Lambert's formula; the
    polygon wholly above every tangent plane)."""
    x = np.asarray(x, np.float64)
    R = [np.asarray(v, np.float64)[None, :] - x for v in poly]
    tot = np.zeros(len(x))
    for i in range(len(R)):
        a, b = R[i], R[(i + 1) % len(R)]
        c = np.cross(a, b)
        cn = np.linalg.norm(c, axis=1)
        beta = np.arctan2(cn, (a * b).sum(1))
        tot += beta * (c * n).sum(1) / cn
    return 0.5 * np.abs(tot)


def test_lambert_polygon_small_limit():
    """The helper itself: a small square of area A at height h straight above: A / h^2."""
    h, e = 10.0, 0.001
    sq = [(-e, h, -e), (e, h, -e), (e, h, e), (-e, h, e)]
    got = lambert_polygon(np.zeros((1, 3)), np.array([0.0, 1.0, 0.0]), sq)[0]
    assert abs(got - (2 * e) ** 2 / h ** 2) < 1e-6 * got
    # a disc-like limit: the whole hemisphere covered (a huge square just above) gives pi
    big = [(-1e6, 1e-3, -1e6), (1e6, 1e-3, -1e6), (1e6, 1e-3, 1e6), (-1e6, 1e-3, 1e6)]
    assert abs(lambert_polygon(np.zeros((1, 3)), np.array([0.0, 1.0, 0.0]), big)[0] - math.pi) < 1e-6


# coplanar (no light shadows another), off-axis emitters of areas != 1 and distinct colours
DI_LIGHTS = [((1.0, 3.0, 0.5), (2.0, 0, 0), (0, 0, 2.0), (0.9, 0.8, 0.7)),
             ((-3.0, 3.0, -1.0), (1.0, 0, 0.2), (0.1, 0, 0.8), (0.2, 0.5, 0.9)),
             ((0.5, 3.0, -3.5), (1.5, 0, 0), (0.4, 0, 1.2), (0.9, 0.3, 0.1))]
DI_ALBEDO = 0.5


def _di_run(spatial, runs=8, frames=5, W=64, H=48):
    """Luminance ratio of the frame-averaged DI radiance to the quadrature for `runs`
    independent renderers (disjoint frame numbers, so disjoint RNG streams, and each
    starting from an empty history); the history and the spatial reuse correlate the
    frames and pixels of one run, not different runs. Returns the run ratios and the
    per-channel ratio of all runs."""
    import oracle
    mats, lum = materials(DI_ALBEDO, [l[3] for l in DI_LIGHTS])
    tris = [down_triangle(c, a, b, light_slot(k)) for k, (c, a, b, _) in enumerate(DI_LIGHTS)]
    s = scene_of(mats, lum, [floor_verts()] + tris)
    cam = camera([0.0, 2.5, 0.0], [0.4, 0.0, 0.3])
    lw = np.array([0.2126, 0.7152, 0.0722])
    want, ratios, tot = None, [], np.zeros(3)
    for k in range(runs):
        r = oracle.Renderer(W, H, 2, mode=1, radius=30.0, di_candidates=8, seed=k)
        acc = np.zeros((W * H, 3))
        for f in range(frames):
            out = r.render(s, cam, 1000 * k + f, nee_type=1, spatial=spatial)
            acc += out["restir"].astype(np.float64) - out["emissivity"][0].astype(np.float64)
        if want is None:
            assert np.all(out["hit_prim"][0] != 0xFFFFFFFF)
            x = out["origin"][1].astype(np.float64)
            want = np.zeros((W * H, 3))
            for tri, (_, _, _, e) in zip(tris, DI_LIGHTS):
                I = lambert_polygon(x, np.array([0.0, 1.0, 0.0]), tri_points(tri))
                want += (DI_ALBEDO / math.pi) * I[:, None] * (1000.0 * np.asarray(e, np.float64))[None, :]
        got = acc / frames
        ratios.append((got @ lw).sum() / (want @ lw).sum())
        tot += got.sum(0)
    return np.array(ratios), tot / (runs * want.sum(0))


def _di_ok(spatial):
    ratios, chan = _di_run(spatial)
    m, se = ratios.mean(), ratios.std(ddof=1) / math.sqrt(len(ratios))
    ok = abs(m - 1) < max(4 * se, 2e-3) and abs(m - 1) < 0.02 and np.all(np.abs(chan - 1) < 0.03)
    return ok, (m, se, chan)


@pytest.mark.parametrize("spatial", [0, 5])
def test_restir_di_matches_quadrature(spatial):
    ok, info = _di_ok(spatial)
    assert ok, info


@pytest.mark.parametrize("mut,spatial", [("di_area", 0), ("di_cosy", 0), ("di_ucw_no_m", 5),
                                         ("di_spatial_m", 5)])
def test_restir_di_catches(mut, spatial):
    import oracle
    with oracle.mutation(mut):
        ok, info = _di_ok(spatial)
    assert not ok, (mut, info)
