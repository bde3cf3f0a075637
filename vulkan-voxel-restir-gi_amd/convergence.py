"""Time-to-equal-RMSE of the learned radiance cache against the plain path tracer.

References. The plain unbiased path tracer (`bounces` bounces, Russian roulette from
`rr_start`, NEE) averaged in FP64 on the GPU over frames whose seeds are disjoint from
every measured run (frames from 2^24 on). Each of its samples is the oracle's, bit for
bit (the frame parity tests). A reference is taken as two halves A and B of spp / 2
frames each: (A - B) / 2 estimates its own error, so its noise variance is
var_ref = mean((A - B)^2) / 4 (per pixel channel), reported as `reference_rmse`. Every
RMSE is reported raw and with that noise removed, sqrt(max(RMSE^2 - var_ref, 0)) (the
reference's noise is independent of the measured image's); the ratios use the latter.

Moving camera. The cache trains online from its first frame while the camera follows the
workload's spline (`cams`, looped for `warmup` frames). At the poses `probe_poses` of the
last loop, its single frame is compared with that pose's reference, beside the plain
tracer's single frame at the same pose; `plain_spp_to_match` is the sample count at which
the plain tracer's RMSE, falling as 1/sqrt(spp), reaches the cache's.

Static camera (pose 0). The plain tracer's running FP64 mean over `frames` frames; the
cache, as it left the moving phase, the same. For a plain checkpoint k, the cache's time
to equal RMSE is the time of the first cache frame whose running-mean RMSE is at most the
plain RMSE after k frames; `time_to_equal_rmse_ratio` divides it by the plain time of k
frames, `..._with_warmup` adds the whole moving phase (every frame the cache trained on)
to the cache's time. Each frame's GPU time is the HIP-event time of the whole frame (BVH
upkeep, trace, NEE pdf, cache inference and training steps, ReSTIR)."""
import numpy as np

import vrgi

LUM = np.array([0.2126, 0.7152, 0.0722])


def _mse(a, b):
    return float(np.mean((a - b) ** 2))


def _frame_ms(ctx):
    return ctx.profile_read(reset=True).frame_ms


def _reference(ctx, cam, spp, frame0, nee, spatial):
    """Mean of spp frames from frame number frame0 on, as two halves; (mean, noise variance)."""
    halves = []
    for h in range(2):
        ctx.frame_set(frame0 + h * (spp // 2))
        ctx.accum_enable(True)
        for k in range(spp // 2):
            ctx.render(cam, nee_type=nee, spatial=spatial)
            if k % 64 == 63:
                ctx.sync()
        ctx.sync()
        m, n = ctx.accum_read()
        assert n == spp // 2
        halves.append(m)
    a, b = halves
    return 0.5 * (a + b), float(np.mean((a - b) ** 2)) / 4.0


def _rm(mse, var):
    return float(np.sqrt(mse)), float(np.sqrt(max(mse - var, 0.0)))


def study(objs, mats, lum, cams, W, H, bounces=8, rr_start=2, ref_spp=1024, frames=32, warmup=32, nee=1,
          spatial=5, probe_poses=(0, 4, 8, 12), probe_ref_spp=256, cache_kw=None, log=print):
    cams = list(cams)
    cfg = vrgi.config(W, H, bounces, rr_start=rr_start, rr_min=0.1)
    out = {"resolution": f"{W}x{H}", "bounces": bounces, "rr_start": rr_start, "reference_spp": ref_spp,
           "frames": frames, "moving_warmup_frames": warmup, "spline_poses": len(cams),
           "probe_poses": list(probe_poses), "probe_reference_spp": probe_ref_spp}
    probe_poses = [p for p in probe_poses if p < len(cams)]
    plain_ms, plain_mse = [], []
    probe_refs, probe_plain = {}, {}
    with vrgi.Context(cfg, mats, lum) as ctx:
        ctx.add_objects(objs)
        ctx.render(cams[0], nee_type=nee, spatial=spatial)  # builds the scene
        ctx.sync()
        log(f"reference at pose 0: {ref_spp} spp")
        ref, ref_var = _reference(ctx, cams[0], ref_spp, 1 << 24, nee, spatial)
        for i, p in enumerate(probe_poses):
            log(f"reference at spline pose {p}: {probe_ref_spp} spp")
            probe_refs[p] = _reference(ctx, cams[p], probe_ref_spp, (1 << 24) + (i + 1) * (1 << 20), nee, spatial)
        # the plain tracer's single frame at each probe pose
        ctx.accum_enable(False)
        for i, p in enumerate(probe_poses):
            ctx.frame_set(1000 + i)
            ctx.render(cams[p], nee_type=nee, spatial=spatial)
            ctx.sync()
            img = ctx.read(vrgi.BUF_RADIANCE).reshape(-1, 3).astype(np.float64)
            probe_plain[p] = _mse(np.where(np.isfinite(img), img, 0.0), probe_refs[p][0])
        # plain: running mean after every frame at pose 0
        ctx.profile(vrgi.PROFILE_STAGES)
        _frame_ms(ctx)
        ctx.frame_set(0)
        ctx.accum_enable(True)
        log(f"plain: {frames} frames")
        for k in range(frames):
            ctx.render(cams[0], nee_type=nee, spatial=spatial)
            ctx.sync()
            plain_ms.append(_frame_ms(ctx))
            m, _ = ctx.accum_read()
            plain_mse.append(_mse(m, ref))
        ctx.profile(0)
    # cache: trained along the moving camera, then measured at pose 0
    cache_ms, cache_mse, probe_cache = [], [], {}
    with vrgi.Context(cfg, mats, lum) as ctx:
        ctx.add_objects(objs)
        ctx.render(cams[0], nee_type=nee, spatial=spatial)
        ctx.sync()
        kw = dict(train_bounces=bounces)
        kw.update(cache_kw or {})
        ctx.cache_enable(**kw)
        ctx.profile(vrgi.PROFILE_STAGES)
        _frame_ms(ctx)
        ctx.frame_set(0)
        log(f"cache: {warmup} frames along the spline")
        warm_ms = 0.0
        last_loop = warmup - len(cams)
        for k in range(warmup):
            p = k % len(cams)
            ctx.render(cams[p], nee_type=nee, spatial=spatial)
            ctx.sync()
            warm_ms += _frame_ms(ctx)
            if k >= last_loop and p in probe_refs:
                img = ctx.read(vrgi.BUF_RADIANCE).reshape(-1, 3).astype(np.float64)
                probe_cache[p] = _mse(np.where(np.isfinite(img), img, 0.0), probe_refs[p][0])
        ctx.accum_enable(True)
        log(f"cache: {frames} frames at pose 0")
        for k in range(frames):
            ctx.render(cams[0], nee_type=nee, spatial=spatial)
            ctx.sync()
            cache_ms.append(_frame_ms(ctx))
            m, _ = ctx.accum_read()
            cache_mse.append(_mse(m, ref))
        info = ctx.cache_info()
        ctx.profile(0)
    plain_r = [_rm(x, ref_var) for x in plain_mse]
    cache_r = [_rm(x, ref_var) for x in cache_mse]
    t_plain = np.cumsum(plain_ms)
    t_cache = np.cumsum(cache_ms)
    ratios, ratios_w = {}, {}
    for k in (1, 4, 16, 64):
        if k > frames:
            continue
        target = plain_r[k - 1][1]
        hit = [j for j in range(frames) if cache_r[j][1] <= target]
        ratios[str(k)] = round(float(t_cache[hit[0]] / t_plain[k - 1]), 4) if hit else None
        ratios_w[str(k)] = round(float((warm_ms + t_cache[hit[0]]) / t_plain[k - 1]), 4) if hit else None
    probes = []
    plain_frame = float(np.mean(plain_ms))
    for p in probe_poses:
        if p not in probe_cache:
            continue
        var = probe_refs[p][1]
        pr, cr = _rm(probe_plain[p], var)[1], _rm(probe_cache[p], var)[1]
        spp = (pr / cr) ** 2 if cr > 0 else None
        probes.append({"pose": p, "plain_rmse": round(pr, 5), "cache_rmse": round(cr, 5),
                       "reference_rmse": round(float(np.sqrt(var)), 5),
                       "plain_spp_to_match": round(spp, 2) if spp else None})
    keys = [k for k in (1, 2, 4, 8, 16, 32, 64) if k <= frames]
    cache_frame = float(np.mean(cache_ms))
    out.update({
        "plain_ms_per_frame": round(plain_frame, 3),
        "cache_ms_per_frame": round(cache_frame, 3),
        "moving_warmup_ms": round(warm_ms, 2),
        "reference_rmse": round(float(np.sqrt(ref_var)), 5),
        "plain_rmse": {str(k): round(plain_r[k - 1][1], 5) for k in keys},
        "cache_rmse": {str(k): round(cache_r[k - 1][1], 5) for k in keys},
        "plain_rmse_raw": {str(k): round(plain_r[k - 1][0], 5) for k in keys},
        "cache_rmse_raw": {str(k): round(cache_r[k - 1][0], 5) for k in keys},
        "time_to_equal_rmse_ratio": ratios,
        "time_to_equal_rmse_ratio_with_warmup": ratios_w,
        "moving_camera_probes": probes,
        "moving_camera_time_ratio": (round(float(np.mean([cache_frame / (q["plain_spp_to_match"] * plain_frame)
                                                          for q in probes if q["plain_spp_to_match"]])), 4)
                                     if probes else None),
        "cache_train_steps": int(info.steps), "cache_last_loss": round(float(info.last_loss), 5),
        "reference_mean_luminance": round(float(np.mean(ref @ LUM)), 5),
    })
    return out


def dynamic_study(blocks, n_types, translucent, mats, lum, cam, W, H, edits, bounces=4, rr_start=2, mode=1,
                  radius=30.0, di_candidates=32, nee=1, spatial=5, probe_frames=(7, 15, 23, 31), probe_ref_spp=64,
                  ref_spp=256, static_frames=16, cache_kw=None, log=print):
    """Time to equal RMSE on a world edited every frame (the large dynamic workload).

    Two device voxel worlds, plain and with the learned cache, take the same voxel writes
    (`edits`, one (xyz, values) pair per frame) and render one frame after each, with a
    static camera; every frame's GPU time (voxel upload, re-mesh of the dirty chunks, BVH
    rebuilds, the frame, the cache's inference and training) comes from HIP events. A
    third world, the reference, follows the edits only at the probe frames: there the
    plain tracer (same estimator as the plain world) accumulates `probe_ref_spp` frames in
    FP64 in two halves on seeds disjoint from the measured frames, its own noise is
    measured and removed as in study(), and each world's single frame is compared with it.
    `plain_spp_to_match` is the sample count at which the plain tracer's error, falling as
    1/sqrt(spp) on that world, reaches the cache's; the dynamic time ratio is the cache's
    frame time over plain_spp_to_match plain frames.

    After the last edit both worlds accumulate `static_frames` frames of the final world
    (running FP64 means) against a `ref_spp` reference, and the time to equal RMSE is
    measured as in study(): the cache's first frame whose running-mean error reaches the
    plain tracer's error after k frames, divided by the plain time of k frames. Both
    worlds rendered the same edited frames before, so neither side counts them."""
    frames = len(edits)
    kw = dict(rr_start=rr_start, rr_min=0.1, mode=mode, radius=radius, di_candidates=di_candidates)
    cfg = vrgi.config(W, H, bounces, **kw)
    pr = [f for f in probe_frames if f < frames]
    out = {"resolution": f"{W}x{H}", "bounces": bounces, "frames": frames, "probe_frames": pr,
           "probe_reference_spp": probe_ref_spp, "reference_spp": ref_spp, "static_frames": static_frames}
    ctxs = [vrgi.Context(cfg, mats, lum) for _ in range(3)]
    try:
        plain, cache, refc = ctxs
        for c in ctxs:
            c.voxels_load(blocks, n_types, translucent)
        cache.render(cam, nee_type=nee, spatial=spatial)  # the cache box comes from the scene
        cache.sync()
        ckw = dict(train_bounces=bounces, query_vertex=0)
        ckw.update(cache_kw or {})
        cache.cache_enable(**ckw)
        info = cache.cache_info()
        out["cache_precision"] = "fp16" if info.precision == vrgi.CACHE_FP16 else "bf16"
        for c in (plain, cache):
            c.frame_set(0)
            c.profile(vrgi.PROFILE_STAGES)
            _frame_ms(c)
        ms = {"plain": [], "cache": []}
        remeshed, tris, probes, pending = [], [], [], []
        _, r0 = plain.voxels_counts()
        for f, (xyz, vals) in enumerate(edits):
            pending.append((xyz, vals))
            for name, c in (("plain", plain), ("cache", cache)):
                c.voxels_set(xyz, vals)
                c.render(cam, nee_type=nee, spatial=spatial)
                c.sync()
                ms[name].append(_frame_ms(c))
            _, r1 = plain.voxels_counts()
            remeshed.append(r1 - r0)
            r0 = r1
            tris.append(int(plain.stats().bvh_nodes + 1) // 2)
            if f in pr:
                imgs = {n: c.read(vrgi.BUF_RADIANCE).reshape(-1, 3).astype(np.float64) for n, c in
                        (("plain", plain), ("cache", cache))}
                for xyz_p, vals_p in pending:
                    refc.voxels_set(xyz_p, vals_p)
                pending = []
                log(f"dynamic study: reference at frame {f}, {probe_ref_spp} spp")
                ref, var = _reference(refc, cam, probe_ref_spp, (1 << 24) + f * (1 << 16), nee, spatial)
                e = {n: _rm(_mse(np.where(np.isfinite(v), v, 0.0), ref), var)[1] for n, v in imgs.items()}
                spp = (e["plain"] / e["cache"]) ** 2 if e["cache"] > 0 else None
                probes.append({"frame": f, "plain_rmse": round(e["plain"], 5), "cache_rmse": round(e["cache"], 5),
                               "reference_rmse": round(float(np.sqrt(var)), 5),
                               "plain_spp_to_match": round(spp, 2) if spp else None})
        for xyz_p, vals_p in pending:
            refc.voxels_set(xyz_p, vals_p)
        log(f"dynamic study: final-world reference, {ref_spp} spp")
        ref, var = _reference(refc, cam, ref_spp, 1 << 25, nee, spatial)
        run = {}
        for name, c in (("plain", plain), ("cache", cache)):
            c.accum_enable(True)
            t, e = [], []
            for k in range(static_frames):
                c.render(cam, nee_type=nee, spatial=spatial)
                c.sync()
                t.append(_frame_ms(c))
                m, _ = c.accum_read()
                e.append(_rm(_mse(m, ref), var)[1])
            run[name] = (np.cumsum(t), e)
            c.profile(0)
        info = cache.cache_info()
    finally:
        for c in ctxs:
            c.close()
    (tp, ep), (tc, ec) = run["plain"], run["cache"]
    ratios = {}
    for k in (1, 4, 16):
        if k > static_frames:
            continue
        hit = [j for j in range(static_frames) if ec[j] <= ep[k - 1]]
        ratios[str(k)] = round(float(tc[hit[0]] / tp[k - 1]), 4) if hit else None
    pm, cm = float(np.mean(ms["plain"])), float(np.mean(ms["cache"]))
    keys = [k for k in (1, 2, 4, 8, 16) if k <= static_frames]
    out.update({
        "plain_ms_per_frame": round(pm, 3), "cache_ms_per_frame": round(cm, 3),
        "plain_ms_per_frame_each": [round(v, 2) for v in ms["plain"]],
        "cache_ms_per_frame_each": [round(v, 2) for v in ms["cache"]],
        "chunks_remeshed_each": remeshed,
        "triangles_each": tris,
        "probes": probes,
        "dynamic_time_ratio": (round(float(np.mean([cm / (q["plain_spp_to_match"] * pm) for q in probes
                                                    if q["plain_spp_to_match"]])), 4) if probes else None),
        "final_reference_rmse": round(float(np.sqrt(var)), 5),
        "static_plain_rmse": {str(k): round(ep[k - 1], 5) for k in keys},
        "static_cache_rmse": {str(k): round(ec[k - 1], 5) for k in keys},
        "static_time_to_equal_rmse_ratio": ratios,
        "cache_train_steps": int(info.steps), "cache_skipped_steps": int(info.skipped_steps),
        "cache_loss_scale": float(info.loss_scale),
    })
    return out
