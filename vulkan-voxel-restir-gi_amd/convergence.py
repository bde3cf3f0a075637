"""Time-to-equal-RMSE of the learned radiance cache against the plain path tracer.

Protocol (static scene, static camera):
  * reference: the plain unbiased path tracer (`bounces` bounces, Russian roulette from
    `rr_start`, NEE) averaged in FP64 on the GPU over `ref_spp` frames whose seeds are
    disjoint from the measured runs (frames from 2^24 on). Each of its samples is the
    oracle's, bit for bit (the frame parity tests), so this is the oracle's converged
    image;
  * plain: the same estimator, running FP64 mean over `frames` frames;
  * cache: the learned cache (training paths of `bounces` bounces, every other path ends
    at its bounce-1 vertex with the cache), trained online for `warmup` frames first, then
    its running FP64 mean over `frames` frames while it keeps training.
Each frame's GPU time is the HIP-event time of the whole frame (BVH upkeep, trace, NEE
pdf, cache inference and training steps, ReSTIR). For a plain checkpoint k, the cache's
time to equal RMSE is the time of the first cache frame whose running-mean RMSE is at
most the plain RMSE after k frames; the ratio divides it by the plain time of k frames.
The cache warm-up is reported separately (`warmup_ms`)."""
import numpy as np

import vrgi


def _rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


def _frame_ms(ctx):
    return ctx.profile_read(reset=True).frame_ms


def study(objs, mats, lum, cam, W, H, bounces=8, rr_start=2, ref_spp=512, frames=64, warmup=32, nee=1,
          spatial=5, cache_kw=None, log=print):
    cfg = vrgi.config(W, H, bounces, rr_start=rr_start, rr_min=0.1)
    out = {"resolution": f"{W}x{H}", "bounces": bounces, "rr_start": rr_start, "reference_spp": ref_spp,
           "frames": frames, "warmup_frames": warmup}
    plain_ms, plain_rmse = [], []
    with vrgi.Context(cfg, mats, lum) as ctx:
        ctx.add_objects(objs)
        ctx.render(cam, nee_type=nee, spatial=spatial)  # builds the scene
        ctx.sync()
        # reference on seeds disjoint from the measured runs
        log(f"reference: {ref_spp} spp")
        ctx.frame_set(1 << 24)
        ctx.accum_enable(True)
        for k in range(ref_spp):
            ctx.render(cam, nee_type=nee, spatial=spatial)
            if k % 64 == 63:
                ctx.sync()
        ctx.sync()
        ref, n = ctx.accum_read()
        assert n == ref_spp
        # plain: running mean after every frame
        ctx.profile(vrgi.PROFILE_STAGES)
        _frame_ms(ctx)
        ctx.frame_set(0)
        ctx.accum_enable(True)
        log(f"plain: {frames} frames")
        for k in range(frames):
            ctx.render(cam, nee_type=nee, spatial=spatial)
            ctx.sync()
            plain_ms.append(_frame_ms(ctx))
            m, _ = ctx.accum_read()
            plain_rmse.append(_rmse(m, ref))
        ctx.profile(0)
    # cache
    cache_ms, cache_rmse = [], []
    with vrgi.Context(cfg, mats, lum) as ctx:
        ctx.add_objects(objs)
        ctx.render(cam, nee_type=nee, spatial=spatial)
        ctx.sync()
        kw = dict(train_bounces=bounces)
        kw.update(cache_kw or {})
        ctx.cache_enable(**kw)
        ctx.profile(vrgi.PROFILE_STAGES)
        _frame_ms(ctx)
        ctx.frame_set(0)
        log(f"cache: {warmup} warm-up frames")
        warm_ms = 0.0
        for k in range(warmup):
            ctx.render(cam, nee_type=nee, spatial=spatial)
            ctx.sync()
            warm_ms += _frame_ms(ctx)
        ctx.accum_enable(True)
        log(f"cache: {frames} frames")
        for k in range(frames):
            ctx.render(cam, nee_type=nee, spatial=spatial)
            ctx.sync()
            cache_ms.append(_frame_ms(ctx))
            m, _ = ctx.accum_read()
            cache_rmse.append(_rmse(m, ref))
        info = ctx.cache_info()
        ctx.profile(0)
    t_plain = np.cumsum(plain_ms)
    t_cache = np.cumsum(cache_ms)
    ratios = {}
    for k in (1, 4, 16, 64):
        if k > frames:
            continue
        target = plain_rmse[k - 1]
        hit = [j for j in range(frames) if cache_rmse[j] <= target]
        ratios[str(k)] = round(float(t_cache[hit[0]] / t_plain[k - 1]), 4) if hit else None
    out.update({
        "plain_ms_per_frame": round(float(np.mean(plain_ms)), 3),
        "cache_ms_per_frame": round(float(np.mean(cache_ms)), 3),
        "warmup_ms": round(warm_ms, 2),
        "plain_rmse": {str(k): round(plain_rmse[k - 1], 5) for k in (1, 2, 4, 8, 16, 32, 64) if k <= frames},
        "cache_rmse": {str(k): round(cache_rmse[k - 1], 5) for k in (1, 2, 4, 8, 16, 32, 64) if k <= frames},
        "time_to_equal_rmse_ratio": ratios,
        "cache_train_steps": int(info.steps), "cache_last_loss": round(float(info.last_loss), 5),
        "reference_mean_luminance": round(float(np.mean(ref @ np.array([0.2126, 0.7152, 0.0722]))), 5),
    })
    return out
