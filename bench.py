"""Frame benchmark of the MI355X voxel path-tracing hot path.

Workload (one "step" = one full frame): the 256x128x256 seeded terrain + 200 boxes
world (~1.37M exposed faces, ~2.73M triangles, ~13k emissive triangles) at 1920x1080,
1 spp, 4 bounces, Russian roulette from bounce 2, light-BVH NEE on every bounce,
ReSTIR GI temporal + spatial (5 neighbours), a full geometry + light-triangle BVH
rebuild every frame, camera on a 16-frame seeded Catmull-Rom spline.

Second part of the metric (field "learned_cache"): time to equal RMSE of the learned
radiance cache against the plain path tracer on the same scene at 1920x1080: the cache
trains online along the 16-frame camera spline (looped twice, every frame's time counted),
its single frames at four spline poses are compared with the plain tracer's, then both
accumulate 32 frames at pose 0; references are the plain 8-bounce tracer in FP64 (1024 spp
at pose 0, 256 spp at the probe poses), their own noise measured and removed.

Third part (field "batched", not the headline): 64 independent seeded scenes of uneven
size (sides 32..160 voxels, 128^2..512^2 pixels), 2 spp, 3 bounces, NEE, ReSTIR DI (32
candidates, 5 spatial neighbours within 30 px), every scene's BVHs rebuilt in one
multi-tree build per batched frame, one launch per stage for all scenes; beside it the
same frames through 64 separate contexts (one render call per scene and pass).

Fourth part (field "large_dynamic", not the headline): the 512x256x512 world (~5.6M
exposed faces, ~40k emissive voxels) as a device-resident voxel world at 3840x2160, 4
bounces, NEE, ReSTIR DI, learned cache in fp16 with dynamic loss scaling; 32 frames, each
after 1% of the voxels are toggled by 64 seeded random walkers, the dirty chunks re-meshed
on the GPU and every BVH rebuilt; beside it the same frames without the cache, both
against FP64 references at four of the frames and on the final world (time to equal
RMSE, convergence.dynamic_study).

Prints ONE JSON line (rank 0). Progress goes to stderr.
  python bench.py [--gpus 1] [--steps 16] [--warmup 3] [--no-cpu-baseline] [--no-cache-study] [--no-batch]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vulkan-voxel-restir-gi_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "ms/frame & Mrays/s @1080p; time-to-equal-RMSE, learned cache vs plain"
W_PX, H_PX, BOUNCES, RR_START, SPATIAL, NEE = 1920, 1080, 4, 2, 5, 1


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class ProfileOut(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("build_ms", C.c_double), ("trace_ms", C.c_double * 16),
                ("nee_ms", C.c_double * 16), ("resolve_ms", C.c_double), ("frame_ms", C.c_double),
                ("rays", C.c_uint64 * 16), ("touched_bytes", C.c_uint64 * 16), ("build_prims", C.c_uint64),
                ("build_leaf_depth_sum", C.c_uint64), ("build_nodes", C.c_uint64),
                ("nee_listed", C.c_uint64 * 16)]


def light_structure_bytes(n_ltri):
    """Bytes of every structure the NEE pdf walk can read, per light triangle: its
    traversal-copy node (64), triangle (48), instance + leaf ids (8), ancestor list
    (<= 64 levels x 4, about 20 used; 80 counted) and the child records of the light
    BVHs (2 x 64). An upper bound of the unique structure bytes of one launch."""
    return n_ltri * (64 + 48 + 8 + 80 + 128)


NEE_LAZY = "k_nee_select + k_nee_near + k_nee_listed"


def kernel_rooflines(prof, touch_bytes, n_ltri, B):
    """Per-launch algorithmic bytes and average launch time of the per-ray kernels.
    k_trace_shade (one launch per bounce): (48 + 150) B per ray -- ray in 24-32 B and hit
    out 16 B, shading outputs and material / triangle reads 150 B -- plus the unique node
    and triangle bytes the launch touches (counted on the GPU).
    NEE pdf of a bounce (the lazy form: k_nee_select over the bounce's queue, k_nee_near over
    the rays whose pdf matters only through a light crossing near the origin, then
    k_nee_listed over the rays listed for the full pdf): 76 B per queued ray (queue id, MIS weight,
    direction, next-vertex radiance, reflectivity, normal, origin; pdf 0 or list entry out)
    plus 52 B per listed ray (list entry, queue id, MIS weight, direction, origin, normal,
    pdf out) plus the light structures; with every pdf computed (--full-nee, one k_nee_pdf
    launch) 48 B per ray plus the light structures."""
    f = max(prof.frames, 1)
    rays = [prof.rays[b] / f for b in range(B)]
    listed = [prof.nee_listed[b] / f for b in range(B - 1)]
    lazy = any(listed)
    trace_ms = [prof.trace_ms[b] / f for b in range(B)]
    nee_ms = [prof.nee_ms[b] / f for b in range(B - 1)]
    trace_bytes = [198 * rays[b] + touch_bytes[b] for b in range(B)]
    if lazy:
        nee_bytes = [76 * rays[b + 1] + 52 * listed[b] + light_structure_bytes(n_ltri) for b in range(B - 1)]
    else:
        nee_bytes = [48 * rays[b + 1] + light_structure_bytes(n_ltri) for b in range(B - 1)]
    nee = {"launches": B - 1, "bytes": sum(nee_bytes) / (B - 1), "ms": sum(nee_ms) / (B - 1)}
    if lazy:
        nee["listed_fraction"] = round(sum(listed) / max(sum(rays[1:]), 1), 4)
    return {
        "k_trace_shade": {"launches": B, "bytes": sum(trace_bytes) / B, "ms": sum(trace_ms) / B},
        (NEE_LAZY if lazy else "k_nee_pdf"): nee,
    }


def build_alg_bytes(prof):
    """BVH build: 72 N (three 24-B vertices per primitive) + 36 S (S = sum of leaf depths:
    28 B of bin input and 8 B of partition traffic per primitive per level) + 40 (2N-1)
    node writes."""
    N, S = prof.build_prims, prof.build_leaf_depth_sum
    return 72 * N + 36 * S + 40 * prof.build_nodes


def calibrate_copy():
    """A plain device-to-device copy of known size (256 MiB read, 256 MiB written per
    launch), run under rocprofv3 --pmc to scale FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch
    n = 64 << 20  # float32 elements = 256 MiB
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}), flush=True)


def cpu_baseline(objs, mats, lum, cams):
    """The oracle (C++ restatement of the reference algorithms) on the host cores, on a
    bounded sample of the same workload: one full geometry + light BVH build of the world
    (single thread, as the reference's per-object builder) and one complete 1920x1080
    frame of the same configuration (4 bounces, RR from bounce 2, NEE pdf, ReSTIR GI
    temporal + 5 spatial; OpenMP over pixels). ms/frame = build + frame."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS") or usable)
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    s = oracle.Scene(mats, lum)
    s.add_objects(objs)
    t0 = time.perf_counter()
    s.build()
    t_build = time.perf_counter() - t0
    r = oracle.Renderer(W_PX, H_PX, BOUNCES, rr_start=RR_START)
    t0 = time.perf_counter()
    r.render(s, cams[0], 0, nee_type=NEE, spatial=SPATIAL, threads=threads)
    t_frame = time.perf_counter() - t0
    ms = 1000.0 * (t_build + t_frame)
    return {"value": round(ms, 1), "unit": "ms/frame", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model,
            "sample": f"oracle: 1 full geometry + light BVH build (1 thread: one SAH tree over the world's "
                      f"triangles, {t_build:.2f} s) + 1 full 1920x1080 frame with NEE and ReSTIR GI "
                      f"({t_frame:.2f} s, {threads} OpenMP threads)"}


def batched_bench(steps, no_graph=False):
    """The 64-scene batched workload: ms per batched frame (2 passes of every scene, BVH
    rebuilds included), closest-hit Mrays/s, and the same 2 frames of every scene through its
    own context (64 render calls per pass)."""
    import vrgi
    n = 64
    log(f"batched workload: generating {n} scenes")
    scenes = [vrgi.batch_scene(i) for i in range(n)]
    flags = vrgi.FLAG_REBUILD_EVERY_FRAME
    ctxs = []
    try:
        for objs, mats, lum, (W, H), cam in scenes:
            c = vrgi.Context(vrgi.config(W, H, 3, mode=1, radius=30.0, di_candidates=32, flags=flags), mats, lum)
            c.add_objects(objs)
            ctxs.append(c)
        cams = [sc[4] for sc in scenes]
        # the same frames through the scenes' own contexts (each rebuilds; 2 frames = 2 spp)
        for c, cam in zip(ctxs, cams):
            c.render(cam, nee_type=NEE, spatial=SPATIAL)
        for c in ctxs:
            c.sync()
        t0 = time.perf_counter()
        for c, cam in zip(ctxs, cams):
            for _ in range(2):
                c.render(cam, nee_type=NEE, spatial=SPATIAL)
        for c in ctxs:
            c.sync()
        seq_ms = 1000.0 * (time.perf_counter() - t0)
        with vrgi.Batch(ctxs, spp=2, flags=vrgi.BATCH_NO_GRAPH if no_graph else 0) as bt:
            log("batched workload: autotuning and warm-up")
            for _ in range(8):
                bt.render(cams, nee_type=NEE, spatial=SPATIAL)
            bt.sync()
            bt.info()
            log(f"batched workload: timing {steps} batched frames")
            t0 = time.perf_counter()
            for _ in range(steps):
                bt.render(cams, nee_type=NEE, spatial=SPATIAL)
            bt.sync()
            dt = time.perf_counter() - t0
            info = bt.info()
    finally:
        for c in ctxs:
            c.close()
    ms = 1000.0 * dt / steps
    return {"workload": f"{n} seeded scenes (sides 32..160 voxels), 128^2..512^2 px, 2 spp, 3 bounces, NEE, "
                        "ReSTIR DI 32 candidates + temporal + 5 spatial (30 px), BVH rebuild every batched frame",
            "scenes": n, "triangles": int(info.triangles), "light_triangles": int(info.light_triangles),
            "pixels": int(info.pixels), "ms_per_batched_frame": round(ms, 3),
            "mrays_per_s": round(info.rays / dt / 1e6, 2),
            "rays_per_block": int(info.rays_per_block), "pixels_per_block": int(info.pixels_per_block),
            "graph": {0: "none", 1: "one graph per pass", 2: "one graph per batched frame (builds and passes)"}[
                int(info.graph)],
            **({"last_frame_ms": round(float(info.passes_ms), 3)} if info.build_ms < 0 else
               {"last_frame_build_ms": round(float(info.build_ms), 3),
                "last_frame_passes_ms": round(float(info.passes_ms), 3)}),
            "scenes_separately_ms": round(seq_ms, 3)}


def large_dynamic_bench(frames=32):
    """The large dynamic workload: 32 frames of the 4K edited world, plain and with the
    fp16 learned cache, ms per frame including the upload of the frame's voxel edits, the
    re-mesh of the dirty chunks, the light and geometry rebuilds and the frame; the error
    of both against FP64 references at four of the frames and the time to equal RMSE
    (convergence.dynamic_study). The edits themselves (64 seeded random walkers) are
    generated beforehand: they are the caller's side (a game's edits)."""
    import vrgi
    import convergence
    log("large dynamic workload: 512x256x512 world")
    blocks0, nt = vrgi.world(4, 0)
    mats, lum = vrgi.world_materials(4, 0)
    tr = vrgi.world_translucent(4, nt)
    nvox = blocks0.size
    steps = nvox // 100  # 1% of the voxels per frame
    dx, dy, dz = blocks0.shape
    cam = vrgi.camera_look([0.1 * dx, 0.9 * dy, 0.1 * dz], [0.6 * dx, 0.15 * dy, 0.6 * dz])
    blocks = blocks0.copy()
    rng = np.random.default_rng(4)
    walkers = np.stack([rng.integers(0, dx, 64), rng.integers(dy // 4, dy // 2, 64), rng.integers(0, dz, 64)],
                       1).astype(np.uint32)
    edits = [vrgi.world_walk(blocks, nt, 9000 + f, walkers, steps) for f in range(frames)]
    out = {"workload": "512x256x512 terrain+400 boxes as a device voxel world, 3840x2160, 1 spp, 4 bounces, "
                       "RR from bounce 2, NEE, ReSTIR DI 32 candidates + temporal + 5 spatial; 1% of the voxels "
                       "toggled per frame by 64 seeded random walkers; dirty chunks re-meshed, all BVHs rebuilt; "
                       "learned cache in fp16 with dynamic loss scaling",
           "voxels_toggled_per_frame": steps,
           "distinct_voxel_writes_per_frame": int(np.mean([len(v) for _, v in edits]))}
    r = convergence.dynamic_study(blocks0, nt, tr, mats, lum, cam, 3840, 2160, edits, bounces=4, rr_start=RR_START,
                                  mode=1, radius=30.0, di_candidates=32, nee=NEE, spatial=SPATIAL,
                                  probe_frames=(7, 15, 23, 31), probe_ref_spp=64, ref_spp=256, static_frames=16,
                                  cache_kw=dict(precision=vrgi.CACHE_FP16), log=log)
    out.update(r)
    # a refit keeps a tree's topology and needs the same triangles; a frame whose edits
    # re-meshed no chunk could take one
    out["frames_with_unchanged_triangles"] = int(sum(1 for k in r["chunks_remeshed_each"] if k == 0))
    out["bvh_update"] = "rebuild (every frame re-meshes chunks: the triangle set changes)"
    log(f"large dynamic: plain {r['plain_ms_per_frame']} ms, cache {r['cache_ms_per_frame']} ms, "
        f"probes {r['probes']}, static ratios {r['static_time_to_equal_rmse_ratio']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cache-study", action="store_true",
                    help="skip the learned-cache time-to-equal-RMSE study")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy (pmc counter scaling)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched workload")
    ap.add_argument("--batch-no-graph", action="store_true",
                    help="launch the batched frames' kernels directly instead of one graph per batched frame "
                         "(for rocprofv3 kernel traces, which end the process inside such a graph)")
    ap.add_argument("--no-large", action="store_true", help="skip the large dynamic workload")
    ap.add_argument("--serial-build", action="store_true",
                    help="one BVH buffer set built on the frame stream (default: frame f + 1 builds beside frame f)")
    ap.add_argument("--serial-nee", action="store_true",
                    help="NEE pdf launches after all bounces on one stream (default: beside the next trace)")
    ap.add_argument("--full-nee", action="store_true",
                    help="NEE pdf of every continued ray (default: only where it can change the outputs)")
    ap.add_argument("--sort-bits", type=int, default=0,
                    help="Morton ray-queue ordering: origin bits per axis (0 = compaction order)")
    a = ap.parse_args()
    if a.calibrate_copy:
        calibrate_copy()
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (a frame would shard by "
              "image rows, a batch by scenes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import vrgi
    log("generating the 256x128x256 world and meshing its chunks")
    blocks, nt = vrgi.world(1, 0)
    mats, lum = vrgi.world_materials(1, 0)
    tr = vrgi.world_translucent(1, nt)
    objs = vrgi.chunk_objects(blocks, nt, tr)
    cams = [vrgi.camera_spline(1, k, 16) for k in range(16)]
    cfg = vrgi.config(W_PX, H_PX, BOUNCES, rr_start=RR_START, rr_min=0.1,
                      flags=vrgi.FLAG_REBUILD_EVERY_FRAME | (vrgi.FLAG_SERIAL_NEE if a.serial_nee else 0)
                      | (vrgi.FLAG_SERIAL_BUILD if a.serial_build else 0) | (vrgi.FLAG_FULL_NEE_PDF if a.full_nee else 0),
                      sort_bits=a.sort_bits)
    L = vrgi.lib()
    with vrgi.Context(cfg, mats, lum) as ctx:
        log("uploading objects (per-object light BVHs)")
        ctx.add_objects(objs)
        log(f"warmup {a.warmup} frames")
        for k in range(a.warmup):
            ctx.render(cams[k % 16], nee_type=NEE, spatial=SPATIAL)
        ctx.sync()
        ctx.stats(reset=True)
        check = vrgi.check
        # per-kernel HIP events on the library's stream during the timed frames
        check(L.vrgi_profile(ctx.h, 1))
        log(f"timing {a.steps} frames")
        t0 = time.perf_counter()
        for k in range(a.steps):
            ctx.render(cams[k % 16], nee_type=NEE, spatial=SPATIAL)
        ctx.sync()
        dt = time.perf_counter() - t0
        st = ctx.stats(reset=True)
        prof = ProfileOut()
        check(L.vrgi_profile_read(ctx.h, C.byref(prof), 1))
        ms = 1000.0 * dt / a.steps
        mrays = st.rays_closest / dt / 1e6
        log(f"{ms:.3f} ms/frame, {mrays:.1f} Mrays/s closest-hit, depth {st.bvh_depth}, nodes {st.bvh_nodes}")
        # one counting frame after the timed region: unique node / triangle bytes per bounce
        check(L.vrgi_profile(ctx.h, 2))
        ctx.render(cams[0], nee_type=NEE, spatial=SPATIAL)
        touch = ProfileOut()
        check(L.vrgi_profile_read(ctx.h, C.byref(touch), 1))
        check(L.vrgi_profile(ctx.h, 0))
        n_ltri = (len(ctx.read(vrgi.BUF_LIGHT_GEOM_BVH, np.uint8)) // 40 + 1) // 2
    f = max(prof.frames, 1)
    P = W_PX * H_PX
    stages = {"build": prof.build_ms / f, "resolve": prof.resolve_ms / f, "frame": prof.frame_ms / f}
    for b in range(BOUNCES):
        stages[f"trace_{b}"] = prof.trace_ms[b] / f
        if b + 1 < BOUNCES:
            stages[f"nee_{b}"] = prof.nee_ms[b] / f
    kr = kernel_rooflines(prof, list(touch.touched_bytes), n_ltri, BOUNCES)
    dom = max(kr, key=lambda k: kr[k]["ms"] * kr[k]["launches"])
    kd = kr[dom]
    achieved = kd["bytes"] / (kd["ms"] * 1e-3) / 1e9 if kd["ms"] > 0 else 0.0
    # FETCH_SIZE + WRITE_SIZE per launch of the same command under rocprofv3 --pmc (one
    # counter per pass, scaled by a plain copy): L2-miss traffic to MALL / HBM
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(dom, {}).get("l2_miss_bytes_per_launch")
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "traffic_source": "profiles/traffic.json: L2-miss traffic per launch (FETCH_SIZE + WRITE_SIZE, "
                                  "scaled by a plain copy) from rocprofv3 --pmc runs of this workload "
                                  "(profile_summary.py collect counters)",
                "alg_bytes_per_launch": int(kd["bytes"]), "avg_launch_ms": round(kd["ms"], 4),
                "launches_per_frame": kd["launches"]}
    others = {k: {"alg_bytes_per_launch": int(v["bytes"]), "avg_launch_ms": round(v["ms"], 4),
                  "achieved_GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 2) if v["ms"] > 0 else 0.0,
                  **({"listed_fraction": v["listed_fraction"]} if "listed_fraction" in v else {})}
              for k, v in kr.items()}
    if prof.build_ms > 0:
        bb = build_alg_bytes(prof)
        others["bvh_build (all builder kernels)"] = {"alg_bytes_per_frame": int(bb), "ms": round(stages["build"], 4),
                                                     "achieved_GBps": round(bb / (stages["build"] * 1e-3) / 1e9, 2)}
    cache = None
    if not a.no_cache_study:
        import convergence
        log("learned cache vs plain: time to equal RMSE (1920x1080, spline camera, FP64 references)")
        cache = convergence.study(objs, mats, lum, cams, W_PX, H_PX, bounces=8, rr_start=RR_START, ref_spp=1024,
                                  frames=32, warmup=32, nee=NEE, spatial=SPATIAL, probe_poses=(0, 4, 8, 12),
                                  probe_ref_spp=256, cache_kw=dict(query_vertex=0), log=log)
    batched = None
    if not a.no_batch:
        batched = batched_bench(a.steps, a.batch_no_graph)
    large = None
    if not a.no_large:
        large = large_dynamic_bench()
    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline (oracle on the host cores, bounded sample)")
        cpu = cpu_baseline(objs, mats, lum, cams)
    out = {
        "metric": METRIC, "value": round(ms, 4), "unit": "ms/frame", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "256x128x256 terrain+200 boxes, 1920x1080, 1 spp, 4 bounces, RR from bounce 2, "
                               "NEE, ReSTIR GI 1 temporal + 5 spatial, BVH rebuild every frame, 16-frame spline",
                   "triangles": int(prof.build_prims), "pixels": P, "bounces": BOUNCES},
        "mrays_per_s": round(mrays, 2),
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "roofline": roofline,
        "kernels": others,
        "learned_cache": cache,
        "batched": batched,
        "large_dynamic": large,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
