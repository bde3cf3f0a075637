"""Frame benchmark of the MI355X voxel path-tracing hot path.

Workload (one "step" = one full frame): the 256x128x256 seeded terrain + 200 boxes
world (~1.37M exposed faces, ~2.73M triangles, ~13k emissive triangles) at 1920x1080,
1 spp, 4 bounces, Russian roulette from bounce 2, light-BVH NEE on every bounce,
ReSTIR GI temporal + spatial (5 neighbours), a full geometry + light-triangle BVH
rebuild every frame, camera on a 16-frame seeded Catmull-Rom spline.

Prints ONE JSON line (rank 0). Progress goes to stderr.
  python bench.py [--gpus 1] [--steps 16] [--warmup 3] [--no-cpu-baseline]
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
                ("build_leaf_depth_sum", C.c_uint64), ("build_nodes", C.c_uint64)]


def alg_bytes(prof, P, B, spatial):
    """Algorithmic (compulsory) bytes per frame of each stage, as defined for this path:
    build 72 N + 36 S + 40 (2N-1) (N primitives, S = sum of leaf depths);
    trace+shade per bounce (48 + 150) R_b + U_b (R_b rays, U_b unique node / triangle
    bytes touched, counted on the GPU); resolve 64 P B + 228 P + (k+2) 80 P + 128 P."""
    N, S = prof.build_prims, prof.build_leaf_depth_sum
    build = 72 * N + 36 * S + 40 * prof.build_nodes
    frames = max(prof.frames, 1)
    trace = [198 * prof.rays[b] / frames + prof.touched_bytes[b] for b in range(B)]
    resolve = 64 * P * B + 228 * P + (spatial + 2) * 80 * P + 128 * P
    return build, trace, resolve


def cpu_baseline(objs, mats, lum, cams, budget_rows=24):
    """The oracle (C++ restatement of the reference algorithms) on the host cores, on a
    bounded sample of the same workload: one full geometry + light BVH build of the
    world (single thread, as the reference's per-object builder) and `budget_rows` rows
    of one 1920x1080 frame (all bounces, NEE pdf; OpenMP over pixels). Extrapolated to
    a full frame: build + rows * (1080 / budget_rows)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    s = oracle.Scene(mats, lum)
    s.add_objects(objs)
    t0 = time.perf_counter()
    s.build()
    t_build = time.perf_counter() - t0
    r = oracle.Renderer(W_PX, H_PX, BOUNCES, rr_start=RR_START)
    y0 = H_PX // 2 - budget_rows // 2
    t0 = time.perf_counter()
    r.render_rows(s, cams[0], 0, y0, y0 + budget_rows, nee_type=NEE, threads=threads)
    t_rows = time.perf_counter() - t0
    ms = 1000.0 * (t_build + t_rows * H_PX / budget_rows)
    return {"value": round(ms, 1), "unit": "ms/frame", "cores": threads, "kind": "port",
            "sample": f"oracle: 1 full BVH build (1 thread, {t_build:.1f} s) + {budget_rows} of 1080 rows "
                      f"({t_rows:.1f} s, {threads} threads), extrapolated to one frame; ReSTIR passes excluded"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=24)
    ap.add_argument("--sort-bits", type=int, default=0,
                    help="Morton ray-queue ordering: origin bits per axis (0 = compaction order)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the frame would shard by "
              "image rows / scenes; see the design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import vrgi
    log("generating the 256x128x256 world and meshing its chunks")
    blocks, nt = vrgi.world(1, 0)
    mats, lum = vrgi.world_materials(1, 0)
    tr = vrgi.world_translucent(1, nt)
    objs = vrgi.chunk_objects(blocks, nt, tr)
    cams = [vrgi.camera_spline(1, k, 16) for k in range(16)]
    cfg = vrgi.config(W_PX, H_PX, BOUNCES, rr_start=RR_START, rr_min=0.1,
                      flags=vrgi.FLAG_REBUILD_EVERY_FRAME, sort_bits=a.sort_bits)
    L = vrgi.lib()
    with vrgi.Context(cfg, mats, lum) as ctx:
        log("uploading objects (per-object light BVHs)")
        ctx.add_objects(objs)
        log(f"warmup {a.warmup} frames")
        for k in range(a.warmup):
            ctx.render(cams[k % 16], nee_type=NEE, spatial=SPATIAL)
        ctx.sync()
        ctx.stats(reset=True)
        log(f"timing {a.steps} frames")
        t0 = time.perf_counter()
        for k in range(a.steps):
            ctx.render(cams[k % 16], nee_type=NEE, spatial=SPATIAL)
        ctx.sync()
        dt = time.perf_counter() - t0
        st = ctx.stats(reset=True)
        ms = 1000.0 * dt / a.steps
        mrays = st.rays_closest / dt / 1e6
        log(f"{ms:.3f} ms/frame, {mrays:.1f} Mrays/s closest-hit, depth {st.bvh_depth}, nodes {st.bvh_nodes}")
        # per-stage HIP-event timing on the library's stream, then one counting frame
        n_prof = min(a.steps, 8)
        check = vrgi.check
        check(L.vrgi_profile(ctx.h, 1))
        for k in range(n_prof):
            ctx.render(cams[k % 16], nee_type=NEE, spatial=SPATIAL)
        prof = ProfileOut()
        check(L.vrgi_profile_read(ctx.h, C.byref(prof), 1))
        check(L.vrgi_profile(ctx.h, 2))
        ctx.render(cams[0], nee_type=NEE, spatial=SPATIAL)
        touch = ProfileOut()
        check(L.vrgi_profile_read(ctx.h, C.byref(touch), 1))
        check(L.vrgi_profile(ctx.h, 0))
        for b in range(BOUNCES):
            prof.touched_bytes[b] = touch.touched_bytes[b]
    f = max(prof.frames, 1)
    P = W_PX * H_PX
    stages = {"build": prof.build_ms / f, "resolve": prof.resolve_ms / f, "frame": prof.frame_ms / f}
    for b in range(BOUNCES):
        stages[f"trace_{b}"] = prof.trace_ms[b] / f
        if b + 1 < BOUNCES:
            stages[f"nee_{b}"] = prof.nee_ms[b] / f
    b_build, b_trace, b_resolve = alg_bytes(prof, P, BOUNCES, SPATIAL)
    t_trace = sum(stages[f"trace_{b}"] for b in range(BOUNCES))
    cands = {"bvh_build": (b_build, stages["build"]), "trace_shade": (sum(b_trace), t_trace),
             "resolve_restir": (b_resolve, stages["resolve"])}
    dom = max(cands, key=lambda k: cands[k][1])
    alg, t_ms = cands[dom]
    achieved = alg / (t_ms * 1e-3) / 1e9 if t_ms > 0 else 0.0
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "alg_bytes_per_frame": int(alg), "ms_per_frame": round(t_ms, 4)}
    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline (oracle on the host cores, bounded sample)")
        cpu = cpu_baseline(objs, mats, lum, cams, a.cpu_rows)
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
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
