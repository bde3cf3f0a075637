"""rocprofv3 runs of the benchmark workloads and their text summaries.

  python3 profile_summary.py collect KIND OUT    # runs rocprofv3 (child processes) into OUT/KIND
  python3 profile_summary.py summary KIND OUT   # prints the summary of OUT/KIND
  python3 profile_summary.py workload cache     # the learned-cache frames profiled by KIND cache

KIND:
  frame     kernel trace of the 1080p frame workload (bench.py, pipelined frame graph)
  serial    the same with every stage serial (--serial-build --serial-nee)
  batched   kernel trace of the 64-scene batched workload
  cache     kernel trace of 24 learned-cache frames at 1080p (8 bounces, spline camera)
  counters  PMC passes over 4 frames of the 1080p workload (L2 hits / misses; waves, VALU
            and wave cycles; FETCH_SIZE; WRITE_SIZE), one counter group per rocprofv3 run,
            plus a plain 256 MiB copy (bench.py --calibrate-copy) that scales FETCH_SIZE
            and WRITE_SIZE to bytes; `summary counters OUT --traffic FILE` also writes the
            per-launch L2-miss traffic that bench.py reports as roofline.traffic.

Counter units (rocprofv3 sums a counter over its instances; GRBM_GUI_ACTIVE is summed over
the 8 XCDs, so one XCD's busy cycles are GRBM_GUI_ACTIVE / 8):
  L2 hit rate   TCC_HIT / (TCC_HIT + TCC_MISS)
  VALU busy     100 * 4 * SQ_ACTIVE_INST_VALU / SIMDs / (GRBM_GUI_ACTIVE / 8)
  occupancy     100 * 4 * SQ_WAVE_CYCLES / (GRBM_GUI_ACTIVE / 8) / CUs / 32 (waves per CU
                against the 32 wave slots of a CU)
  FETCH_SIZE / WRITE_SIZE  KiB of L2-miss traffic (requests from the L2 to the memory side:
                the MALL, then HBM), scaled by the copy's reading: bytes = KiB x 1024 x
                (copy bytes / (copy KiB x 1024))."""
import collections
import csv
import glob
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")
SELF = os.path.abspath(__file__)
CUS, SIMDS, XCDS = 256, 1024, 8

HEAD = ("1080p frame workload (256x128x256 world + 200 boxes, 2.75M triangles, 1920x1080, 1 spp, 4 bounces, RR "
        "from bounce 2, NEE every bounce, ReSTIR GI temporal + 5 spatial, BVH rebuild every frame, 16-frame "
        "camera spline)")
FRAME_ARGS = ["--no-cpu-baseline", "--no-cache-study", "--no-batch", "--no-large"]
BUILDER = ("k_flatten", "k_flatten_lights", "k_light_ancestry", "k_init", "k_roots", "k_items", "k_clear_node_data",
           "k_bounds", "k_bin", "k_split", "k_count", "k_part_a", "k_part_b", "k_children", "k_subtree",
           "k_advance", "k_power_pairs", "k_wide", "k_wide4")
PASSES = [("p1", "TCC_HIT_sum TCC_MISS_sum"),
          ("p2", "SQ_WAVES SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_WAVE_CYCLES GRBM_GUI_ACTIVE"),
          ("p3", "FETCH_SIZE"),
          ("p4", "WRITE_SIZE")]


def _cmds(kind):
    """(dir, rocprofv3 arguments, program) of each run of KIND."""
    py = ["python3"]
    if kind == "frame":
        return [("trace", ["--kernel-trace", "--stats"], py + [BENCH] + FRAME_ARGS)]
    if kind == "serial":
        return [("trace", ["--kernel-trace", "--stats"], py + [BENCH] + FRAME_ARGS + ["--serial-build", "--serial-nee"])]
    if kind == "batched":
        return [("trace", ["--kernel-trace", "--stats"],
                 py + [BENCH, "--no-cpu-baseline", "--no-cache-study", "--no-large", "--batch-no-graph"])]
    if kind == "cache":
        return [("trace", ["--kernel-trace", "--stats"], py + [SELF, "workload", "cache"])]
    if kind == "counters":
        short = py + [BENCH, "--steps", "4", "--warmup", "1"] + FRAME_ARGS
        runs = [(p, ["--pmc"] + c.split(), short) for p, c in PASSES]
        runs += [("calib_fetch", ["--pmc", "FETCH_SIZE"], py + [BENCH, "--calibrate-copy"]),
                 ("calib_write", ["--pmc", "WRITE_SIZE"], py + [BENCH, "--calibrate-copy"])]
        return runs
    raise SystemExit(f"unknown kind {kind}")


def _cmdline(args, prog):
    rel = lambda a: a.replace(ROOT, "/root/repo")
    return "rocprofv3 " + " ".join(args) + " --output-format csv -d OUT -o run -- " + " ".join(rel(a) for a in prog)


def collect(kind, out):
    """One rocprofv3 run per entry of _cmds(kind), each as a child process under its own
    time limit; the program's stdout (the bench line) goes to OUT/KIND/<dir>/stdout.txt."""
    env = dict(os.environ, TMPDIR="/tmp")
    for d, args, prog in _cmds(kind):
        dst = os.path.join(out, kind, d)
        os.makedirs(dst, exist_ok=True)
        limit = "120" if "--pmc" in args else "600"
        cmd = ["timeout", "-s", "KILL", limit, "rocprofv3"] + args + ["--output-format", "csv", "-d", dst, "-o",
                                                                     "run", "--"] + prog
        print(f"[profile] {kind}/{d}: {' '.join(cmd)}", flush=True)
        with open(os.path.join(dst, "stdout.txt"), "w") as fo, open(os.path.join(dst, "stderr.txt"), "w") as fe:
            r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=fo, stderr=fe)
        print(f"[profile] {kind}/{d}: exit {r.returncode}", flush=True)
        if r.returncode != 0:
            sys.exit(r.returncode)


def _find(dirname, suffix):
    f = sorted(glob.glob(os.path.join(dirname, "**", "*" + suffix), recursive=True))
    if not f:
        raise SystemExit(f"no *{suffix} under {dirname}")
    return f[0]


def _short(name):
    n = name.split("(")[0].replace("void ", "").replace("vrgi::", "")
    return n.strip()


def _trace(dirname):
    rows = []
    for r in csv.DictReader(open(_find(dirname, "kernel_trace.csv"))):
        rows.append((_short(r["Kernel_Name"]), int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                     int(r.get("Stream_Id", r.get("Queue_Id", 0)) or 0)))
    rows.sort(key=lambda x: x[1])
    return rows


def _bench_line(dirname):
    try:
        lines = [l for l in open(os.path.join(dirname, "stdout.txt")) if l.startswith("{")]
        return json.loads(lines[-1]) if lines else None
    except OSError:
        return None


def _table(rows, total=None, limit=40):
    agg = collections.defaultdict(lambda: [0, 0.0, 1e30, 0.0])
    for n, s, e, _ in rows:
        a = agg[n]
        d = (e - s) / 1e3
        a[0] += 1
        a[1] += d
        a[2] = min(a[2], d)
        a[3] = max(a[3], d)
    tot = total or sum(a[1] for a in agg.values()) or 1.0
    out = [f"{'kernel':56s} {'calls':>6s} {'total ms':>10s} {'avg us':>9s} {'min us':>8s} {'max us':>9s} {'%':>6s}"]
    for n, a in sorted(agg.items(), key=lambda kv: -kv[1][1])[:limit]:
        out.append(f"{n[:56]:56s} {a[0]:6d} {a[1] / 1e3:10.3f} {a[1] / a[0]:9.1f} {a[2]:8.1f} {a[3]:9.1f} "
                   f"{100 * a[1] / tot:6.2f}")
    return out


def _frame_summary(kind, out, title):
    d = os.path.join(out, kind, "trace")
    rows = _trace(d)
    bl = _bench_line(d)
    (_, args, prog), = _cmds(kind)
    lines = [f"Kernel time, {HEAD}.", title, "Made by (files of this repository only):",
             f"  python3 vulkan-voxel-restir-gi_amd/profile_summary.py collect {kind} OUT; "
             f"python3 vulkan-voxel-restir-gi_amd/profile_summary.py summary {kind} OUT",
             "which runs, from /tmp with TMPDIR=/tmp:", "  " + _cmdline(args, prog)]
    warm, steps = 3, 16
    if bl:
        warm, steps = bl["warmup"], bl["steps"]
        rf = bl["roofline"]
        lines.append(f"The same run's bench line: {bl['value']} ms/frame, {bl['mrays_per_s']} Mrays/s; dominant "
                     f"kernel {rf['kernel']} {rf['avg_launch_ms']} ms/launch, {rf['achieved']} GB/s "
                     f"(frac {rf['frac']}); build {bl['stages_ms']['build']} ms/frame (HIP events on the "
                     f"launch streams over the {steps} timed frames)")
        for k, v in bl["kernels"].items():
            lines.append(f"  bench {k}: " + ", ".join(f"{a} {b}" for a, b in v.items()))
    # per-ray kernels over the timed frames, by launch index
    B = 4
    per_ray = [("k_trace_shade<32, false, false>", B), ("k_nee_select", B - 1), ("k_nee_near<32>", B - 1),
               ("k_nee_listed<32>", B - 1), ("k_nee_pdf<32>", B - 1)]
    lines.append(f"rocprof kernel trace over the {steps} timed frames (frames {warm}..{warm + steps - 1}):")
    for name, per in per_ray:
        ds = [(e - s) / 1e3 for n, s, e, _ in rows if n == name]
        sel = ds[warm * per:(warm + steps) * per]
        if sel:
            lines.append(f"  {name}: {len(sel)} launches, avg {sum(sel) / len(sel) / 1e3:.4f} ms/launch")
    # one timed frame's build: the build stream's launches from its k_flatten to the next one
    fl = [i for i, r in enumerate(rows) if r[0] == "k_flatten"]
    f = warm + steps // 2
    if len(fl) > f + 1:
        i0, i1 = fl[f], fl[f + 1]
        st = rows[i0][3]
        b = [r for r in rows[i0:i1] if r[3] == st]
        last = max(i for i, r in enumerate(b) if r[0] in BUILDER or r[0].startswith("k_subtree"))
        b = b[:last + 1]
        wall = (b[-1][2] - b[0][1]) / 1e6
        kt = sum(e - s for _, s, e, _ in b) / 1e6
        lines += ["", f"BVH build of timed frame {f} (stream {st}, from its k_flatten to its last builder "
                      f"launch): wall {wall:.3f} ms over {len(b)} launches, kernel time {kt:.3f} ms"]
        lines += ["  " + l for l in _table(b, limit=30)]
    lines += ["", "Whole run (warm-up, timed and counting frames, uploads):"] + _table(rows)
    return lines


def _batched_summary(out):
    d = os.path.join(out, "batched", "trace")
    rows = _trace(d)
    bl = _bench_line(d)
    (_, args, prog), = _cmds("batched")
    bi = [i for i, r in enumerate(rows) if "_b<" in r[0] or r[0].endswith("_b")]
    lines = ["Kernel time, batched workload: 64 seeded scenes (sides 32..160 voxels), 128^2..512^2 px, 2 spp, "
             "3 bounces, NEE, ReSTIR DI 32 candidates + temporal + 5 spatial (30 px), BVH rebuild every batched "
             "frame.", "Made by (files of this repository only):",
             "  python3 vulkan-voxel-restir-gi_amd/profile_summary.py collect batched OUT; "
             "python3 vulkan-voxel-restir-gi_amd/profile_summary.py summary batched OUT",
             "which runs, from /tmp with TMPDIR=/tmp:", "  " + _cmdline(args, prog),
             "The batched frames' kernels are launched directly here (--batch-no-graph): under rocprofv3's kernel "
             "trace the process ends inside the graph of a batched frame (builds and passes, ~1500 nodes); the "
             "bench itself runs one graph per batched frame."]
    if bl and bl.get("batched"):
        b = bl["batched"]
        parts = (f"last frame build {b['last_frame_build_ms']} ms, passes {b['last_frame_passes_ms']} ms"
                 if "last_frame_build_ms" in b else f"last frame {b.get('last_frame_ms')} ms")
        lines.append(f"The same run's bench line: {b['ms_per_batched_frame']} ms per batched frame, "
                     f"{b['mrays_per_s']} Mrays/s, {parts}; the same frames through 64 contexts "
                     f"{b['scenes_separately_ms']} ms")
    if bi:
        w = rows[bi[0]:bi[-1] + 1]
        lines += ["Window: from the first to the last batched launch (warm-up, the autotuner's trial frames "
                  "and the timed frames; builds included). One batched frame = 2 spp passes.", ""] + _table(w)
    return lines


def _cache_summary(out):
    d = os.path.join(out, "cache", "trace")
    rows = _trace(d)
    (_, args, prog), = _cmds("cache")
    lines = ["Kernel time, learned radiance cache at 1080p: the 256x128x256 world, 8-bounce training paths from "
             "3% of the pixels, the other paths end in the cache at their first hit, 4 training steps of 16384 "
             "records per frame, 24 frames along the 16-frame camera spline (the cache trains from its first "
             "frame).", "Made by (files of this repository only):",
             "  python3 vulkan-voxel-restir-gi_amd/profile_summary.py collect cache OUT; "
             "python3 vulkan-voxel-restir-gi_amd/profile_summary.py summary cache OUT",
             "which runs, from /tmp with TMPDIR=/tmp:", "  " + _cmdline(args, prog)]
    try:
        lines += ["Frame times (HIP events around each frame): " +
                  " ".join(l.split()[2] for l in open(os.path.join(d, "stdout.txt")) if l.startswith("frame "))]
    except OSError:
        pass
    nrc = [i for i, r in enumerate(rows) if "nrc" in r[0]]
    if nrc:
        w = rows[nrc[0]:]
        lines += ["Window: from the first cache kernel on (all cache frames; kernels of the training stream "
                  "run beside the frame's rays, so durations overlap).", ""] + _table(w)
    return lines


def _counters(out):
    base = os.path.join(out, "counters")
    per = collections.defaultdict(lambda: collections.defaultdict(lambda: collections.defaultdict(float)))
    for p, _ in PASSES:
        for r in csv.DictReader(open(_find(os.path.join(base, p), "counter_collection.csv"))):
            per[_short(r["Kernel_Name"])][r["Counter_Name"]][int(r["Dispatch_Id"])] += float(r["Counter_Value"])
    calib = {}
    for d, c in (("calib_fetch", "FETCH_SIZE"), ("calib_write", "WRITE_SIZE")):
        vals = [float(r["Counter_Value"]) for r in csv.DictReader(open(_find(os.path.join(base, d),
                                                                               "counter_collection.csv")))
                if "copyBuffer" in r["Kernel_Name"] and r["Counter_Name"] == c]
        calib[c] = max(vals) if vals else 0.0
    copy_bytes = 256 << 20
    scale = {c: (copy_bytes / (v * 1024) if v else 1.0) for c, v in calib.items()}
    mean = lambda d: sum(d.values()) / len(d) if d else 0.0
    res = {}
    for k, cs in per.items():
        gui = mean(cs.get("GRBM_GUI_ACTIVE", {})) / XCDS
        hit, miss = mean(cs.get("TCC_HIT_sum", {})), mean(cs.get("TCC_MISS_sum", {}))
        res[k] = {
            "n": max(len(v) for v in cs.values()),
            "l2_hit": hit / (hit + miss) if hit + miss else None,
            "valu_busy": 100 * 4 * mean(cs.get("SQ_ACTIVE_INST_VALU", {})) / SIMDS / gui if gui else None,
            "occupancy": 100 * 4 * mean(cs.get("SQ_WAVE_CYCLES", {})) / gui / CUS / 32 if gui else None,
            "fetch": mean(cs.get("FETCH_SIZE", {})) * 1024 * scale["FETCH_SIZE"],
            "write": mean(cs.get("WRITE_SIZE", {})) * 1024 * scale["WRITE_SIZE"],
        }
    return res, calib, scale, _bench_line(os.path.join(base, "p3"))


def _counters_summary(out, traffic_file=None):
    res, calib, scale, bl = _counters(out)
    lines = [f"Hardware counters per dispatch, {HEAD}; rocprofv3 PMC, one counter group per run.",
             "Made by (files of this repository only):",
             "  python3 vulkan-voxel-restir-gi_amd/profile_summary.py collect counters OUT; "
             "python3 vulkan-voxel-restir-gi_amd/profile_summary.py summary counters OUT --traffic "
             "profiles/traffic.json", "which runs, from /tmp with TMPDIR=/tmp:"]
    lines += ["  " + _cmdline(a, p) for _, a, p in _cmds("counters")]
    lines += [f"Calibration, a plain 256 MiB copy (268435456 B read and written): FETCH_SIZE reads "
              f"{calib.get('FETCH_SIZE', 0):.1f} KiB, WRITE_SIZE {calib.get('WRITE_SIZE', 0):.1f} KiB -> bytes = "
              f"KiB x 1024 x {scale['FETCH_SIZE']:.4f} (fetch), x {scale['WRITE_SIZE']:.4f} (write).",
              "fetch / write: L2-miss traffic (the L2's requests to the memory side: MALL, then HBM), MB per "
              "dispatch. L2 hit: TCC_HIT / (TCC_HIT + TCC_MISS). VALU busy and occupancy: see the module "
              "docstring of profile_summary.py.", ""]
    alg = {}
    if bl:
        for k, v in bl["kernels"].items():
            if "alg_bytes_per_launch" in v:
                alg[k] = v["alg_bytes_per_launch"]
    lines.append(f"{'kernel':48s} {'n':>4s} {'L2 hit':>7s} {'VALU%':>6s} {'occ%':>6s} {'fetch MB':>9s} "
                 f"{'write MB':>9s}")
    for k, r in sorted(res.items(), key=lambda kv: -(kv[1]["fetch"] + kv[1]["write"])):
        if r["fetch"] + r["write"] < 1e6 and not any(s in k for s in ("trace_shade", "nee_")):
            continue
        f = lambda v, fmt: (fmt % v) if v is not None else "-"
        lines.append(f"{k[:48]:48s} {r['n']:4d} {f(r['l2_hit'], '%7.3f')} {f(r['valu_busy'], '%6.1f')} "
                     f"{f(r['occupancy'], '%6.1f')} {r['fetch'] / 1e6:9.2f} {r['write'] / 1e6:9.2f}")
    # the per-ray kernels against their algorithmic bytes
    ts = res.get("k_trace_shade<32, false, false>")
    sel, near, lst = res.get("k_nee_select"), res.get("k_nee_near<32>"), res.get("k_nee_listed<32>")
    traffic = {"_scale": {"fetch_bytes_per_KiB_reading": round(1024 * scale["FETCH_SIZE"], 2),
                          "write_bytes_per_KiB_reading": round(1024 * scale["WRITE_SIZE"], 2),
                          "calibration": "plain 256 MiB copy (bench.py --calibrate-copy)",
                          "meaning": "L2-miss traffic per launch (FETCH_SIZE + WRITE_SIZE), to MALL / HBM"}}
    lines.append("")
    if ts:
        t = ts["fetch"] + ts["write"]
        traffic["k_trace_shade"] = {"l2_miss_bytes_per_launch": int(t), "fetch_bytes": int(ts["fetch"]),
                                    "write_bytes": int(ts["write"]), "l2_hit_rate": ts["l2_hit"]}
        a = alg.get("k_trace_shade")
        lines.append(f"k_trace_shade: L2-miss traffic {t / 1e6:.1f} MB per launch" +
                     (f" against {a / 1e6:.1f} MB algorithmic ({t / a:.2f}x)" if a else ""))
    if sel and near and lst:
        by = lambda r: r["fetch"] + r["write"]
        t = by(sel) + by(near) + by(lst)
        key = "k_nee_select + k_nee_near + k_nee_listed"
        traffic[key] = {"l2_miss_bytes_per_launch": int(t), "select_bytes": int(by(sel)),
                        "near_bytes": int(by(near)), "listed_bytes": int(by(lst))}
        a = alg.get(key)
        lines.append(f"{key} (one lazy NEE pdf of a bounce): L2-miss traffic {t / 1e6:.1f} MB" +
                     (f" against {a / 1e6:.1f} MB algorithmic ({t / a:.2f}x)" if a else ""))
    if traffic_file:
        with open(traffic_file, "w") as fo:
            json.dump(traffic, fo, indent=1)
    return lines


def cache_workload():
    """24 learned-cache frames at 1080p along the camera spline (the cache study's setting)."""
    sys.path.insert(0, os.path.dirname(SELF))
    import vrgi
    blocks, nt = vrgi.world(1, 0)
    mats, lum = vrgi.world_materials(1, 0)
    tr = vrgi.world_translucent(1, nt)
    objs = vrgi.chunk_objects(blocks, nt, tr)
    cams = [vrgi.camera_spline(1, k, 16) for k in range(16)]
    with vrgi.Context(vrgi.config(1920, 1080, 8, rr_start=2, rr_min=0.1), mats, lum) as ctx:
        ctx.add_objects(objs)
        ctx.render(cams[0], nee_type=1, spatial=5)
        ctx.sync()
        ctx.cache_enable(train_bounces=8, query_vertex=0)
        ctx.profile(vrgi.PROFILE_STAGES)
        for k in range(24):
            ctx.render(cams[k % 16], nee_type=1, spatial=5)
            ctx.sync()
            print(f"frame {k} {ctx.profile_read(reset=True).frame_ms:.3f} ms", flush=True)


def main(argv):
    if len(argv) >= 2 and argv[0] == "workload" and argv[1] == "cache":
        return cache_workload()
    if len(argv) < 3 or argv[0] not in ("collect", "summary"):
        raise SystemExit(__doc__)
    op, kind, out = argv[0], argv[1], argv[2]
    if op == "collect":
        return collect(kind, out)
    if kind == "frame":
        lines = _frame_summary("frame", out, "Pipelined frame graph (the bench's default): the NEE pdf passes run "
                                             "beside the next bounce's trace and frame f + 1's build beside frame "
                                             "f's rays, so a launch's duration includes the time it shares the GPU.")
    elif kind == "serial":
        lines = _frame_summary("serial", out, "Serial frame graph (--serial-build --serial-nee): every launch runs "
                                              "alone, so its duration is its own.")
    elif kind == "batched":
        lines = _batched_summary(out)
    elif kind == "cache":
        lines = _cache_summary(out)
    elif kind == "counters":
        tf = argv[argv.index("--traffic") + 1] if "--traffic" in argv else None
        lines = _counters_summary(out, tf)
    else:
        raise SystemExit(f"unknown kind {kind}")
    print("\n".join(lines))


if __name__ == "__main__":
    main(sys.argv[1:])
