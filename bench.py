#!/usr/bin/env python3
"""PBF particle-steps/s on one MI355X: 1,048,576-particle dam break, 4 solver
iterations, XSPH 0.01, marching cubes on the 256^3 density field every step.

One "step" = predict + Morton key + radix sort + reorder + cell table, 4 x
(lambda, delta), finalise, XSPH, then density-field sampling, case
classification + count, device scan and triangle emission. Inputs are
resident in HBM before the timed region; nothing crosses PCIe inside it.

Prints one JSON line. The warm-up steps time every launch with HIP events on
the library's own stream and name the kernel with the largest share of device
time; inside the timed steps only that kernel's launches carry events (events
around every launch cost ~8 % of a step), and the roofline entry prices its
mean launch time with its algorithmic bytes (the byte model below). The
per-kernel table comes from the same steps repeated from the same start with
every launch timed. The CPU baseline is the
CPU restatement (oracle/, FP32, OpenMP on the host cores) on the first steps
of the same scene.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(name, n, G, S, T):
    """Compulsory HBM bytes of one launch, the per-kernel table of the path's
    measurement definition: each input array read once and each output written
    once; vec3 = 12 B, scalars and keys 4 B, type 1 B. n particles, G
    cell-table entries, S lattice samples, T triangles.

    The grid build's kernels (predict_count, the bucket scan, scatter,
    large-bucket sort, reorder) together replace key + sort + reorder + cell
    start, priced as a whole at 20 n + 16 n + 88 n + (4 n + 4 G); each kernel's
    share below is its own compulsory traffic."""
    cubes = S  # (S-1)^3 ~ S^3 at 256^3
    table = {
        "predict_count": 20 * n + 4 * n,        # p* (from x, v) -> key, rank; counts
        "bucket_reduce": 4 * G,                 # counts
        "bucket_scan": 4 * G + 4 * G,           # counts -> table
        "bucket_scatter": 8 * n + 8 * n,        # key, rank -> key, index
        "reorder": 88 * n,                      # state in order, p* and keys out
        "pbf_lambda": 21 * n + 4 * G,           # p*, mass/type, key, table -> lambda
        "pbf_lambda_kahan": 21 * n + 4 * G,
        "pbf_lambda_fast": 21 * n + 4 * G,
        "pbf_lambda_kahan_fast": 21 * n + 4 * G,
        "pbf_delta": 29 * n + 4 * G,            # p*, lambda, key, table -> p*
        "pbf_delta_fast": 29 * n + 4 * G,
        # the last delta pass of a step with finalise fused in: + finalise's 60 n
        "pbf_delta_final": 29 * n + 4 * G + 60 * n,
        "pbf_delta_final_fast": 29 * n + 4 * G + 60 * n,
        "pbf_finalise": 60 * n,
        "pbf_xsph": 37 * n + 4 * G,
        "pbf_diffuse": 17 * n + 4 * G,
        "mc_density_field": 4 * S + 16 * n + 4 * G,  # field out; particles + table in
        "mc_classify_count": 4 * S + cubes,     # field in, one case byte per cube out
        "mc_emit": cubes + 36 * T,              # case bytes in, triangle positions out
    }
    return table.get(name)


GRID_BUILD = ("predict_count", "bucket_reduce", "bucket_scan", "bucket_scatter", "sort_big",
              "reorder")


def grid_build_bytes(n, G):
    """Key + one logical sort permutation + reorder + cell start."""
    return 20 * n + 16 * n + 88 * n + 4 * n + 4 * G


# The kernel behind each launch name of the library: a rocprof name prefix, or
# a pattern of the full name where template arguments tell launches apart.
KERNEL_SYMBOL = {
    "predict_count": "k_predict_count", "bucket_reduce": "k_bucket_reduce",
    "bucket_scan": "k_bucket_scan", "bucket_scatter": "k_bucket_scatter",
    "sort_big": "k_sort_big", "reorder": "k_reorder", "pbf_build_list": "k_build_list",
    # k_lambda<threads, KAHAN, arith, PACKED16, tiles>, k_delta<threads, arith, FINAL, PACKED16,
    # tiles>; arith 0: IEEE sequences, 1: their shortened correctly rounded forms, 2: FAST_MATH
    "pbf_lambda": r"k_lambda<\d+, false, [01], false, false>",
    "pbf_lambda_kahan": r"k_lambda<\d+, true, [01], false, false>",
    "pbf_delta": r"k_delta<\d+, [01], false, false, false>",
    "pbf_delta_final": r"k_delta<\d+, [01], true, false, false>",
    "pbf_finalise": "k_finalise", "pbf_xsph": "k_xsph",
    "pbf_diffuse": "k_diffuse", "mc_density_field": "k_mc_density", "mc_classify_count": "k_mc_count",
    "mc_scan": "k_mc_scan", "mc_emit": "k_mc_emit", "copy_probe": "k_copy_probe",
}
# The context stream each launch runs on: the density-field marching cubes on
# their own stream (beside the next step), the rest on the main stream.
KERNEL_STREAM = {"mc_density_field": "mc", "mc_classify_count": "mc", "mc_scan": "mc",
                 "mc_emit": "mc"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_dam_break_1m.txt")


def measured_traffic():
    """HBM bytes per dispatch by kernel symbol, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (already scaled by the copy-kernel
    calibration); {} when absent."""
    out = {}
    if not os.path.exists(TRAFFIC_FILE):
        return out
    for line in open(TRAFFIC_FILE):
        f = line.split("|")
        if len(f) < 6 or not f[1].strip().startswith("k_"):
            continue
        try:
            out[f[1].strip()] = int(f[4].strip())
        except ValueError:
            pass
    return out


def summarize_pmc(fetch_dir, write_dir, copy_bytes=1 << 28):
    """Per-kernel mean FETCH_SIZE / WRITE_SIZE per dispatch (KiB) from two
    rocprofv3 --pmc runs, scaled so the copy-probe kernel (copy_bytes read,
    copy_bytes written per dispatch) reads exactly its size. Prints a
    markdown table: kernel | fetch KiB | write KiB | HBM bytes | factors."""
    import csv
    import glob
    import re

    def load(d, counter):
        acc = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                if r["Counter_Name"] != counter:
                    continue
                m = re.search(r"(k_[a-z_0-9]+(<[^>]*>)?)\(", r["Kernel_Name"])
                if not m:
                    continue
                acc.setdefault(m.group(1), []).append(float(r["Counter_Value"]))
        return {k: sum(v) / len(v) for k, v in acc.items()}

    fetch, write = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    kib = copy_bytes / 1024.0
    ff = kib / fetch["k_copy_probe"] if fetch.get("k_copy_probe") else 1.0
    fw = kib / write["k_copy_probe"] if write.get("k_copy_probe") else 1.0
    print("| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes per dispatch | scale |")
    print("|---|---:|---:|---:|---|")
    for k in sorted(set(fetch) | set(write)):
        fb, wb = fetch.get(k, 0.0), write.get(k, 0.0)
        total = int(round((fb * ff + wb * fw) * 1024))
        print("| %s | %.0f | %.0f | %d | fetch x%.3f, write x%.3f |" % (k, fb, wb, total, ff, fw))


SQ_COUNTERS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CYCLES",
               "SQ_WAVE_CYCLES", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT", "SQ_INSTS_VMEM_RD")


def summarize_sq(sq_dir, trace_csv, cus=256, simds=4, ghz=2.4):
    """Per-kernel means per dispatch of a rocprofv3 --pmc run of SQ_COUNTERS,
    with the kernel's mean duration from a kernel trace of the same command:
    `VALU time` = SQ_INSTS_VALU x 4 cycles (a wave64 VALU instruction occupies
    a SIMD16 for 4 cycles) / (cus x simds x ghz), the time the kernel would take
    if VALU issue were its only limit, and `VALU share` = that time / the mean
    duration. Prints a markdown table."""
    import csv
    import glob
    import re
    acc = {}
    for f in glob.glob(os.path.join(sq_dir, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            m = re.search(r"(k_[a-z_0-9]+(<[^>]*>)?)\(", r["Kernel_Name"])
            if m:
                acc.setdefault(m.group(1), {}).setdefault(r["Counter_Name"], []).append(
                    float(r["Counter_Value"]))
    dur = {}
    for name, a, b in load_trace(trace_csv):
        m = re.search(r"(k_[a-z_0-9]+(<[^>]*>)?)\(", name)
        if m:
            dur.setdefault(m.group(1), []).append((b - a) / 1000.0)
    print("| kernel | waves | VALU / wave | VMEM rd / wave | LDS / wave | LDS bank conflict cycles "
          "| VALU time us | trace avg us | VALU share |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|---:|")
    for k in sorted(acc):
        c = {n: sum(v) / len(v) for n, v in acc[k].items()}
        waves = c.get("SQ_WAVES", 0.0)
        per = lambda n: c.get(n, 0.0) / waves if waves else 0.0  # noqa: E731
        valu_us = c.get("SQ_INSTS_VALU", 0.0) * 4.0 / (cus * simds * ghz * 1e3)
        d = dur.get(k)
        avg = sum(d) / len(d) if d else None
        print("| %s | %.0f | %.0f | %.1f | %.1f | %.0f | %.1f | %s | %s |" % (
            k, waves, per("SQ_INSTS_VALU"), per("SQ_INSTS_VMEM_RD"), per("SQ_INSTS_LDS"),
            c.get("SQ_LDS_BANK_CONFLICT", 0.0), valu_us, "%.1f" % avg if avg else "-",
            "%.0f%%" % (100.0 * valu_us / avg) if avg else "-"))


def load_trace(path):
    """(name, start_ns, end_ns) of every kernel dispatch of a rocprofv3
    --kernel-trace run, in start order: the SQLite results database (ROCm 7's
    default output) or a kernel_trace.csv."""
    if path.endswith(".csv"):
        import csv
        rows = [(r["Kernel_Name"], int(r["Start_Timestamp"]), int(r["End_Timestamp"]))
                for r in csv.DictReader(open(path))]
    else:
        import sqlite3
        db = sqlite3.connect(path)
        rows = db.execute("select name, start, end from kernels").fetchall()
    return sorted(rows, key=lambda r: r[1])


def summarize_trace(path, steps, warmup, iterations=4):
    """Per-kernel launch statistics of a rocprofv3 kernel trace of
    `bench.py --steps K --warmup W --no-cpu --batch-scenes 0 --no-combined`, as a
    markdown table: over every launch of the process, and over the timed window
    alone (the K steps after the W warm-up steps of the second pass over the
    scene, the steps bench.py's value and roofline are measured on). Steps are
    counted by the launches of k_predict_count (one per step)."""
    rows = load_trace(path)
    # "void lf::(anonymous namespace)::k_delta<256, ...>(lf::Work, ...)" -> "lf::k_delta<256, ...>"
    sym = lambda name: re.sub(r"\(.*", "", re.sub(r"^void ", "", name.replace(  # noqa: E731
        "(anonymous namespace)::", "")))
    starts = [r[1] for r in rows if sym(r[0]).startswith("lf::k_predict_count")]
    per_pass = warmup + steps
    t0 = starts[per_pass + warmup] if len(starts) > per_pass + warmup else None
    t1 = starts[2 * per_pass] if len(starts) > 2 * per_pass else (rows[-1][2] + 1)

    def stats(sel):
        acc = {}
        for name, a, b in sel:
            k = sym(name).replace("lf::", "")
            acc.setdefault(k, []).append((b - a) / 1000.0)
        return acc

    every = stats(rows)
    window = stats([r for r in rows if t0 is not None and t0 <= r[1] < t1])
    print("| kernel | launches (all) | mean us (all) | launches (timed steps) | mean us (timed steps) "
          "| us per timed step |")
    print("|---|---:|---:|---:|---:|---:|")
    for k in sorted(every, key=lambda k: -sum(window.get(k, every[k]))):
        a, w = every[k], window.get(k, [])
        print("| %s | %d | %.2f | %d | %s | %s |" % (
            k, len(a), sum(a) / len(a), len(w), "%.2f" % (sum(w) / len(w)) if w else "-",
            "%.1f" % (sum(w) / steps) if w else "-"))
    if t0 is not None:
        print("\ntimed window: %d steps, %.3f ms of trace time (%.3f ms per step)" % (
            steps, (t1 - t0) / 1e6, (t1 - t0) / 1e6 / steps))


def cpu_baseline(scene, mc, budget_s=20.0):
    """The CPU restatement (FP32, OpenMP) on the first steps of the scene,
    marching cubes included; returns (particle-steps/s, steps, threads)."""
    import oracle
    p = scene.params()
    o = oracle.Pbf(p)
    o.load(scene.position, scene.velocity, scene.mass, scene.type, scene.colour)
    steps = 0
    t = 0.0
    while steps == 0 or (t < budget_s and steps < scene.steps):
        t0 = time.perf_counter()
        o.step()
        field = o.mc_density(mc.samples)
        origin, step = o.density_lattice(mc.samples)
        oracle.mc_emit(field, mc.isolevel, origin, step, scene.scale)
        t += time.perf_counter() - t0
        steps += 1
        print("bench: cpu baseline step %d: %.2f s" % (steps, t), file=sys.stderr, flush=True)
        if t > budget_s / 2 and steps >= 1:
            break
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return scene.n * steps / t, steps, threads


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


MIXED_FLAGS = None  # set in main: FAST_MATH | KAHAN


def _batched_ctx(lf, sc, flags, log, tag):
    ctx = lf.Context(0)
    ctx.batch_upload([s.params(flags=flags) for s in sc],
                     [dict(position=s.position, velocity=s.velocity, mass=s.mass, type=s.type,
                           colour=s.colour) for s in sc])
    ctx.batch_step(1)
    tune = ctx.autotune(3)
    log("%s: autotune %s" % (tag, tune))
    ctx.batch_step(1)
    ctx.graph_capture(2)
    return ctx, tune


def _time_graph(ctx, steps, warmup):
    ctx.graph_launch(max(1, warmup // 2))
    ctx.sync()
    reps = max(1, steps // 2)
    t0 = time.perf_counter()
    ctx.graph_launch(reps)
    ctx.sync()
    return 2 * reps, time.perf_counter() - t0


def bench_batched(n_scenes, steps, warmup, seed=3, log=None):
    """Aggregate particle-steps/s of a batch of uneven scenes, three ways:
    plain = one context (own stream) per scene, FP32, stepped scene after
    scene (one solver per scene, as libfluid runs); batched = all scenes in one
    context, autotuned, whole steps replayed from a HIP graph, FP32; batched
    mixed = the same with FAST_MATH + KAHAN."""
    import libfluid_amd as lf
    from libfluid_amd import scenes
    sc = scenes.batch_scenes(n_scenes, seed=seed)
    total = sum(s.n for s in sc)
    log("batched: %d scenes, %d particles" % (len(sc), total))
    ctxs = []
    for s in sc:
        c = lf.Context(0)
        c.upload(s.position, s.velocity, s.mass, s.type, s.colour)
        ctxs.append((c, s.params()))
    for _ in range(warmup):
        for c, p in ctxs:
            c.step(p, 1)
    for c, _ in ctxs:
        c.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        for c, p in ctxs:
            c.step(p, 1)
    for c, _ in ctxs:
        c.sync()
    plain_s = time.perf_counter() - t0
    for c, _ in ctxs:
        c.close()
    log("batched: plain per-scene %.3f s" % plain_s)
    out = {"workload": "%d scenes, %d particles (log-uniform %d..%d, seed %d), %d iterations"
                       % (len(sc), total, min(s.n for s in sc), max(s.n for s in sc), seed,
                          sc[0].iterations),
           "unit": "particle-steps/s", "steps": steps,
           "plain_per_scene_fp32": total * steps / plain_s}
    for tag, flags in (("batched_graph_fp32", 0), ("batched_graph_mixed", MIXED_FLAGS),
                       ("batched_graph_mixed_packed16", MIXED_FLAGS | lf.LF_FLAG_PACKED16)):
        ctx, tune = _batched_ctx(lf, sc, flags, log, tag)
        k, sec = _time_graph(ctx, steps, warmup)
        ctx.close()
        out[tag] = total * k / sec
        out[tag + "_autotune"] = tune
        log("%s: %d steps %.3f s" % (tag, k, sec))
    out["speedup_batched_fp32"] = out["batched_graph_fp32"] / out["plain_per_scene_fp32"]
    out["speedup_batched_mixed_vs_plain_fp32"] = (out["batched_graph_mixed"]
                                                  / out["plain_per_scene_fp32"])
    out["speedup_batched_mixed_packed16_vs_plain_fp32"] = (out["batched_graph_mixed_packed16"]
                                                           / out["plain_per_scene_fp32"])
    return out


def bench_combined(steps, mc_steps, log):
    """32 scenes totalling ~8M particles, mixed precision, 6 iterations, graph
    replayed steps; then batched marching cubes at 512^3 per scene (FP32
    density field, positions only) over every scene: particle-steps/s and
    triangles/s."""
    import libfluid_amd as lf
    from libfluid_amd import scenes
    sc = scenes.combined_scenes()
    total = sum(s.n for s in sc)
    log("combined: %d scenes, %d particles" % (len(sc), total))
    ctx, tune = _batched_ctx(lf, sc, MIXED_FLAGS, log, "combined")
    k, sec = _time_graph(ctx, steps, 2)
    log("combined: %d steps %.3f s" % (k, sec))
    ctx.mc_set_capacity(48 << 20)
    m = sc[0].mc(attributes=0)
    tris = 0
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(mc_steps):
        for s in range(len(sc)):
            ctx.batch_select(s)
            ctx.mc_sample(None, m)
            ctx.mc_classify_count(m)
            ctx.mc_emit(None, m)
            tris += ctx.mc_count()
    ctx.sync()
    mc_sec = time.perf_counter() - t0
    ctx.close()
    log("combined: marching cubes %d x %d scenes %.3f s, %d triangles" % (mc_steps, len(sc),
                                                                        mc_sec, tris))
    return {"workload": "%d scenes, %d particles (log-uniform 50k..500k rescaled to ~8M, seed 5), "
                        "FAST_MATH+KAHAN, 6 iterations; marching cubes 512^3 per scene"
                        % (len(sc), total),
            "particle_steps_per_s": total * k / sec, "steps": k, "autotune": tune,
            "mc_passes": mc_steps, "mc_samples_per_s": mc_steps * len(sc) * 512 ** 3 / mc_sec,
            "triangles_per_s": tris / mc_sec, "triangles_per_pass": tris // mc_steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--scene", default="dam_break_1m")
    ap.add_argument("--batch-scenes", type=int, default=64,
                    help="scenes of the batched comparison (0: skip it)")
    ap.add_argument("--no-combined", action="store_true",
                    help="skip the 32-scene ~8M combined case")
    ap.add_argument("--skin", type=float, default=None,
                    help="neighbour-list skin (fraction of h) of the 1M runs (default: the "
                         "library's, 0.15)")
    ap.add_argument("--persistent", action="store_true",
                    help="work-queue launches for the pair kernels of the 1M runs")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="print the per-kernel HBM traffic table of two rocprofv3 --pmc runs")
    ap.add_argument("--summarize-sq", nargs=2, metavar=("SQ_DIR", "TRACE_CSV"),
                    help="print the per-kernel SQ counter table of a rocprofv3 --pmc run with "
                         "the durations of a kernel trace of the same command")
    ap.add_argument("--summarize-trace", metavar="RESULTS",
                    help="print the per-kernel table of a rocprofv3 --kernel-trace run of this "
                         "script (results .db or kernel_trace.csv) over the --steps/--warmup window")
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc)
        return
    if args.summarize_sq:
        summarize_sq(*args.summarize_sq)
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, args.steps, args.warmup)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import libfluid_amd as lf
    from libfluid_amd import scenes
    global MIXED_FLAGS
    MIXED_FLAGS = lf.LF_FLAG_FAST_MATH | lf.LF_FLAG_KAHAN

    scene = getattr(scenes, args.scene)()
    p = scene.params()
    mc = scene.mc()
    n = scene.n
    print("bench: %s n=%d steps=%d warmup=%d" % (scene.name, n, args.steps, args.warmup),
          file=sys.stderr, flush=True)

    ctx = lf.Context(0)
    if args.skin is not None or args.persistent:
        ctx.exec_config(256, args.persistent, 0.15 if args.skin is None else args.skin)
    ctx.upload(scene.position, scene.velocity, scene.mass, scene.type, scene.colour)

    def one_step():
        ctx.step(p, 1)
        ctx.mc_sample(p, mc)
        ctx.mc_classify_count(mc)
        ctx.mc_emit(p, mc)

    # pass 1: warm-up, then the K steps with every launch timed: the per-kernel
    # table, the device time per step, and the kernel with the largest total
    for _ in range(args.warmup):
        one_step()
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(args.steps):
        one_step()
    ctx.sync()
    stats = ctx.kernel_stats()
    # the dominant kernel: the largest total time on the main stream, the
    # step's critical path (the marching cubes run on the context's other
    # stream beside it, and their event-timed durations include the time they
    # share the GPU)
    main = {k: v for k, v in stats.items() if KERNEL_STREAM.get(k, "main") == "main"}
    dom_name = max(main.items(), key=lambda kv: kv[1][1])[0] if main else None
    print("bench: profiled pass done, dominant kernel %s" % dom_name, file=sys.stderr, flush=True)

    # pass 2, timed: the same steps from the same start, HIP events around the
    # dominant kernel's launches only (events on every launch cost ~8 % of a step)
    ctx.profile(False)
    ctx.upload(scene.position, scene.velocity, scene.mass, scene.type, scene.colour)
    for _ in range(args.warmup):
        one_step()
    ctx.profile_only(dom_name)
    ctx.profile(True)
    ctx.profile_reset()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_step()
    ctx.sync()
    elapsed = time.perf_counter() - t0
    print("bench: %d timed steps in %.3f s" % (args.steps, elapsed), file=sys.stderr, flush=True)
    dom_stats = ctx.kernel_stats()
    ntri = ctx.mc_count()
    ctx.profile_only(None)
    ctx.profile(False)
    G = ctx.grid_info()[0]
    S = int(np.prod(mc.samples))

    traffic = measured_traffic()

    def traffic_of(name):
        sym = KERNEL_SYMBOL.get(name)
        if not sym:
            return None
        if "<" in sym:
            hits = [v for k, v in traffic.items() if re.fullmatch(sym, k)]
        else:
            hits = [v for k, v in traffic.items() if k == sym or k.startswith(sym + "<")]
        return hits[0] if hits else None

    value = n * args.steps / elapsed
    ms_per_step = 1000.0 * elapsed / args.steps
    device_ms = sum(v[1] for v in stats.values())
    dom_count, dom_ms = dom_stats[dom_name]
    dom_avg_s = dom_ms / dom_count / 1000.0
    dom_bytes = algorithmic_bytes(dom_name, n, G, S, ntri)
    achieved = dom_bytes / dom_avg_s / 1e9 if dom_bytes else None
    kernels = {}
    for name, (cnt, ms) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        b = algorithmic_bytes(name, n, G, S, ntri)
        avg = ms / cnt / 1000.0
        kernels[name] = {"launches": cnt, "avg_us": round(avg * 1e6, 2),
                         "stream": KERNEL_STREAM.get(name, "main"),
                         "share": round(ms / device_ms, 4),
                         "alg_GBs": round(b / avg / 1e9, 1) if b else None,
                         "hbm_GBs": (round(traffic_of(name) / avg / 1e9, 1)
                                     if traffic_of(name) else None)}

    out = {
        "metric": "PBF particle-steps/s @1M, 4 iters (% HBM peak); batched-mixed vs plain FP32",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded jittered lattice dam break)",
        "config": {"workload": "%s: plain FP32 PBF dam break, %d particles (a 128x128x64 "
                               "jittered lattice: the 1,048,576 count kept, a 128^3 lattice would "
                               "hold 2,097,152), world spacing %g, world h %g (scale %g; solver "
                               "h %g), %gx%gx%g box, %d iterations, XSPH %g, dt 1/120, marching "
                               "cubes on the %dx%dx%d density field every step; steps %d-%d of "
                               "the simulation timed"
                               % (scene.name, n, scene.spacing, scene.h * scene.scale,
                                  scene.scale, scene.h, *scene.max_bound, scene.iterations,
                                  scene.xsph_c, *mc.samples, args.warmup,
                                  args.warmup + args.steps),
                   "particles": n, "iterations": scene.iterations, "scale": scene.scale,
                   "world_h": scene.h * scene.scale, "world_spacing": scene.spacing,
                   "box": list(scene.max_bound), "mc_samples": list(mc.samples),
                   "triangles_per_step": ntri},
        "roofline": {"bound": "hbm", "kernel": dom_name,
                     "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": traffic_of(dom_name), "alg_bytes_per_launch": dom_bytes,
                     "byte_model": "per-launch compulsory bytes of the path's measurement table "
                                   "(lambda: 21 n + 4 G)",
                     "avg_launch_us": round(dom_avg_s * 1e6, 2),
                     "window": "the %d timed steps (simulation steps %d-%d), HIP events on the "
                               "library's stream" % (args.steps, args.warmup,
                                                     args.warmup + args.steps),
                     "selection": "largest total time among the main-stream kernels (the step's "
                                  "critical path; the other streams run beside it)"},
        "kernels": kernels,
        "kernels_pass": "the timed steps run beforehand from the same start with every launch timed",
        "grid_build": {
            "kernels": [k for k in GRID_BUILD if k in stats],
            "us_per_step": round(sum(stats[k][1] for k in GRID_BUILD if k in stats)
                                 / args.steps * 1000.0, 2),
            "alg_bytes": grid_build_bytes(n, G),
            "alg_GBs": round(grid_build_bytes(n, G) / (sum(stats[k][1] for k in GRID_BUILD
                                                            if k in stats) / args.steps / 1e3)
                             / 1e9, 1) if any(k in stats for k in GRID_BUILD) else None},
        "device_ms_per_step": device_ms / args.steps,
        "main_stream_ms_per_step": sum(v[1] for k, v in stats.items()
                                       if KERNEL_STREAM.get(k, "main") == "main") / args.steps,
    }
    log = lambda m: print("bench: " + m, file=sys.stderr, flush=True)  # noqa: E731
    # the same 1M workload with the mixed-precision options (and with the
    # 16-bit cell-relative neighbour positions)
    for key, flags, tag in (("mixed_1m", MIXED_FLAGS, "FAST_MATH|KAHAN"),
                            ("mixed_packed16_1m", MIXED_FLAGS | lf.LF_FLAG_PACKED16,
                             "FAST_MATH|KAHAN|PACKED16")):
        pm = scene.params(flags=flags)
        ctx.upload(scene.position, scene.velocity, scene.mass, scene.type, scene.colour)
        ctx.profile(False)
        for _ in range(args.warmup):
            ctx.step(pm, 1)
            ctx.mc_sample(pm, mc)
            ctx.mc_classify_count(mc)
            ctx.mc_emit(pm, mc)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.step(pm, 1)
            ctx.mc_sample(pm, mc)
            ctx.mc_classify_count(mc)
            ctx.mc_emit(pm, mc)
        ctx.sync()
        mixed_s = time.perf_counter() - t0
        out[key] = {"flags": tag, "value": n * args.steps / mixed_s,
                    "ms_per_step": 1000.0 * mixed_s / args.steps}
        log("%s: %.3f ms/step" % (key, 1000.0 * mixed_s / args.steps))
    # the same plain FP32 steps without the marching cubes
    ctx.upload(scene.position, scene.velocity, scene.mass, scene.type, scene.colour)
    for _ in range(args.warmup):
        ctx.step(p, 1)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(p, args.steps)
    ctx.sync()
    nomc_s = time.perf_counter() - t0
    out["no_mc_1m"] = {"value": n * args.steps / nomc_s, "ms_per_step": 1000.0 * nomc_s / args.steps,
                       "note": "plain FP32, the same steps without the marching cubes"}
    log("no_mc_1m: %.3f ms/step" % (1000.0 * nomc_s / args.steps))
    if args.batch_scenes > 0:
        out["batched"] = bench_batched(args.batch_scenes, max(20, args.steps), args.warmup,
                                       log=log)
    if not args.no_combined:
        out["combined"] = bench_combined(max(20, args.steps // 5), 1, log)
    copy_bytes = 1 << 28
    copy_ms = ctx.copy_probe(copy_bytes, reps=20)
    out["copy_probe"] = {"bytes_moved": 2 * copy_bytes, "ms": copy_ms,
                         "GBs": round(2 * copy_bytes / (copy_ms / 1e3) / 1e9, 1)}
    ctx.close()
    if not args.no_cpu:
        cv, csteps, threads = cpu_baseline(scene, mc)
        out["cpu_baseline"] = {
            "value": cv, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "first %d step(s) of the same %d-particle scene incl. 256^3 marching "
                      "cubes, FP32 CPU restatement, OpenMP" % (csteps, n),
            "cpu": cpu_model(), "host_threads": os.cpu_count()}
        out["gpu_over_cpu"] = value / cv
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
