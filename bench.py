"""Benchmark of the CURFIL hot path on one MI355X.

Headline workload (batched level-wise training): 800 synthetic RGB-D frames
of 640x480 (30 Voronoi sites, 4 classes + void, 8% missing depth), 5 trees
trained together to depth 20, 4000 class-uniform samples per image and tree
(void ignored), 4000 candidate features (offsets <= 120 px, regions <= 20
px) x 32 thresholds, minSampleCount 50.

A step is one full training of the 5-tree forest (feature generation, split
evaluation and partition of every level) from the resident integral images
and the sampled root pixels; `value` is feature-threshold evaluations per
second of the batched level-wise trainer.  Also reported: the plain
full-frame workload (3 trees to depth 18, F = 2000 x T = 20) on the same
frames, the plain per-node path against the batched one on the headline
configuration (tree 0, depth 12), dense prediction (the trained forest, and
10 complete trees of depth 20 on 64-frame batches) with its roofline, the
serving loop under Poisson arrivals, the roofline of the split-evaluation
kernel, the CPU oracle timed on bounded samples of the same workloads, and
the file paths around the hot path (PNG datasets, tree files, the
curfil_predict tool) on a few of the frames.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
BYTES_PER_RESPONSE = 48.0      # 2 boxes x 4 corners: colour 4 B, depth 8 B (sum + count); 50/50 mix


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=800)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=120000)
    ap.add_argument("--no-extras", action="store_true",
                    help="only the training steps (no plain-vs-batched or prediction legs); for kernel profiles")
    ap.add_argument("--graph-step", action="store_true",
                    help="with --no-extras: also a training step with level graphs (CURFIL_LEVEL_GRAPH=1, profiling off)")
    ap.add_argument("--serve-seconds", type=float, default=6.0,
                    help="arrival window per serving rate (the plain one-frame path takes the first 2 s of it)")
    ap.add_argument("--io-frames", type=int, default=32, help="frames written as PNG files for the I/O leg")
    ap.add_argument("--dense-only", action="store_true",
                    help="only the dense-prediction leg (10 trees of depth 20, 64-frame batches); for kernel profiles")
    ap.add_argument("--predict-only", action="store_true",
                    help="only the dense-prediction and serving legs")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the serving loop (no dense timings, no plain one-frame path); for tuning the loop")
    ap.add_argument("--serve-rates", default="1000,2000,3000,4000",
                    help="with --serve-only: offered rates (frames/s) of the mix; the 640x480 stream runs those of "
                         "2000 and 3000 among them")
    ap.add_argument("--plain-full", action="store_true",
                    help="with --no-extras: also grow the whole batched workload (5 trees to depth 20) with the plain "
                         "per-node path (part of the default run; about 100 s)")
    ap.add_argument("--no-plain-full", action="store_true",
                    help="skip the plain per-node growth of the whole batched workload")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only copy a device buffer of this many MiB 3 times (scale for FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration_mb:
        import torch
        n = args.copy_calibration_mb << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {n} bytes 3 times")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); multi-GPU runs are not part of "
              f"this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import curfil_amd as cf

    if args.serve_only:
        ctx = cf.Context(0)
        forest10 = dense_forest(ctx)
        rates = tuple(int(r) for r in args.serve_rates.split(","))
        serving = serving_leg(ctx, forest10, args.serve_seconds, plain=False, rates=rates)
        rates640 = tuple(r for r in rates if r in (2000, 3000))
        serving640 = (serving_leg(ctx, forest10, args.serve_seconds, plain=False, mix=(0, 1, 0), rates=rates640)
                      if rates640 else None)
        print(json.dumps({"serving": serving, "serving_640x480": serving640}), flush=True)
        return
    if args.dense_only or args.predict_only:
        ctx = cf.Context(0)
        predict10, forest10 = dense_prediction_leg(ctx)
        serving = serving640 = None
        if args.predict_only:
            serving = serving_leg(ctx, forest10, args.serve_seconds)
            serving640 = serving_leg(ctx, forest10, args.serve_seconds, mix=(0, 1, 0), rates=(2000, 3000))
        print(json.dumps({"predict_forest10": predict10, "serving": serving, "serving_640x480": serving640}),
              flush=True)
        return

    W, H, N = 640, 480, args.frames
    base_kw = dict(randomSeed=4711, minSampleCount=50, boxRadius=120, regionSize=20, ignored_labels=(0,))
    # the batched workload: 5 trees to depth 20, F = 4000 x T = 32 (samplesPerImage counts all trees)
    trees = 5
    cfg_kw = dict(base_kw, samplesPerImage=4000 * trees, featureCount=4000, thresholds=32, maxDepth=20)
    # the plain full-frame workload: 3 trees to depth 18, F = 2000 x T = 20
    full_kw = dict(base_kw, samplesPerImage=4000 * 3, featureCount=2000, thresholds=20, maxDepth=18)
    ctx = cf.Context(0)

    t0 = time.time()
    rgb, depth, labels = cf.synth_dataset(N, W, H, sites=30, classes=4, zero_frac=0.08, seed0=4711)
    log(f"synthesised {N} frames {W}x{H} in {time.time() - t0:.1f}s")
    t0 = time.time()
    images = cf.Images(ctx, N, W, H)
    for i in range(N):
        images.set_rgbd_u8(i, rgb[i], depth[i])
        images.set_labels(i, labels[i])
    ctx.synchronize()
    log(f"uploaded + integrated {N} frames in {time.time() - t0:.1f}s")

    cfg = cf.TrainingConfiguration(**cfg_kw)
    t0 = time.time()
    forest = cf.RandomForestImage.sample(ctx, images, cfg, trees, num_labels=5)
    roots = [len(forest.root_samples(t)) for t in range(trees)]
    log(f"sampled root pixels {roots} in {time.time() - t0:.1f}s")

    for w in range(args.warmup):
        t0 = time.time()
        forest.grow(cfg, cf.MODE_BATCHED)
        log(f"warmup step {w}: {time.time() - t0:.2f}s")

    ctx.set_profiling(True)
    step_times, profiles = [], []
    for k in range(args.steps):
        ctx.synchronize()
        t0 = time.perf_counter()
        forest.grow(cfg, cf.MODE_BATCHED)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        step_times.append(dt)
        profiles.append(forest.profile())
        log(f"step {k}: {dt:.3f}s  {profiles[-1]}")
    ctx.set_profiling(False)
    prof = profiles[-1]
    evals = prof["evaluations"]
    total_time = sum(step_times)
    value = evals * len(step_times) / total_time
    nodes = [len(forest.tree(t)[0]) for t in range(trees)]

    roofline = eval_roofline(profiles, N, W, H, cfg_kw["thresholds"])
    shapes = cf.eval_shapes()
    tuning = {"shapes": [dict(zip(("lanes", "samples_per_lane", "features_per_step", "workgroups_per_cu"), s))
                         for s in shapes],
              "table": cf.eval_tuning(), "shape_used": int(prof["eval_shape"]),
              "note": "the tuner times every shape on the first pass-1 launches of each (threshold width, level "
                      "size class) key (the warm-up step here) and keeps the fastest; shape 0 is the default"}
    log(f"tuning: {tuning['table']} (shape used {tuning['shape_used']})")

    plain_full = None
    if (args.plain_full or not args.no_extras) and not args.no_plain_full:
        ctx.synchronize()
        t0 = time.perf_counter()
        forest.grow(cfg, cf.MODE_PLAIN)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        pp = forest.profile()
        plain_full = {"workload": "the batched workload grown with the plain path: one node per launch and host "
                                  "sync per node (the reference GPU path's shape)",
                      "seconds": dt, "evals_per_s": pp["evaluations"] / dt, "nodes": int(pp["nodes"]),
                      "batched_evals_per_s": value, "speedup": value / (pp["evaluations"] / dt),
                      "same_node_counts": [len(forest.tree(t)[0]) for t in range(trees)] == nodes}
        log(f"plain per-node path on the whole workload: {dt:.1f}s, {plain_full['evals_per_s']:.3e} evals/s, "
            f"batched is {plain_full['speedup']:.2f}x")

    full_frame = plain_vs_batched = predict = predict10 = serving = serving640 = io = graph_step = None
    if not args.no_extras:
        full_frame = full_frame_leg(ctx, images, full_kw, N, W, H)
        plain_vs_batched, predict = extra_legs(ctx, images, forest, cfg_kw, N, W, H, trees)
        predict10, forest10 = dense_prediction_leg(ctx)
        serving = serving_leg(ctx, forest10, args.serve_seconds)
        serving640 = serving_leg(ctx, forest10, args.serve_seconds, mix=(0, 1, 0), rates=(2000, 3000))
        io = io_leg(forest, cfg, rgb, depth, labels, args.io_frames)

    # last of the GPU legs: once a graph has been instantiated the runtime holds a hardware queue, which
    # would slow the serving loop's streams created after it
    if not args.no_extras or args.graph_step:
        graph_step = level_graph_step(ctx, forest, cfg, nodes, step_times)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(forest, rgb, depth, labels, cfg_kw, args.cpu_sample,
                           forest10 if not args.no_extras else None)

    out = {
        "metric": "feature-threshold evals/s (train)",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": len(step_times),
        "warmup": args.warmup,
        "ms_per_step": total_time / len(step_times) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"batched level-wise training: {N} frames {W}x{H}, {trees} trees depth 20 "
                               f"trained together, 4000 samples/image/tree, F=4000 x T=32, boxRadius 120, "
                               f"regionSize 20, minSampleCount 50, 4 classes + void, u32 histograms, f64 "
                               f"responses against f32 thresholds",
                   "mode": "batched level-wise", "evaluations_per_step": evals, "nodes_per_tree": nodes,
                   "root_samples_per_tree": roots[0], "levels": int(prof["levels"])},
        "roofline": roofline,
        "tuning": tuning,
        "cpu_baseline": cpu,
        "phase_ms": {k: round(prof[k], 1) for k in ("featgen_ms", "eval_kernel_ms", "hist_ms", "score_ms",
                                                    "partition_ms")},
        "level_graph_step": graph_step,
        "full_frame_training": full_frame,
        "plain_vs_batched": plain_vs_batched,
        "plain_whole_workload": plain_full,
        "predict": predict,
        "predict_forest10": predict10,
        "serving": serving,
        "serving_640x480": serving640,
        "io": io,
    }
    print(json.dumps(out), flush=True)


def level_graph_step(ctx, forest, cfg, nodes, step_times):
    """One more training step with CURFIL_LEVEL_GRAPH=1 and profiling off: every level whose pass-1 shape
    is settled runs as one captured graph (the timed steps launch kernel by kernel, the library's
    default).  Reports its time beside the timed steps' and whether it grew the same trees."""
    import curfil_amd as cf
    old = os.environ.get("CURFIL_LEVEL_GRAPH")
    os.environ["CURFIL_LEVEL_GRAPH"] = "1"
    try:
        ctx.synchronize()
        t0 = time.perf_counter()
        forest.grow(cfg, cf.MODE_BATCHED)
        ctx.synchronize()
        dt = time.perf_counter() - t0
    finally:
        if old is None:
            del os.environ["CURFIL_LEVEL_GRAPH"]
        else:
            os.environ["CURFIL_LEVEL_GRAPH"] = old
    p = forest.profile()
    same = [len(forest.tree(t)[0]) for t in range(len(nodes))] == nodes
    out = {"seconds": dt, "timed_steps_mean_s": sum(step_times) / len(step_times), "level_graphs": int(p["level_graphs"]),
           "levels": int(p["levels"]), "same_node_counts": same}
    log(f"step with level graphs: {dt:.3f}s ({out['level_graphs']} of {out['levels']} levels captured), "
        f"same trees: {same}")
    return out


def eval_roofline(profiles, N, W, H, T):
    """Roofline of the response kernel (pass 1 of the level evaluation: the integral gathers), from its
    HIP-event time on the library's stream: 48 B of corner reads per (sample, feature) response."""
    eval_s = sum(p["eval_kernel_ms"] for p in profiles) / 1e3
    responses = sum(p["responses"] for p in profiles)
    launches = sum(p["eval_launches"] for p in profiles)
    achieved = responses * BYTES_PER_RESPONSE / eval_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic("k_eval_bins", N, W, H, T),
            "kernel": "k_eval_bins", "launches": int(launches),
            "avg_launch_ms": round(sum(p["eval_kernel_ms"] for p in profiles) / max(1, launches), 4),
            "responses_per_launch": responses / max(1, launches),
            "bytes_per_response": BYTES_PER_RESPONSE}


def full_frame_leg(ctx, images, full_kw, N, W, H):
    """The plain full-frame workload on the same frames (3 trees to depth 18, F = 2000 x T = 20), trained
    batched: warm-up steps until the launch-shape tuner has settled this workload's keys (its 85 pass-1
    launches per step are fewer than the tuner's 108), then one timed step."""
    import curfil_amd as cf
    cfg = cf.TrainingConfiguration(**full_kw)
    f = cf.RandomForestImage.sample(ctx, images, cfg, 3, num_labels=5)
    for _ in range(3):
        f.grow(cfg, cf.MODE_BATCHED)
        if all(r["shape"] >= 0 for r in cf.eval_tuning()):
            break
    ctx.set_profiling(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    f.grow(cfg, cf.MODE_BATCHED)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ctx.set_profiling(False)
    p = f.profile()
    out = {"workload": f"{N} frames {W}x{H}, 3 trees depth 18, 4000 samples/image/tree, F=2000 x T=20, "
                       f"minSampleCount 50", "seconds": dt, "evals_per_s": p["evaluations"] / dt,
           "evaluations": p["evaluations"], "nodes_per_tree": [len(f.tree(t)[0]) for t in range(3)],
           "roofline": eval_roofline([p], N, W, H, 20)}
    log(f"full-frame workload (3 trees depth 18, F=2000 T=20): {dt:.2f}s, {out['evals_per_s']:.3e} evals/s, "
        f"k_eval_bins {out['roofline']['avg_launch_ms']:.2f} ms per launch, frac {out['roofline']['frac']}")
    return out


def measured_traffic(kernel, N, W, H, T):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter passes (FETCH_SIZE and
    WRITE_SIZE per dispatch, each scaled by the same counter's reading for a copy kernel of known size),
    for this workload only."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    for e in t.get("entries", [t]):
        w = e.get("workload", {})
        if (w.get("frames"), w.get("width"), w.get("height"), w.get("thresholds")) == (N, W, H, T):
            return e.get("bytes_per_launch")
    return None


def classify_roofline(batch_ms, n, W, H):
    """HBM roofline of the dense-prediction launch.  achieved = its algorithmic bytes per launch (each
    pixel's integral planes read and its outputs written once, 41 B, plus the forest read once) over the
    launch time measured here (the device-output call, which waits for the launch); traffic = its HBM
    bytes per launch (FETCH_SIZE and WRITE_SIZE of the committed counter passes, scaled by the copy
    calibration).  The kernel is bound by the vector L1 (its lookups per CU and cycle, from the same file),
    not by HBM: the fraction is low because the forest's blocks and the frames' planes are served from the
    caches."""
    path = os.path.join(ROOT, "profiles", "k_classify_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    for e in t.get("entries", []):
        w = e.get("workload", {})
        if (w.get("frames"), w.get("width"), w.get("height"), w.get("trees"), w.get("depth")) == (n, W, H, 10, 20):
            alg = e["algorithmic_bytes_per_launch"]
            achieved = alg / (batch_ms * 1e-3) / 1e9
            return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": e["bytes_per_launch"],
                    "algorithmic_bytes": alg, "kernel": "k_classify", "avg_launch_ms": round(batch_ms, 3),
                    "l1_tag_lookups_per_cu_cycle": e.get("l1_tag_lookups_per_cu_cycle"),
                    "note": "achieved = algorithmic bytes (41 B per pixel + the forest once) per launch; traffic = "
                            "FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per launch (profiles/k_classify_traffic.json); "
                            "the launch is bound by the vector L1 (about one tag lookup per CU and cycle), see "
                            "profiles/vector_memory_path_counters.md"}
    return None


IO_PALETTE = [(0, 0, 0), (128, 0, 0), (0, 128, 0), (128, 128, 0), (0, 0, 128)]


def io_leg(forest, cfg, rgb, depth, labels, frames):
    """The file paths around the hot path on the first `frames` frames: the frames written as
    *_colors.png / *_depth.png / *_ground_truth.png, loaded back as a dataset (loadImages), the
    trained forest exported to tree<N>.json.gz and imported again, and curfil_predict run on the
    PNG folder (forest import, PNG decode, integrals, classification, prediction PNGs, metrics)."""
    import shutil
    import tempfile
    import numpy as np
    import curfil_amd as cf

    root = tempfile.mkdtemp(prefix="curfil_io_")
    try:
        test, pred, trees = (os.path.join(root, d) for d in ("test", "pred", "trees"))
        os.makedirs(test)
        os.makedirs(pred)
        pal = np.array(IO_PALETTE, np.uint8)
        t0 = time.perf_counter()
        for i in range(frames):
            cf.write_png(os.path.join(test, f"{i:04d}_colors.png"), rgb[i])
            cf.write_png(os.path.join(test, f"{i:04d}_depth.png"), depth[i])
            cf.write_png(os.path.join(test, f"{i:04d}_ground_truth.png"), pal[labels[i]])
        write_s = time.perf_counter() - t0
        png_bytes = sum(os.path.getsize(os.path.join(test, f)) for f in os.listdir(test))
        t0 = time.perf_counter()
        ds = cf.load_images(test)
        load_s = time.perf_counter() - t0
        assert ds.count == frames
        ds.close()
        t0 = time.perf_counter()
        paths = forest.export(trees, cfg, palette=pal)
        export_s = time.perf_counter() - t0
        tree_bytes = sum(os.path.getsize(p) for p in paths)
        t0 = time.perf_counter()
        back = cf.RandomForestImage.from_files(paths)
        import_s = time.perf_counter() - t0
        nodes = sum(len(back.tree(t)[0]) for t in range(back.num_trees))
        t0 = time.perf_counter()
        rc = cf.run_tool("predict", [pred, test] + paths)
        predict_s = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"curfil_predict exited with {rc}")
        out = {"frames": frames, "png_bytes": png_bytes,
               "png_write_frames_per_s": frames / write_s,
               "load_images_frames_per_s": frames / load_s,
               "load_images_mb_per_s": png_bytes / load_s / 1e6,
               "tree_files": len(paths), "tree_file_bytes": tree_bytes, "tree_nodes": nodes,
               "export_s": export_s, "import_s": import_s,
               "curfil_predict_frames_per_s": frames / predict_s,
               "note": "CPU threads of this host for PNG coding and JSON; curfil_predict end to end, prediction "
                       "PNGs written"}
        log(f"io: load {out['load_images_frames_per_s']:.0f} frames/s ({out['load_images_mb_per_s']:.0f} MB/s of "
            f"PNG), export {export_s:.2f}s import {import_s:.2f}s ({nodes} nodes), curfil_predict "
            f"{out['curfil_predict_frames_per_s']:.0f} frames/s")
        return out
    finally:
        shutil.rmtree(root, ignore_errors=True)


def extra_legs(ctx, images, forest, cfg_kw, N, W, H, trees):
    """Plain per-node vs batched training on one tree, and dense prediction."""
    import numpy as np
    import curfil_amd as cf

    # plain per-node path vs batched on the same single tree (tree 0, depth 12)
    cfg1 = cf.TrainingConfiguration(**dict(cfg_kw, samplesPerImage=4000, maxDepth=12))
    f1 = cf.RandomForestImage.sample(ctx, images, cfg1, 1, num_labels=5)
    t0 = time.perf_counter()
    f1.grow(cfg1, cf.MODE_PLAIN)
    plain_s = time.perf_counter() - t0
    plain_evals = f1.profile()["evaluations"]
    plain_nodes = f1.profile()["nodes"]
    t0 = time.perf_counter()
    f1.grow(cfg1, cf.MODE_BATCHED)
    batched1_s = time.perf_counter() - t0
    log(f"plain (1 tree, depth 12, {int(plain_nodes)} nodes): {plain_s:.2f}s; batched: {batched1_s:.2f}s")

    # the same tree with horizontal flipping (every root sample evaluated in both orientations)
    cfgf = cf.TrainingConfiguration(**dict(cfg_kw, samplesPerImage=4000, maxDepth=12, horizontalFlipping=True))
    ff = cf.RandomForestImage.sample(ctx, images, cfgf, 1, num_labels=5)
    t0 = time.perf_counter()
    ff.grow(cfgf, cf.MODE_BATCHED)
    flip_s = time.perf_counter() - t0
    pf = ff.profile()
    log(f"horizontal flipping (1 tree, depth 12, {int(pf['nodes'])} nodes): {flip_s:.2f}s")

    # dense prediction with the trained forest
    forest.normalizeHistograms(0.0, True)
    batch = 64
    forest.predict_batch(images, list(range(min(batch, N))), probabilities=False)   # warm (forest upload)
    t0 = time.perf_counter()
    for b0 in range(0, N, batch):
        forest.predict_batch(images, list(range(b0, min(N, b0 + batch))), probabilities=False)
    pred_s = time.perf_counter() - t0
    mpix = N * W * H / pred_s / 1e6
    lat = []
    for i in range(min(N, 100)):
        t1 = time.perf_counter()
        forest.predict(images, i, probabilities=True)
        lat.append((time.perf_counter() - t1) * 1e3)
    lat = np.array(lat)
    log(f"predict: {mpix:.1f} Mpix/s batched ({N / pred_s:.1f} frames/s); single-frame p50 "
        f"{np.percentile(lat, 50):.2f} ms p99 {np.percentile(lat, 99):.2f} ms")
    plain_vs_batched = {"workload": "tree 0 only, depth 12", "plain_evals_per_s": plain_evals / plain_s,
                        "batched_evals_per_s": plain_evals / batched1_s,
                        "speedup": plain_s / batched1_s, "nodes": int(plain_nodes),
                        "horizontal_flipping": {"seconds": flip_s, "nodes": int(pf["nodes"]),
                                                "responses_per_s": pf["responses"] / flip_s,
                                                "note": "batched two-pass; responses_per_s counts a sample "
                                                        "once although Both samples are evaluated in both "
                                                        "orientations"}}
    predict = {"frames": N, "trees": trees, "batch": batch, "mpix_per_s": mpix, "frames_per_s": N / pred_s,
               "single_frame_p50_ms": float(np.percentile(lat, 50)),
               "single_frame_p99_ms": float(np.percentile(lat, 99)),
               "note": "batched = labels to host; single frame = posteriors + labels to host"}
    return plain_vs_batched, predict


def dense_forest(ctx, frames=False):
    """10 complete trees of depth 20 with random splits (thresholds = responses at random pixels,
    Dirichlet(1) leaves over 5 classes, seed 4711) drawn on 64 synthetic 640x480 frames (12 Voronoi sites,
    5 classes, 5% missing depth); with frames=True also the frames' images."""
    import numpy as np
    import curfil_amd as cf
    W, H, n = 640, 480, 64
    rgb, depth, _ = cf.synth_dataset(n, W, H, sites=12, classes=5, zero_frac=0.05, seed0=5711)
    im = cf.Images(ctx, n, W, H)
    for i in range(n):
        im.set_rgbd_u8(i, rgb[i], depth[i])
    t0 = time.time()
    forest = cf.synth_forest(ctx, im, depth.astype(np.int32), trees=10, depth=20, classes=5, seed=4711)
    forest.normalizeHistograms(0.0, False)
    forest.predict_batch(im, list(range(n)), probabilities=True)   # forest upload + warm
    log(f"forest of 10 trees x depth 20 built in {time.time() - t0:.1f}s")
    return (forest, im) if frames else forest


def dense_prediction_leg(ctx):
    """Batched dense prediction with the forest of dense_forest on its 64 frames."""
    import numpy as np
    import curfil_amd as cf
    W, H, n = 640, 480, 64
    forest, im = dense_forest(ctx, frames=True)
    out = {"workload": "64 frames 640x480, 10 complete trees of depth 20 (random splits), 5 classes"}
    # outputs to host: into page-locked arrays (DMA at the link's rate), and into ordinary numpy arrays
    pl = cf.pinned_empty(ctx, (n, H, W), np.uint8)
    pp = cf.pinned_empty(ctx, (n, 5, H, W), np.float32)
    for probs in (False, True):
        for pinned in (True, False):
            ts = []
            for _ in range(5):
                t1 = time.perf_counter()
                if pinned:
                    forest.predict_batch(im, list(range(n)), out=(pl, pp if probs else None))
                else:
                    forest.predict_batch(im, list(range(n)), probabilities=probs)
                ts.append(time.perf_counter() - t1)
            ms = float(np.median(ts)) * 1e3
            key = ("with_posteriors" if probs else "labels_only") + ("" if pinned else "_pageable")
            out[key] = {"batch_ms": ms, "frames_per_s": n / ms * 1e3, "mpix_per_s": n * W * H / ms / 1e3}
    dp = cf.DeviceBuffer(ctx, n * 5 * H * W * 4)
    dl = cf.DeviceBuffer(ctx, n * H * W)
    ts = []
    for _ in range(5):
        t1 = time.perf_counter()
        forest.predict_batch_device(im, list(range(n)), dp.ptr, dl.ptr)
        ts.append(time.perf_counter() - t1)
    ms = float(np.median(ts)) * 1e3
    out["device_outputs"] = {"batch_ms": ms, "frames_per_s": n / ms * 1e3, "mpix_per_s": n * W * H / ms / 1e3}
    out["roofline"] = classify_roofline(ms, n, W, H)
    dp.close()
    dl.close()
    lat = []
    for i in range(100):
        t1 = time.perf_counter()
        forest.predict(im, i % n, probabilities=True)
        lat.append((time.perf_counter() - t1) * 1e3)
    out["single_frame_p50_ms"] = float(np.percentile(lat, 50))
    out["single_frame_p99_ms"] = float(np.percentile(lat, 99))
    out["note"] = ("labels_only / with_posteriors: outputs copied to page-locked host arrays (cf_host_alloc); "
                   "*_pageable: to ordinary numpy arrays; device_outputs: posteriors and labels written to "
                   "device memory (caller-owned), no host copy")
    log(f"dense prediction (10 trees depth 20): {out['device_outputs']['frames_per_s']:.0f} frames/s on device, "
        f"{out['labels_only']['frames_per_s']:.0f} frames/s labels to host, "
        f"{out['with_posteriors']['frames_per_s']:.0f} frames/s with posteriors; single frame p50 "
        f"{out['single_frame_p50_ms']:.2f} ms p99 {out['single_frame_p99_ms']:.2f} ms")
    return out, forest


def worker_stats(before, after):
    """The serving worker's passes and allocations between two stats() readings (the longest pass is the
    longest since the server started)."""
    keys = ("passes", "passes_over_1ms", "slot_allocs", "pinned_allocs", "event_allocs")
    out = {k: int(after[k] - before[k]) for k in keys}
    out["longest_pass_ms"] = round(after["longest_pass_ms"], 3)
    return out


def serving_leg(ctx, forest, seconds, plain=True, mix=(0.4, 0.5, 0.1), rates=(1000, 2000, 3000, 4000)):
    """Poisson arrivals of RGB-D frames (by default 320x240 40%, 640x480 50%, 1280x960 10%) at 1000-4000
    frames/s into the serving loop (8 GB budget), against the plain path (one frame at a time, in arrival
    order).  Serving latency = results in host memory - the submit call at the frame's arrival (the server's
    clock); plain latency = results in host memory - arrival."""
    import threading
    import numpy as np
    import curfil_amd as cf
    sizes = [(320, 240), (640, 480), (1280, 960)]
    mix = np.array(mix, dtype=np.float64)
    frames = []
    for k, (W, H) in enumerate(sizes):
        rgb, depth, _ = cf.synth_dataset(8, W, H, sites=12, classes=5, zero_frac=0.05, seed0=7711 + 100 * k)
        frames.append((rgb, depth))
    singles = []
    for k, (W, H) in enumerate(sizes):
        im = cf.Images(ctx, 1, W, H)
        singles.append(im)
    sys.setswitchinterval(0.0002)
    shares = " / ".join(f"{W}x{H} ({100 * m:.0f}%)" for (W, H), m in zip(sizes, mix) if m > 0)
    results = {"workload": f"Poisson arrivals, frames {shares}, forest of 10 trees x depth 20, posteriors + "
                           f"labels to host", "rates": []}
    rng = np.random.default_rng(4711)
    server = cf.Server(ctx, forest, budget_mb=8192.0, max_batch=64)
    # warm-up: the slots and pinned buffers a burst needs are allocated before timing
    warm = [server.submit(frames[k][0][i % 8], frames[k][1][i % 8]) for k in range(3) for i in range(40)]
    for rid in warm:
        server.wait(rid, probabilities=False, labels=False)
    # the client's own pauses are not the library's latency: no cyclic garbage collection while timing
    import gc
    gc.collect()
    gc_was_enabled = gc.isenabled()
    gc.disable()
    for lam in rates:
        n = int(lam * seconds)
        arrivals = np.cumsum(rng.exponential(1.0 / lam, n))
        kinds = rng.choice(3, n, p=mix)
        picks = rng.integers(0, 8, n)
        row = {"rate_fps": lam, "frames": n}
        # serving loop
        st0 = server.stats()
        ids = [None] * n
        lat = np.zeros(n)

        def submit_all():
            t0 = time.perf_counter()
            for i in range(n):
                while time.perf_counter() - t0 < arrivals[i]:
                    pass
                rgb, depth = frames[kinds[i]]
                ids[i] = server.submit(rgb[picks[i]], depth[picks[i]])
        th = threading.Thread(target=submit_all)
        t_start = time.perf_counter()
        th.start()
        parts = np.zeros((n, 3))
        for i in range(n):
            while ids[i] is None:
                time.sleep(0.0001)
            _, _, lat[i] = server.wait(ids[i], probabilities=False, labels=False)
            parts[i] = server.last_breakdown
        th.join()
        slow = lat >= np.percentile(lat, 99)
        wall = time.perf_counter() - t_start
        px = float(sum(sizes[k][0] * sizes[k][1] for k in kinds))
        row["serving"] = {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                          "frames_per_s": n / wall, "mpix_per_s": px / wall / 1e6,
                          "launches": server.stats()["launches"] - st0["launches"],
                          "worker": worker_stats(st0, server.stats()),
                          "p99_tail_parts_ms": {"queued": float(parts[slow, 0].mean()),
                                                "to_launch": float(parts[slow, 1].mean()),
                                                "device_and_copy": float(parts[slow, 2].mean())},
                          "median_parts_ms": {"queued": float(np.median(parts[:, 0])),
                                              "to_launch": float(np.median(parts[:, 1])),
                                              "device_and_copy": float(np.median(parts[:, 2]))},
                          "by_size": {f"{sizes[k][0]}x{sizes[k][1]}": {
                              "frames": int(np.sum(kinds == k)),
                              "p50_ms": float(np.percentile(lat[kinds == k], 50)),
                              "p99_ms": float(np.percentile(lat[kinds == k], 99)),
                              "share_of_p99_tail": float(np.mean(kinds[slow] == k))}
                              for k in range(3) if np.any(kinds == k)}}
        if not plain:
            log(f"serving {lam} fps offered: {row['serving']['frames_per_s']:.0f} fps, p50 "
                f"{row['serving']['p50_ms']:.2f} p99 {row['serving']['p99_ms']:.2f} ms")
            results["rates"].append(row)
            continue
        # plain path: one frame at a time in arrival order (upload + integrate + predict + results to host), on
        # the first 2 s of the same arrivals (it saturates near 1000 frames/s, so longer only lengthens the run)
        n2 = min(n, int(lam * 2.0))
        lat2 = np.zeros(n2)
        t0 = time.perf_counter()
        for i in range(n2):
            now = time.perf_counter() - t0
            if now < arrivals[i]:
                while time.perf_counter() - t0 < arrivals[i]:
                    pass
            k = kinds[i]
            rgb, depth = frames[k]
            singles[k].set_rgbd_u8(0, rgb[picks[i]], depth[picks[i]])
            forest.predict(singles[k], 0, probabilities=True)
            lat2[i] = (time.perf_counter() - t0 - arrivals[i]) * 1e3
        wall2 = time.perf_counter() - t0
        px2 = float(sum(sizes[k][0] * sizes[k][1] for k in kinds[:n2]))
        row["plain"] = {"frames": n2, "p50_ms": float(np.percentile(lat2, 50)), "p99_ms": float(np.percentile(lat2, 99)),
                        "frames_per_s": n2 / wall2, "mpix_per_s": px2 / wall2 / 1e6}
        log(f"serving {lam} fps offered: {row['serving']['frames_per_s']:.0f} fps, p50 {row['serving']['p50_ms']:.2f} "
            f"p99 {row['serving']['p99_ms']:.2f} ms; plain {row['plain']['frames_per_s']:.0f} fps, "
            f"p50 {row['plain']['p50_ms']:.1f} p99 {row['plain']['p99_ms']:.1f} ms")
        results["rates"].append(row)
    server.close()
    if gc_was_enabled:
        gc.enable()
    return results


def cpu_threads():
    """Host threads for the multi-thread CPU legs: the host's processors, at most 16 (the host-CPU share a
    single GPU gets on the benchmark machine)."""
    return max(1, min(16, os.cpu_count() or 1))


class _OracleThreads:
    """The oracle's thread count (ORC_THREADS, read per call) for the duration of a block."""

    def __init__(self, n):
        self.n = n

    def __enter__(self):
        self.old = os.environ.get("ORC_THREADS")
        os.environ["ORC_THREADS"] = str(self.n)

    def __exit__(self, *exc):
        if self.old is None:
            os.environ.pop("ORC_THREADS", None)
        else:
            os.environ["ORC_THREADS"] = self.old


def cpu_baseline(forest, rgb, depth, labels, cfg_kw, sample, forest10=None):
    """The oracle (C++ restatement of the reference CPU path) on bounded samples of the workloads:

    * the headline: the root node of tree 0 on its first root samples, all F x T feature-threshold pairs,
      on 1 thread (`sample` / 4 samples) and split over the host threads (`sample` samples), per-thread
      counters summed afterwards as the reference's TBB loop does (random_tree_image.cpp:339-371);
    * the CPU-runnable workload end to end (4 frames 160x120: sampling, 1 tree to depth 8, prediction of
      all 4 frames);
    * dense prediction with the 10-tree depth-20 forest on 4 of the 64 frames, rows split over the host
      threads (random_forest_image.cpp:310)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import curfil_amd as cf
    import oracle as orc

    threads = cpu_threads()
    roots = forest.root_samples(0)[:sample].copy()
    used = np.unique(roots["image"])
    remap = np.zeros(int(used.max()) + 1, np.uint32)
    remap[used] = np.arange(len(used))
    roots["image"] = remap[roots["image"]]
    ds = orc.Dataset.from_u8(rgb[used], depth[used], labels[used])
    F, T = cfg_kw["featureCount"], cfg_kw["thresholds"]
    feats, thr = ds.generate_features(roots, 12345, F, T, cfg_kw["boxRadius"], cfg_kw["regionSize"])
    one = roots[:max(1, len(roots) // 4)]
    with _OracleThreads(1):
        t0 = time.perf_counter()
        ds.node_counters(one, feats, thr, 5)
        dt1 = time.perf_counter() - t0
    ev1 = len(one) * F * T
    log(f"cpu baseline, 1 thread: {ev1:.3e} evals in {dt1:.2f}s")
    with _OracleThreads(threads):
        t0 = time.perf_counter()
        ds.node_counters(roots, feats, thr, 5)
        dtm = time.perf_counter() - t0
    evm = len(roots) * F * T
    log(f"cpu baseline, {threads} threads: {evm:.3e} evals in {dtm:.2f}s")
    out = {"value": evm / dtm, "unit": "evals/s", "cores": threads, "kind": "port",
           "sample": f"root node of tree 0 of the batched workload: first {len(roots)} root samples x {F} "
                     f"features x {T} thresholds ({evm:.2e} evaluations), {threads} threads; {cpu_model()}, "
                     f"{os.cpu_count()} processors visible",
           "single_thread": {"value": ev1 / dt1, "cores": 1,
                             "sample": f"first {len(one)} root samples x {F} x {T} ({ev1:.2e} evaluations)"}}

    # the CPU-runnable workload end to end
    from helpers import CONFIG0
    r0, d0, l0 = cf.synth_dataset(4, 160, 120, sites=12, classes=5, zero_frac=0.05, seed0=4711)
    with _OracleThreads(threads):
        t0 = time.perf_counter()
        ds0 = orc.Dataset.from_u8(r0, d0, l0)
        of = ds0.train_forest(orc.config(**CONFIG0), 1, 6)
        t_train = time.perf_counter() - t0
        on, oh = of.tree(0)
        post = [np.array([orc.normalize_histogram(h, of.prior(), 0.0, True) for h in oh])]
        t1 = time.perf_counter()
        for i in range(4):
            ds0.predict(i, [on], post, 6)
        t_pred = time.perf_counter() - t1
    out["small_end_to_end"] = {"workload": "4 frames 160x120, 1 tree depth 8, F=100 x T=10, predict all 4 frames",
                               "train_s": t_train, "predict_s": t_pred, "nodes": len(on), "cores": threads}
    log(f"cpu baseline, small workload end to end: train {t_train:.2f}s ({len(on)} nodes), predict 4 frames "
        f"{t_pred:.2f}s")

    if forest10 is not None:
        W, H = 640, 480
        r3, d3, _ = cf.synth_dataset(4, W, H, sites=12, classes=5, zero_frac=0.05, seed0=5711)
        ds3 = orc.Dataset.from_u8(r3, d3)
        trees = [forest10.tree(t)[0] for t in range(forest10.num_trees)]
        posts = [forest10.posteriors(t) for t in range(forest10.num_trees)]
        with _OracleThreads(threads):
            t0 = time.perf_counter()
            for i in range(4):
                ds3.predict(i, trees, posts, 5)
            dt = time.perf_counter() - t0
        out["dense_prediction"] = {"workload": "4 of the 64 frames 640x480, 10 complete trees of depth 20",
                                   "mpix_per_s": 4 * W * H / dt / 1e6, "frames_per_s": 4 / dt, "cores": threads}
        log(f"cpu baseline, dense prediction: {4 * W * H / dt / 1e6:.2f} Mpix/s ({4 / dt:.2f} frames/s)")
    return out


if __name__ == "__main__":
    main()
