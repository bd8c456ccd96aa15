"""Benchmark of the batched thinning path on one MI355X.

Workload (default): the mixed batch -- 512 images with sides U{128..2048}
(polyline content, 60 polylines of width 3-15 px, 0.05% salt) and 16 volumes
with sides U{64..384} (vessel trees + 20 ellipsoids, ~3% foreground), seeded,
uint8 {0,1}, all resident in HBM before timing.  One step = thinning the whole
batch to convergence (pack, passes, unpack) with one batched launch sequence.
Reported beside it: the same items looped through the plain per-image
drop-in path, p50/p99 latency of 1024x1024 tree frames through the per-image
path and of the drop-in CenterlineExtraction on the 1024x1024 frame and the
256^3 volume, the dominant kernel's roofline, and the CPU baseline (the oracle, the
reference's rules restated word-parallel in C, item-parallel on every host
core, and on one core over a sample) on the whole workload, which also checks
every skeleton bit for bit.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
import argparse
import hashlib
import json
import os
import re
import statistics
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
DOMINANT = "thin2d_queue_pass"
# HBM bytes per launch of the dominant kernel, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs scaled by a copy of known size (see the file for the runs)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "thin2d_queue_pass.mixed_batch.pmc.txt")
# VALU instructions per launch of the dominant kernel, from a rocprofv3 SQ
# counter run (see the file for the run)
VALU_FILE = os.path.join(ROOT, "profiles", "thin2d_queue_pass.mixed_batch.sq.txt")
# one wave64 VALU instruction per cycle per CU: 256 CUs x 2.4 GHz
VALU_CEIL = 256 * 2.4e9


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def gen_items():
    from begputhinning_amd import synth
    d2, d3 = synth.batch_dims()
    n = len(d2) + len(d3)
    with ThreadPoolExecutor(min(16, os.cpu_count() or 4)) as ex:
        items = list(ex.map(lambda i: synth.workload4_item(i, d2, d3), range(n)))
    return items, len(d2)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def measured_per_launch(path, key):
    """(figure `key` per launch of the dominant kernel, launch shape it was
    measured with) from a committed counter summary, or (None, None)."""
    if not os.path.exists(path):
        return None, None
    text = open(path).read()
    m = re.search(r"^" + key + r":\s*([0-9.eE+]+)", text, re.M)
    sh = re.search(r"^launch_shape:\s*waves=(\d+) rows=(\d+) iters_per_pass=(\d+)", text, re.M)
    shape = tuple(int(v) for v in sh.groups()) if sh else None
    return (float(m.group(1)) if m else None), shape


def serve_pool(synth):
    """Distinct inputs the serving leg draws requests from: frames with sides
    U{256..1024} (polyline or tree content, half each) including 1024x1024
    ones, and volumes with sides U{64..512} (tree content)."""
    rng = np.random.default_rng(0xBE6E0004)
    specs = []
    for i in range(40):
        w, h = (int(v) for v in rng.integers(256, 1025, 2))
        specs.append(("poly" if i % 2 else "tree", w, h, 0))
    specs += [("tree", 1024, 1024, 0)] * 4 + [("poly", 1024, 1024, 0)] * 4
    for i in range(12):
        w, h, d = (int(v) for v in rng.integers(64, 513, 3))
        specs.append(("tree3d", w, h, d))

    def make(k):
        kind, w, h, d = specs[k]
        seed = synth.item_seed(4, 10000 + k)
        if kind == "poly":
            return synth.polylines(w, h, seed)
        if kind == "tree":
            return synth.tree2d(w, h, seed)
        return synth.tree3d(w, h, d, seed)

    with ThreadPoolExecutor(min(16, os.cpu_count() or 4)) as ex:
        return list(ex.map(make, range(len(specs))))


def serve_leg(args, torch, B, synth, dev):
    """Poisson arrivals of 80% frames / 20% volumes drawn from the pool,
    submitted open-loop to the server, at each rate of --serve-qps.  Frames
    join the server's running 2D queue at a pass boundary.  Each rate runs
    twice: for skeleton-ready latency (submit-to-done on the server's clock),
    and with the server also tracing every served skeleton into its segment
    list (its tracer threads: the device pre-pass, then the host trace) for
    segments-ready latency (submit-to-segments), which takes host cores the
    engine's threads would otherwise have.  Every served output is compared
    with the per-image device result of its input."""
    import threading
    from begputhinning_amd.serve import ThinningServer
    pool = serve_pool(synth)
    ins = [torch.from_numpy(a).to(dev) for a in pool]
    single = []
    for t in ins:
        single.append((B.thin2d_device(t) if t.dim() == 2 else B.thin3d_device(t))[0])
    torch.cuda.synchronize()
    frames = [k for k, a in enumerate(pool) if a.ndim == 2]
    vols = [k for k, a in enumerate(pool) if a.ndim == 3]
    n = args.serve_requests
    levels = []

    def run_level(qps, traced):
        """One open-loop run at `qps`: (skeleton ms, segments ms or None,
        outputs equal, Mvoxel/s, worker batches, picks)."""
        gaps, mix = synth.arrivals(n, qps, 0xBE6E0005)
        rng = np.random.default_rng(0xBE6E0006)
        pick = [int(frames[rng.integers(len(frames))]) if m < 0.8 else int(vols[rng.integers(len(vols))])
                for m in mix]
        outs = [torch.empty_like(ins[k]) for k in pick]
        torch.cuda.synchronize()
        srv = ThinningServer(budget_bytes=32 << 30, workers2d=2, workers3d=2, max_items=64)
        if traced:
            srv.enable_segments(args.serve_tracers)
        tickets = [None] * n
        t_sub = [0.0] * n
        lat = np.zeros(n)
        seg_lat = np.zeros(n)
        submitted = threading.Semaphore(0)
        nxt = [0]
        nxt_mu = threading.Lock()
        errors = []

        def waiter():
            try:
                while True:
                    with nxt_mu:
                        i = nxt[0]
                        nxt[0] += 1
                    if i >= n:
                        return
                    submitted.acquire()  # one release per submitted request
                    while tickets[i] is None:  # submitted in order; i may be a little ahead
                        time.sleep(1e-5)
                    if traced:
                        _, lat[i], seg_lat[i], _ = srv.wait_segments(tickets[i])
                    else:
                        _, lat[i] = srv.wait(tickets[i])
            except Exception as exc:  # reported after the run
                errors.append(exc)

        waiters = [threading.Thread(target=waiter) for _ in range(12)]
        for w in waiters:
            w.start()
        t0 = time.perf_counter()
        due = t0
        for i in range(n):
            due += gaps[i]
            while True:
                left = due - time.perf_counter()
                if left <= 0:
                    break
                if left > 2e-4:
                    time.sleep(left - 1e-4)
            t_sub[i] = time.perf_counter()
            tickets[i] = srv.submit(ins[pick[i]], outs[i])
            submitted.release()
        for w in waiters:
            w.join()
        # skeletons served per second: the submissions' span to the last
        # skeleton ready (the server's submit-to-done added to the host's
        # submit times)
        span = max(t_sub[i] + lat[i] / 1e3 for i in range(n)) - t0
        st = srv.stats()
        srv.close()
        if errors:
            raise errors[0]
        same = all(torch.equal(o, single[k]) for o, k in zip(outs, pick))
        vox = sum(pool[k].size for k in pick)
        return lat, (seg_lat if traced else None), same, vox / span / 1e6, st["batches"], pick

    for qps in args.serve_qps:
        lat, _, same, mvox, nb, pick = run_level(qps, False)
        tlat, seg_lat, same_t, _, _, _ = run_level(qps, True)
        is2 = np.array([pool[k].ndim == 2 for k in pick])
        big = np.array([pool[k].ndim == 2 and pool[k].shape == (1024, 1024) for k in pick])

        def pct(v, mask):
            return ({"p50": round(float(np.percentile(v[mask], 50)), 3),
                     "p99": round(float(np.percentile(v[mask], 99)), 3), "n": int(mask.sum())}
                    if mask.any() else None)

        lv = {"qps": qps, "requests": n,
              "skeleton_ms": {"2d": pct(lat, is2), "1024x1024": pct(lat, big), "3d": pct(lat, ~is2)},
              "segments_ms": {"2d": pct(seg_lat, is2), "1024x1024": pct(seg_lat, big),
                              "3d": pct(seg_lat, ~is2)},
              "skeleton_ms_while_tracing": {"1024x1024": pct(tlat, big), "3d": pct(tlat, ~is2)},
              "mvoxel_per_s": round(mvox, 1), "served_equals_single_item": same and same_t,
              "worker_batches": nb}
        log(f"serving {qps:.0f} QPS: skeleton 2D {lv['skeleton_ms']['2d']}, 1024^2 "
            f"{lv['skeleton_ms']['1024x1024']}, 3D {lv['skeleton_ms']['3d']}; with tracing: segments "
            f"1024^2 {lv['segments_ms']['1024x1024']}, 3D {lv['segments_ms']['3d']}, skeleton 1024^2 "
            f"{lv['skeleton_ms_while_tracing']['1024x1024']}; {lv['mvoxel_per_s']} Mvoxel/s, "
            f"equal {lv['served_equals_single_item']}")
        levels.append(lv)
    ok = [lv["qps"] for lv in levels if lv["skeleton_ms"]["1024x1024"]
          and lv["skeleton_ms"]["1024x1024"]["p99"] < 5.0]
    ok_seg = [lv["qps"] for lv in levels if lv["segments_ms"]["1024x1024"]
              and lv["segments_ms"]["1024x1024"]["p99"] < 5.0]
    return {"mix": "80% 2D U{256..1024} (a sixth of the pool 1024x1024), 20% 3D U{64..512}",
            "workers": f"frames: the running 2D queue (64 slots); volumes: 2 workers; "
                       f"{args.serve_tracers} tracer threads",
            "budget_bytes": 32 << 30, "levels": levels,
            "max_qps_1024x1024_p99_under_5ms": max(ok) if ok else None,
            "max_qps_1024x1024_segments_p99_under_5ms": max(ok_seg) if ok_seg else None}


def frame_leg(args, torch, B, synth, dev):
    """p50/p99 latency of 1024x1024 tree frames (distinct seeds) through the
    per-image device path, and its algorithmic HBM rate at p50: the frames'
    mean 2N + 2 ceil(N/8) I bytes (N pixels, I iterations) per p50 latency."""
    frames = [torch.from_numpy(synth.workload2(i)).to(dev) for i in range(min(args.frames, 64))]
    fout = torch.empty_like(frames[0])
    iters = [B.thin2d_device(f, out=fout)[1] for f in frames]
    torch.cuda.synchronize()
    lat = []
    for k in range(args.frames):
        f = frames[k % len(frames)]
        t2 = time.perf_counter()
        B.thin2d_device(f, out=fout)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t2) * 1e3)
    p50 = float(np.percentile(lat, 50))
    p99 = float(np.percentile(lat, 99))
    n = frames[0].numel()
    b_alg = float(np.mean([2 * n + 2 * ((n + 7) // 8) * i for i in iters]))
    gbs = b_alg / (p50 / 1e3) / 1e9
    log(f"1024x1024 frame latency p50 {p50:.3f} ms p99 {p99:.3f} ms; mean {np.mean(iters):.1f} "
        f"iterations, {gbs:.1f} GB/s algorithmic at p50")
    return {"p50": round(p50, 3), "p99": round(p99, 3), "frames": args.frames,
            "path": "per-image device-resident", "mean_iterations": round(float(np.mean(iters)), 2),
            "hbm_gbs_algorithmic_at_p50": round(gbs, 1),
            "b_alg_bytes": int(b_alg)}


def trace_leg(args, torch, B, synth, dev):
    """Latency of the segment tracer over skeletons already in device memory
    (bgt_extract_segments2d/3d_device: GPU pre-pass, copy of the point graph,
    host walk): 1024x1024 polyline and tree skeletons, 256^3 and 512^3 vessel
    trees, each the skeleton of a seeded input."""
    cases = [("2d_1024x1024_polylines", [synth.polylines(1024, 1024, synth.item_seed(4, 10000 + k))
                                         for k in range(4)], 40),
             ("2d_1024x1024_tree", [synth.workload2(k) for k in range(4)], 40),
             ("3d_256^3_tree", [synth.workload3(k) for k in range(2)], 12),
             ("3d_512^3_tree", [synth.tree3d(512, 512, 512, synth.item_seed(5, 90000))], 6)]
    st = torch.cuda.Stream()
    out = {"path": "bgt_extract_segments2d/3d_device on the skeleton in device memory"}
    for key, inputs, reps in cases:
        thin = B.thin2d_device if inputs[0].ndim == 2 else B.thin3d_device
        skels = [thin(torch.from_numpy(a).to(dev))[0] for a in inputs]
        torch.cuda.synchronize()
        for sk in skels:
            r = B.extract_segments_device(sk, stream=st)
        ts = []
        for k in range(reps):
            t0 = time.perf_counter()
            r = B.extract_segments_device(skels[k % len(skels)], stream=st)
            ts.append((time.perf_counter() - t0) * 1e3)
        out[key] = {"p50": round(float(np.percentile(ts, 50)), 3), "p99": round(float(np.percentile(ts, 99)), 3),
                    "calls": reps, "points": int(len(r["x"])), "segments": int(len(r["start"]) - 1)}
        del skels
    log("segment tracing p50: " + ", ".join(f"{k} {v['p50']} ms" for k, v in out.items() if isinstance(v, dict)))
    return out


def host_cores():
    """Threads the host gives this process: the affinity mask, capped by
    OMP_NUM_THREADS where the environment sets the process's CPU share."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(cap))) if cap.isdigit() and int(cap) > 0 else n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(items, got, iters, total_vox):
    """The oracle (the reference's rules restated word-parallel in C) on the
    host: item-parallel over every core the process has (dynamic
    scheduling; the oracle releases the GIL), over the whole workload, which
    also checks every skeleton and iteration count; and on one core over a
    sample of every 16th item."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O  # the checker and the CPU baseline

    def one(k):
        a = items[k]
        ref, rit = (O.thin2d(a) if a.ndim == 2 else O.thin3d(a))
        return int(np.array_equal(ref, got[k]) and rit == int(iters[k]))

    cores = host_cores()
    t3 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        ok = sum(ex.map(one, range(len(items))))
    all_s = time.perf_counter() - t3
    sample = list(range(0, len(items), 16))
    sample_vox = sum(items[k].size for k in sample)
    t4 = time.perf_counter()
    for k in sample:
        a = items[k]
        O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
    one_s = time.perf_counter() - t4
    model = cpu_model()
    cpu = {"value": round(total_vox / all_s / 1e6, 2), "unit": "Mvoxel/s", "cores": cores,
           "kind": "port", "cpu_model": model,
           "sample": f"whole workload ({len(items)} items, {total_vox / 1e6:.0f} Mvoxel) item-parallel "
                     f"on {cores} threads, dynamic scheduling: {all_s:.2f} s wall",
           "one_core": {"value": round(sample_vox / one_s / 1e6, 2), "unit": "Mvoxel/s", "cores": 1,
                        "sample": f"every 16th item ({len(sample)} items, {sample_vox / 1e6:.0f} "
                                  f"Mvoxel), {one_s:.2f} s"}}
    log(f"cpu baseline ({model}): {cpu['value']} Mvoxel/s on {cores} threads, "
        f"{cpu['one_core']['value']} Mvoxel/s on one; bit-exact items {ok}/{len(items)}")
    return cpu, ok == len(items)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=200, help="1024x1024 frames for p50/p99")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graph", type=int, default=0, help="replay pass chunks as a hipGraph")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of tuning them on the batch")
    ap.add_argument("--only-batch", action="store_true",
                    help="time the batched step only, at the library's launch shapes (no autotune, "
                         "per-image loop, latency, serving or CPU leg)")
    ap.add_argument("--serve-qps", type=float, nargs="*",
                    default=[1000.0, 2000.0, 4000.0, 6000.0, 8000.0, 12000.0],
                    help="Poisson arrival rates of the serving leg (none or 0 = skip it)")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-tracers", type=int, default=6,
                    help="the server's segment tracer threads (6 leave the engine threads "
                         "enough host cores at high rates)")
    ap.add_argument("--only-serving", action="store_true",
                    help="run only the serving leg and print its figures")
    ap.add_argument("--only-frames", action="store_true",
                    help="run only the 1024x1024 per-image latency leg (for counter runs of the "
                         "per-image path) and print its figures")
    ap.add_argument("--only-trace", action="store_true",
                    help="run only the segment-tracing latency leg and print its figures")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a kernel of "
                         "known traffic, to scale hardware byte counters)")
    args = ap.parse_args()
    if args.only_batch:
        args.no_cpu_baseline = True
        args.serve_qps = []
        args.no_autotune = True
    args.serve_qps = [q for q in args.serve_qps if q > 0]
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path "
              "(items shard across GPUs with no communication)", file=sys.stderr)
        sys.exit(2)

    import torch
    import begputhinning_amd as B
    from begputhinning_amd import synth
    from begputhinning_amd.batch import ThinningBatch

    assert torch.cuda.is_available(), "no GPU"
    dev = torch.device("cuda:0")
    if args.only_frames:
        print(json.dumps({"latency_1024x1024_ms": frame_leg(args, torch, B, synth, dev)}))
        return
    if args.only_serving:
        print(json.dumps({"serving": serve_leg(args, torch, B, synth, dev)}))
        return
    if args.only_trace:
        print(json.dumps({"trace_ms": trace_leg(args, torch, B, synth, dev)}))
        return
    t0 = time.time()
    items, n2 = gen_items()
    total_vox = sum(a.size for a in items)
    log(f"generated {len(items)} items, {total_vox / 1e6:.1f} Mvoxel in {time.time() - t0:.1f}s")

    ins = [torch.from_numpy(a).to(dev) for a in items]
    outs = [torch.empty_like(t) for t in ins]
    torch.cuda.synchronize()
    if args.copy_probe_mib:
        src = torch.ones(args.copy_probe_mib << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mib} MiB device-to-device")
        del src, dst

    # ---- batched path: the timed steps -------------------------------------
    batch = ThinningBatch()
    for t, o in zip(ins, outs):
        if t.dim() == 2:
            batch.add2d(t, o)
        else:
            batch.add3d(t, o)
    stream = torch.cuda.current_stream()
    tuned = None
    if not args.no_autotune:
        t_tune = time.perf_counter()
        tuned = batch.autotune(stream, reps=3)  # untimed; results do not depend on it
        log(f"autotune: {tuned} in {time.perf_counter() - t_tune:.2f}s")
    for _ in range(args.warmup):
        batch.run(stream, use_graph=bool(args.graph))
    step_ms = []
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ev0.record(stream)
        batch.run(stream, use_graph=bool(args.graph))
        ev1.record(stream)
        ev1.synchronize()
        step_ms.append(ev0.elapsed_time(ev1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    iters = batch.iterations()
    ms_step = wall * 1e3 / args.steps
    value = total_vox * args.steps / wall / 1e6
    log(f"batched: {ms_step:.2f} ms/step (events median {statistics.median(step_ms):.2f} ms), "
        f"{value:.0f} Mvoxel/s")

    # ---- kernel timing: HIP events around every launch of the 2D pass kernel
    # and every 3D iteration, on the streams they run on, in one more step
    batch.set_timing(True)
    ev0.record(stream)
    batch.run(stream, use_graph=False)
    ev1.record(stream)
    ev1.synchronize()
    timed_step_ms = ev0.elapsed_time(ev1)
    ms2d, l2d, ms3d, l3d = batch.kernel_time()
    batch.set_timing(False)
    n2d_vox = [a.size for a in items[:n2]]
    n3d_vox = [a.size for a in items[n2:]]
    it2, it3 = iters[:n2], iters[n2:]
    # algorithmic bytes: every iteration reads and writes the packed state once
    bytes_iter_2d = sum(2 * ((n + 7) // 8) * int(i) for n, i in zip(n2d_vox, it2))
    bytes_iter_3d = sum(2 * ((n + 7) // 8) * int(i) for n, i in zip(n3d_vox, it3))
    b_alg = 2 * total_vox + bytes_iter_2d + bytes_iter_3d
    avg_launch_s = ms2d / 1e3 / max(l2d, 1)
    bytes_per_launch = bytes_iter_2d / max(l2d, 1)
    achieved = bytes_per_launch / avg_launch_s if avg_launch_s > 0 else 0.0
    log(f"kernels: {DOMINANT} {ms2d:.2f} ms / {l2d} launches, 3D iterations {ms3d:.2f} ms / "
        f"{l3d}; timed step {timed_step_ms:.2f} ms")

    per_image = latency = centerline = tracing = None
    outs_pi = None
    if not args.only_batch:
        # ---- per-image drop-in path on the same items -----------------------
        outs_pi = [torch.empty_like(t) for t in ins]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for t, o in zip(ins, outs_pi):
            if t.dim() == 2:
                B.thin2d_device(t, out=o)
            else:
                B.thin3d_device(t, out=o)
        torch.cuda.synchronize()
        per_image_s = time.perf_counter() - t1
        per_image_value = total_vox / per_image_s / 1e6
        log(f"per-image loop: {per_image_s * 1e3:.1f} ms, {per_image_value:.0f} Mvoxel/s")
        per_image = {"value": round(per_image_value, 1), "unit": "Mvoxel/s",
                     "ratio_batched_over_per_image": round(value / per_image_value, 2)}

        latency = frame_leg(args, torch, B, synth, dev)

        # ---- the drop-in CenterlineExtraction on host arrays (C2, C3) -------
        cl = {"path": "drop-in CenterlineExtraction(data, 10, 3) on host uint8 arrays: "
                      "copy in, GPU thinning, GPU segment pre-pass, host trace/prune/smooth"}
        for key, a, reps in [("2d_1024x1024", synth.workload2(), 40), ("3d_256^3", synth.workload3(), 8)]:
            B.CenterlineExtraction(a, 10, 3)
            ts = []
            for _ in range(reps):
                t3 = time.perf_counter()
                B.CenterlineExtraction(a, 10, 3)
                ts.append((time.perf_counter() - t3) * 1e3)
            cl[key] = {"p50": round(float(np.percentile(ts, 50)), 3),
                       "p99": round(float(np.percentile(ts, 99)), 3), "calls": reps}
        log(f"CenterlineExtraction: 1024^2 p50 {cl['2d_1024x1024']['p50']} ms, "
            f"256^3 p50 {cl['3d_256^3']['p50']} ms")
        centerline = cl
        tracing = trace_leg(args, torch, B, synth, dev)

    serving = None
    if args.serve_qps:
        serving = serve_leg(args, torch, B, synth, dev)

    # ---- results back; bit-exactness vs the oracle (the CPU baseline leg) --
    got = [o.cpu().numpy() for o in outs]
    same_pi = (all(np.array_equal(a, o.cpu().numpy()) for a, o in zip(got, outs_pi))
               if outs_pi is not None else None)
    gold_path = os.path.join(ROOT, "tests", "golden", "workload4.json")
    gold = json.load(open(gold_path)) if os.path.exists(gold_path) else None
    hashes = [sha(a) for a in got]
    gold_ok = None
    if gold:
        gold_ok = (len(gold["skeleton_sha"]) == len(hashes)
                   and all(h == g for h, g in zip(hashes, gold["skeleton_sha"]))
                   and list(map(int, iters)) == gold["iterations"])

    cpu = None
    exact = None
    if not args.no_cpu_baseline:
        cpu, exact = cpu_baseline(items, got, iters, total_vox)

    traffic, traffic_shape = measured_per_launch(TRAFFIC_FILE, "traffic_bytes_per_launch")
    valu, valu_shape = measured_per_launch(VALU_FILE, "valu_insts_per_launch")
    # the library's default 2D shape (4 waves, 16 rows, K = 3) unless set by its environment
    shape = ((tuned["waves"], tuned["rows"], tuned["iters_per_pass"]) if tuned
             else tuple(int(os.environ.get(k, v)) for k, v in
                        (("BGT_2D_WAVES", 4), ("BGT_2D_ROWS", 16), ("BGT_2D_K", 3))))
    if traffic is not None and traffic_shape != shape:
        log(f"traffic figure was measured at launch shape {traffic_shape}, this run uses "
            f"{shape}: not reported")
        traffic = None
    if valu is not None and valu_shape != shape:
        valu = None
    # VALU issue: the profiled instructions per launch over what the CUs can
    # issue in the live average launch time
    valu_frac = (round(valu / (avg_launch_s * VALU_CEIL), 4)
                 if valu is not None and avg_launch_s > 0 else None)
    line = {
        "metric": "bit-exact thinned Mvoxel/s, batched vs per-image, 1 MI355X; p50/p99 ms",
        "value": round(value, 1),
        "unit": "Mvoxel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "mixed batch: 512 images U{128..2048}^2 + 16 volumes "
                               "U{64..384}^3, uint8, seeded", "items": len(items),
                   "mvoxel": round(total_vox / 1e6, 1), "graph": bool(args.graph),
                   "launch_shapes": tuned or "defaults"},
        "per_image": per_image,
        "latency_1024x1024_ms": latency,
        "centerline_ms": centerline,
        "trace_ms": tracing,
        "serving": serving,
        "bit_exact": {"vs_oracle": exact, "vs_committed_hashes": gold_ok,
                      "batched_equals_per_image": same_pi},
        "roofline": {"bound": "hbm", "kernel": DOMINANT, "achieved": round(achieved / 1e9, 1),
                     "peak": 8000.0, "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": traffic, "launches": l2d,
                     "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                     "bytes_per_launch": int(bytes_per_launch),
                     "timing": "HIP events around each launch on its stream, one extra step",
                     "valu_issue_frac": valu_frac,
                     "valu_insts_per_launch": valu},
        "three_d_iterations": {"launch_groups": l3d, "ms": round(ms3d, 3),
                               "avg_iteration_ms": round(ms3d / max(l3d, 1), 4)},
        "b_alg": {"bytes": int(b_alg), "frac_of_hbm_peak_per_step": round(
            b_alg / (ms_step / 1e3) / HBM_PEAK, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
