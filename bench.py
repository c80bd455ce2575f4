"""Benchmark: body-steps/s of the batched+served rigid-body step vs the plain single-world
GPU path, p50/p99 step latency of the served loop, roofline of the dominant launch group, and
the reference engine's multithreaded CPU step on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps 120] [--warmup 10]

Prints one JSON line.  Workload (value): 4096 independent worlds, body counts log-uniform in
8..512 (boxes with U(0.2,1.0) m edges and spheres with r U(0.1,0.6) m, p = 1/2, ~496k bodies),
each on its own kinematic ground box, dt = 1/60, 10 solver iterations, FP32, one
hipGraph-replayed batched step per timed step, steps 0..120 from the drop; launch schedule from
the tuning table (or --tune).  Beside it: the same worlds settled (steps 150..270), the plain
path (10k box rain, the same worlds one at a time, the 50k shape mix), the served loop at two
arrival rates, and the reference's CPU step in FP32 and FP64.

Profiling: --profile batched|box_rain_10k|shape_mix runs one leg's timed window alone;
--traffic-summary FETCH WRITE turns two rocprofv3 --pmc passes over `--profile batched
--calibrate-copy` into profiles/batched_traffic.json.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "physics-engine_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


class StdoutToStderr:
    """Sends file descriptor 1 to stderr while the compiled reference engine runs: it logs to
    stdout, which must carry nothing but the result line.  The C library's buffers are flushed
    before fd 1 is restored."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        import ctypes
        ctypes.CDLL(None).fflush(None)
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def stage_bytes(n, n_dyn, pairs, points, hull_bytes=0):
    """Algorithmic (compulsory) HBM bytes of one step per launch group: every stage reads its
    inputs once and writes its outputs once, one Gauss-Seidel sweep counted.
      K1 integrate + broadphase: 148 B / dynamic body + 64 B / body + 8 B / emitted pair
      K2 narrowphase:            96 B / candidate pair + hull bytes + 36 B / contact point
      K3 setup + solve:          216 B + 200 B / contact point"""
    return [148 * n_dyn + 64 * n + 8 * pairs, 96 * pairs + hull_bytes + 36 * points, 416 * points]


def step_bytes(n, n_dyn, pairs, points, hull_bytes=0):
    return sum(stage_bytes(n, n_dyn, pairs, points, hull_bytes))


def batched_schedule(args, worlds):
    """The launch schedule of the batched leg: the tuning table's entry for this workload
    (pe_amd/schedules.json, chosen by Batch.autotune on an MI355X), or with --tune the autotuner's
    choice made now on a separate copy of the worlds after args.settle settling steps."""
    from pe_amd.batch import Batch, load_schedule
    if not args.tune:
        return load_schedule("batched_4096_worlds"), "tuning table pe_amd/schedules.json: batched_4096_worlds"
    tb = Batch(worlds)
    tb.step(args.settle)
    tuned = tb.autotune(steps=2)
    sched = tb.schedule()
    tb.close()
    log(f"autotuned (block, solver group bodies) per class: {tuned}")
    if args.write_schedule:
        with open(args.write_schedule, "w") as f:
            json.dump(sched, f)
    return sched, f"autotuned in this run (Batch.autotune(steps=2) after {args.settle} settling steps)"


def run_batched(args, worlds, sched):
    """The timed window is the workload's own: steps 0..K from the drop.  The graph is captured and
    warmed up on a snapshot of step 0, which is restored before the timed replays.  The stage
    breakdown (hipEvents between the three launch groups) is measured over the same window, again
    from the snapshot, with the pair and contact counts of every step for its algorithmic bytes.
    Then the worlds run on to step args.settle and K more steps are timed there (settled piles)."""
    from pe_amd.batch import Batch
    t0 = time.perf_counter()
    b = Batch(worlds)
    log(f"batch created: {len(worlds)} worlds, {b.n_bodies} bodies in {time.perf_counter() - t0:.1f}s")
    b.apply_schedule(sched)
    b.snapshot()
    b.capture_graph()
    b.step(args.warmup)
    b.sync()
    b.restore()
    K = args.steps
    b.sync()
    w0 = time.perf_counter()
    dev_ms = b.step_timed(K)
    b.sync()
    wall = time.perf_counter() - w0
    c1 = b.counts()
    if c1["overflow_worlds"]:
        raise SystemExit(f"bench.py: {c1['overflow_worlds']} worlds exceeded a capacity; the step is not valid")
    out = dict(wall=wall, dev_ms=dev_ms, counts=c1, n=b.n_bodies, n_dyn=b.n_dynamic, batch=b)
    b.restore()
    st_ms, st_b = [0.0] * 3, [0.0] * 3
    pairs = points = 0
    for _ in range(K):
        ms = b.stage_times(1)
        c = b.counts()
        sb = stage_bytes(c["bodies"], c["dynamic"], c["pairs"], c["points"])
        for i in range(3):
            st_ms[i] += ms[i]
            st_b[i] += sb[i]
        pairs += c["pairs"]
        points += c["points"]
    out.update(stage_ms=[x / K for x in st_ms], stage_bytes=[x / K for x in st_b], pairs=pairs / K, points=points / K)
    if args.profile == "batched":
        return out
    if args.settle > K:
        b.step(args.settle - K)
    b.sync()
    s0 = b.steps_done
    t1 = time.perf_counter()
    b.step_timed(K)
    b.sync()
    sw = time.perf_counter() - t1
    cs = b.counts()
    if cs["overflow_worlds"]:
        raise SystemExit(f"bench.py: {cs['overflow_worlds']} worlds exceeded a capacity; the step is not valid")
    out["settled"] = {"window": f"steps {s0}..{s0 + K}", "body_steps_per_s": b.n_dynamic * K / sw,
                      "ms_per_step": 1000 * sw / K, "pairs": cs["pairs"], "points": cs["points"]}
    return out


def plain_stage_roofline(scene, start, steps):
    """Stage breakdown of the plain single-world path over steps start..start+steps of the scene
    (a fresh World; hipEvents between its four stages, every step), with each step's algorithmic
    bytes from its pair and contact counts; roofline of the costliest stage."""
    import pe_amd
    w = pe_amd.World(scene)
    if start:
        w.step(start)
    w.sync()
    w.set_timing(True)
    names = ["integrate", "broadphase (sweep-and-prune)", "narrowphase", "solve (colouring + Gauss-Seidel)"]
    t_prev = w.stage_times()
    ms, by = [0.0] * 4, [0.0] * 4
    hull = scene_mesh_bytes(scene)
    for _ in range(steps):
        w.step(1)
        t = w.stage_times()
        c = w.counts()
        for i in range(4):
            ms[i] += t[i] - t_prev[i]
        t_prev = t
        k1, k2, k3 = stage_bytes(c["bodies"], c["dynamic"], c["pairs"], c["points"], hull)
        integ = 148 * c["dynamic"]
        for i, v in enumerate((integ, k1 - integ, k2, k3)):
            by[i] += v
    w.close()
    k = max(range(4), key=lambda i: ms[i])
    ach = by[k] / (ms[k] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": None, "kernel": names[k], "algorithmic_bytes_per_launch": by[k] / steps,
            "launch_ms": ms[k] / steps, "stage_ms": [x / steps for x in ms],
            "stage_algorithmic_bytes": [x / steps for x in by], "window": f"steps {start}..{start + steps}",
            "whole_step": {"algorithmic_bytes": sum(by) / steps, "ms": sum(ms) / steps,
                           "achieved_gbs": sum(by) / (sum(ms) * 1e-3) / 1e9}}


def scene_mesh_bytes(scene):
    """Device bytes of every mesh of a scene (FP32 vertices and face normals, int32 face starts and
    indices): the hull and triangle data the narrowphase reads, counted once per step."""
    tot = 0
    for m in scene.meshes or []:
        v, st, ix, fn = m.flat()
        tot += 4 * (v.size + st.size + ix.size + fn.size)
    return tot


def run_plain(scene, steps, warmup=0):
    import pe_amd
    w = pe_amd.World(scene)
    if warmup:
        w.step(warmup)
    w.sync()
    t0 = time.perf_counter()
    w.step(steps)
    w.sync()
    dt = time.perf_counter() - t0
    counts = w.counts()
    w.close()
    return dt, counts


def run_plain_sample(worlds, steps, settle=0):
    """Body-steps/s of the plain single-world path on the given worlds one at a time: each world
    steps `settle` untimed steps (so the timed ones see the contact load the batched or served
    steps see), then `steps` timed ones."""
    import pe_amd
    tot = 0.0
    bodies = 0
    for sc in worlds:
        w = pe_amd.World(sc)
        if settle:
            w.step(settle)
        w.sync()
        t0 = time.perf_counter()
        w.step(steps)
        w.sync()
        tot += time.perf_counter() - t0
        bodies += sc.n_dynamic()
        w.close()
    return bodies * steps / tot


def served_pool(args):
    from pe_amd import scenes
    rng = np.random.default_rng(1004)
    sizes = scenes.log_uniform_counts(rng, args.serve_pool, 16, 4096)
    t0 = time.perf_counter()
    # mixed shapes as the 50k shape mix (boxes, spheres, cylinders, hulls over a heightfield); a pool of
    # distinct request scenes, sized from the same distribution, is cycled through
    pool = [scenes.served_shape_world(int(s), scenes.world_seed(1004, i)) for i, s in enumerate(sizes)]
    log(f"served: {len(pool)} distinct mixed-shape request worlds generated in {time.perf_counter() - t0:.1f}s")
    return pool


def run_served(args, pool, lam, seed):
    """Poisson arrivals at lam requests/s for args.serve_seconds; horizons U(30, 600); the request
    loop submits what has arrived, ticks (one step of every live world), and collects results.
    Reports throughput, per-tick latency and the per-request figures (admission wait, submission to
    result, queue depth), and the offered load: lam x the mean body-steps a request asks for."""
    from pe_amd.server import Server
    rng = np.random.default_rng(seed)
    T = args.serve_seconds
    n_req = int(lam * T * 1.5) + 8
    arrivals = np.cumsum(rng.exponential(1.0 / lam, n_req))
    horizons = rng.integers(30, 601, n_req)
    reqs = [pool[i % len(pool)] for i in range(n_req)]
    dyn = np.array([r.n_dynamic() for r in reqs], float)
    srv = Server(hbm_budget_bytes=int(args.serve_budget_gb * (1 << 30)), max_worlds=16384, max_world_bodies=4096)
    nxt = done = ticks = 0
    start = time.perf_counter()
    while True:
        now = time.perf_counter() - start
        if now >= T:
            break
        while nxt < n_req and arrivals[nxt] <= now:
            srv.submit(reqs[nxt], int(horizons[nxt]))
            nxt += 1
        live, _ = srv.tick()
        ticks += 1
        for rid in srv.poll():
            srv.result(rid)
            done += 1
        if live == 0 and nxt < n_req:
            time.sleep(max(0.0, min(0.01, arrivals[nxt] - (time.perf_counter() - start))))
    elapsed = time.perf_counter() - start
    st = srv.stats()
    rq = srv.request_stats()
    srv.close()
    offered = lam * float(np.mean(dyn * horizons))
    achieved = st["body_steps"] / elapsed
    return dict(rate=lam, seconds=elapsed, submitted=nxt, completed=done, ticks=ticks,
                body_steps=st["body_steps"], body_steps_per_s=achieved,
                offered_body_steps_per_s=offered, offered_over_achieved=offered / achieved,
                p50_ms=st["p50_ms"], p99_ms=st["p99_ms"], mean_ms=st["mean_ms"],
                queued_at_end=st["queued"], live_at_end=st["live"],
                admission_wait_p50_ms=rq["wait_p50_ms"], admission_wait_p99_ms=rq["wait_p99_ms"],
                submit_to_result_p50_ms=rq["latency_p50_ms"], submit_to_result_p99_ms=rq["latency_p99_ms"],
                queue_depth_mean=rq["queue_mean"], queue_depth_max=rq["queue_max"],
                capacity_errors=rq["capacity_errors"])


def run_shape_mix(args):
    """Mixed-shape world (50k boxes / spheres / cylinders / convex hulls over a 64 x 64
    heightfield) on the plain single-world path: args.mix_warm steps to let the drop land and
    pile up, then args.mix_steps timed steps.  With the reference built and the CPU leg on, the
    reference engine is timed for args.mix_cpu_steps steps from the same state (its contact
    caches start empty there), on args.cpu_threads threads."""
    from pe_amd import World, scenes
    t0 = time.perf_counter()
    sc = scenes.shape_mix_scene(args.mix_bodies, seed=1002)
    log(f"shape mix: {sc.n} bodies, {len(sc.meshes)} meshes generated in {time.perf_counter() - t0:.1f}s")
    w = World(sc)
    w.step(args.mix_warm)
    w.sync()
    t0 = time.perf_counter()
    w.step(args.mix_steps)
    w.sync()
    dt = time.perf_counter() - t0
    out = {"workload": f"{sc.n - 1} mixed bodies (40% boxes, 25% spheres, 15% cylinders, 20% hulls of 32 Gaussian "
                       f"points) over a 64x64 heightfield, steps {args.mix_warm}..{args.mix_warm + args.mix_steps}",
           "body_steps_per_s": sc.n_dynamic() * args.mix_steps / dt, "ms_per_step": 1000 * dt / args.mix_steps,
           "counts": w.counts()}
    if args.profile is None:
        out["roofline"] = plain_stage_roofline(sc, args.mix_warm, args.mix_steps)
    if not args.no_cpu:
        from pyoracle import RefWorld, ref_available
        if ref_available(32):
            with StdoutToStderr():
                r = RefWorld(sc, 32, threads=cpu_threads(args))
                r.set_state(*[np.asarray(x, np.float64) for x in w.state()])
                cdt = r.step_timed(args.mix_cpu_steps)
                del r
            out["cpu_reference"] = {"body_steps_per_s": sc.n_dynamic() * args.mix_cpu_steps / cdt,
                                    "ms_per_step": 1000 * cdt / args.mix_cpu_steps, "cores": cpu_threads(args),
                                    "sample": f"{args.mix_cpu_steps} steps from the GPU state at step "
                                              f"{args.mix_warm + args.mix_steps}"}
    w.close()
    return out


def cpu_threads(args):
    """Threads the reference's pool gets: --cpu-threads, at most the cores this process may run on."""
    return max(1, min(args.cpu_threads, len(os.sched_getaffinity(0))))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(worlds, steps, threads, budget_s, bits):
    """The reference engine itself (oracle/_ref/libperef{32,64}.so, compiled from its sources) when
    present, else the oracle restatement: the batched workload's worlds one after another, each
    for the same steps 0..steps from the drop that the timed GPU window covers, each with the
    reference's thread pool, until about budget_s seconds of CPU work."""
    from pyoracle import OracleWorld, RefWorld, ref_available
    kind = "reference" if ref_available(bits) else "port"
    tot_t, body_steps, nw = 0.0, 0, 0
    for sc in worlds:
        if kind == "reference":
            with StdoutToStderr():
                w = RefWorld(sc, bits, threads=threads)
                dt = w.step_timed(steps)
                del w
        else:
            w = OracleWorld(sc, bits)
            t0 = time.perf_counter()
            w.step(steps)
            dt = time.perf_counter() - t0
        tot_t += dt
        body_steps += sc.n_dynamic() * steps
        nw += 1
        if tot_t >= budget_s:
            break
    return dict(value=body_steps / tot_t, unit="body-steps/s", cores=threads if kind == "reference" else 1,
                kind=kind, sample=f"first {nw} of the 4096 worlds ({sum(w.n_dynamic() for w in worlds[:nw])} dynamic "
                                  f"bodies), steps 0..{steps} from the drop each, one world after another, "
                                  f"FP{bits}, " + (f"the reference's ThreadPool of {threads} threads"
                                                   if kind == "reference" else "single thread"))


def copy_calibration():
    """Three device-to-device copies of 1 GiB (hipMemcpy): a kernel of known traffic, ahead of
    everything else, for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE readings."""
    from pe_amd.hip import DeviceArray, synchronize
    x = DeviceArray((1 << 28,), np.float32)
    y = DeviceArray((1 << 28,), np.float32)
    for _ in range(3):
        y.copy_from(x)
    synchronize()
    x.close()
    y.close()
    log("calibration copy: 3 x (1 GiB read + 1 GiB written)")


def _pmc_rows(path, counter):
    import csv
    import glob
    files = [path] if os.path.isfile(path) else glob.glob(os.path.join(path, "**", "*counter_collection.csv"),
                                                              recursive=True)
    rows = []
    for fn in files:
        with open(fn) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] == counter:
                    rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return rows


def traffic_summary(fetch_path, write_path, steps, out_path):
    """HBM bytes per step of each launch group from two rocprofv3 --pmc passes (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch) over `python bench.py --profile batched --calibrate-copy`: each
    counter is scaled by its reading for the three 1 GiB calibration copies (the first three copy
    dispatches), and summed over the dispatches of the last `steps` steps (one k_list_reset opens a
    step; the profiled run ends with the staged pass over steps 0..K, the timed window's work)."""
    stages = ("k_world_broad", "k_narrow_list", "k_group_solve")
    res = {}
    for name, path in (("fetch", fetch_path), ("write", write_path)):
        rows = _pmc_rows(path, "FETCH_SIZE" if name == "fetch" else "WRITE_SIZE")
        copies = [r for r in rows if "copyBuffer" in r[1]][:3]
        kib = sum(r[2] for r in copies) / len(copies)
        scale = (1 << 30) / (kib * 1024)
        resets = [i for i, r in enumerate(rows) if "k_list_reset" in r[1]]
        window = rows[resets[-steps]:]
        per = [0.0, 0.0, 0.0]
        for _, kname, v in window:
            for k, sname in enumerate(stages):
                if sname in kname:
                    per[k] += v * 1024 * scale
        res[name] = dict(kib_per_gib_copy=kib, scale=scale, stage_bytes_per_step=[x / steps for x in per])
    tot = [res["fetch"]["stage_bytes_per_step"][k] + res["write"]["stage_bytes_per_step"][k] for k in range(3)]
    from pe_amd.batch import load_schedule
    out = {"steps_counted": steps, "window": f"steps 0..{steps}",
           "schedule": load_schedule("batched_4096_worlds"), "stages": list(stages),
           "stage_bytes_per_launch": tot, "fetch": res["fetch"], "write": res["write"]}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


def trace_summary(path, steps):
    """Per-step spans of the three launch groups (first start to last end of their kernels) and
    per-kernel mean durations in a rocprofv3 --kernel-trace of `python bench.py --profile batched`
    (one k_list_reset opens a step): its timed graph replays (steps 0..K), then the same K steps
    again, eager, with events between the launch groups (what the bench's stage figures time)."""
    import csv
    import glob
    files = [path] if os.path.isfile(path) else glob.glob(os.path.join(path, "**", "*kernel_trace.csv"), recursive=True)
    rows = []
    for fn in files:
        with open(fn) as f:
            for r in csv.DictReader(f):
                rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"],
                             int(r["Workgroup_Size_X"]), int(r["Grid_Size_X"]) // max(1, int(r["Workgroup_Size_X"]))))
    rows.sort()
    resets = [i for i, r in enumerate(rows) if "k_list_reset" in r[2]]
    out = {"steps": steps}
    for label, first in (("graph_replays", len(resets) - 2 * steps), ("eager_staged", len(resets) - steps)):
        out[label] = _trace_window(rows, resets, first, steps)
    print(json.dumps(out, indent=1))


def _trace_window(rows, resets, first, steps):
    stages = ("k_world_broad", "k_narrow_list", "k_group_solve")
    spans = [[] for _ in stages]
    whole = []
    per = {}
    for k in range(first, first + steps):
        lo = resets[k]
        hi = resets[k + 1] if k + 1 < len(resets) else len(rows)
        step = rows[lo:hi]
        whole.append((max(r[1] for r in step) - min(r[0] for r in step)) / 1e6)
        for si, sname in enumerate(stages):
            ks = [r for r in step if sname in r[2]]
            if ks:
                spans[si].append((max(r[1] for r in ks) - min(r[0] for r in ks)) / 1e6)
        for r in step:
            key = (r[2].replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0], r[3], r[4])
            per.setdefault(key, []).append((r[1] - r[0]) / 1e3)
    return {"median_span_ms": {s_: float(np.median(v)) for s_, v in zip(stages, spans)},
            "mean_span_ms": {s_: float(np.mean(v)) for s_, v in zip(stages, spans)},
            "median_step_ms": float(np.median(whole)), "mean_step_ms": float(np.mean(whole)),
            "kernels_mean_us": sorted([[k[0], k[1], k[2], float(np.mean(v)), len(v)] for k, v in per.items()],
                                      key=lambda x: -x[3])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120, help="timed steps: the batched workload's window from the drop")
    ap.add_argument("--warmup", type=int, default=10, help="untimed graph replays, from a snapshot of step 0")
    ap.add_argument("--settle", type=int, default=150,
                    help="the settled-pile figure times steps settle..settle+steps; --tune tunes after settle steps")
    ap.add_argument("--tune", action="store_true", help="run the autotuner now instead of using the tuning table")
    ap.add_argument("--write-schedule", default=None, help="with --tune: write the chosen schedule (JSON) here")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-rates", default="150,40", help="Poisson arrival rates (requests/s), one served run each")
    ap.add_argument("--serve-seconds", type=float, default=12.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-pool", type=int, default=192, help="distinct request worlds the arrivals cycle through")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU work per precision of the CPU baseline")
    ap.add_argument("--plain-steps", type=int, default=600)
    ap.add_argument("--no-mix", action="store_true")
    ap.add_argument("--mix-bodies", type=int, default=50000)
    ap.add_argument("--mix-warm", type=int, default=240)
    ap.add_argument("--mix-steps", type=int, default=60)
    ap.add_argument("--mix-cpu-steps", type=int, default=2)
    ap.add_argument("--profile", choices=["batched", "box_rain_10k", "shape_mix"], default=None,
                    help="run only this leg's timed window (for a profiler) and print its own line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device-to-device three times (a kernel of known traffic, "
                         "for scaling rocprof FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--traffic-summary", nargs=2, metavar=("FETCH", "WRITE"), default=None,
                    help="summarise two rocprofv3 --pmc outputs of --profile batched into profiles/batched_traffic.json")
    ap.add_argument("--trace-summary", default=None, metavar="TRACE",
                    help="per-step launch-group spans and kernel means of a rocprofv3 --kernel-trace of --profile batched")
    args = ap.parse_args()
    if args.trace_summary:
        trace_summary(args.trace_summary, args.steps)
        return
    if args.traffic_summary:
        traffic_summary(*args.traffic_summary, args.steps, os.path.join(ROOT, "profiles", "batched_traffic.json"))
        return
    if args.gpus != 1:
        print(f"bench.py: this build steps one GPU; --gpus {args.gpus} is not supported "
              f"(worlds shard across GPUs with no step-path traffic; see the design notes)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        copy_calibration()

    from pe_amd import scenes
    if args.profile == "box_rain_10k":
        c1 = scenes.box_rain_scene(10000, seed=1001)
        dt1, cnt1 = run_plain(c1, args.plain_steps)
        print(json.dumps({"profile": args.profile, "steps": args.plain_steps,
                          "ms_per_step": 1000 * dt1 / args.plain_steps, "final_counts": cnt1}), flush=True)
        return
    if args.profile == "shape_mix":
        args.no_cpu = True
        print(json.dumps({"profile": args.profile, **run_shape_mix(args)}), flush=True)
        return
    t0 = time.perf_counter()
    worlds = scenes.batch_worlds(4096, 8, 512, seed=1003)
    log(f"generated 4096 worlds in {time.perf_counter() - t0:.1f}s")
    sched, sched_src = batched_schedule(args, worlds)
    bt = run_batched(args, worlds, sched)
    K = args.steps
    value = bt["n_dyn"] * K / bt["wall"]
    ms_step = 1000.0 * bt["wall"] / K
    if args.profile == "batched":
        bt["batch"].close()
        print(json.dumps({"profile": "batched", "steps": K, "ms_per_step": ms_step,
                          "device_ms_per_step": bt["dev_ms"] / K, "body_steps_per_s": value,
                          "stage_ms": bt["stage_ms"], "stage_algorithmic_bytes": bt["stage_bytes"]}), flush=True)
        return
    dev_step_ms = bt["dev_ms"] / K
    B = sum(bt["stage_bytes"])
    names = ["k_world_broad (integrate + sweep-and-prune, all size classes)", "k_narrow_list (narrowphase)",
             "k_group_solve (solver setup + colouring + Gauss-Seidel, all size classes)"]
    k = max(range(3), key=lambda i: bt["stage_ms"][i])
    achieved = bt["stage_bytes"][k] / (bt["stage_ms"][k] * 1e-3) / 1e9
    log(f"batched: {value:.3e} body-steps/s, {ms_step:.3f} ms/step wall, {dev_step_ms:.3f} ms/step device, "
        f"mean pairs {bt['pairs']:.0f} points {bt['points']:.0f}; stage ms {['%.3f' % x for x in bt['stage_ms']]}, "
        f"dominant {names[k]} {achieved:.1f} GB/s; settled {bt['settled']}")
    bt["batch"].close()

    # HBM bytes per launch group from the rocprofv3 FETCH_SIZE / WRITE_SIZE passes over
    # `--profile batched` (profiles/batched_traffic.json), when they ran this schedule and window
    traffic, traffic_note = None, "no counter passes recorded"
    tp = os.path.join(ROOT, "profiles", "batched_traffic.json")
    if os.path.exists(tp):
        with open(tp) as f:
            tj = json.load(f)
        keys = ("class_upper_bodies", "block", "solver_group_bodies")
        same = all(list(tj.get("schedule", {}).get(kk, [])) == list(sched[kk]) for kk in keys)
        if same and tj.get("window") == f"steps 0..{K}":
            traffic, traffic_note = tj["stage_bytes_per_launch"][k], "counter passes over this schedule and window"
        else:
            traffic_note = "counter passes recorded for another schedule or window: not used"

    out = {
        "metric": "body-steps/s batched+served vs plain GPU at matched accuracy; p50/p99 step latency",
        "value": value, "unit": "body-steps/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batched: 4096 worlds, log-uniform 8-512 bodies (boxes + spheres), "
                               f"~496k bodies, dt 1/60, 10 iterations, hipGraph step, steps 0..{K} from the drop",
                   "worlds": len(worlds), "bodies": bt["n"], "dynamic_bodies": bt["n_dyn"],
                   "schedule": sched, "schedule_source": sched_src, "overflow_worlds": bt["counts"]["overflow_worlds"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note,
                     "kernel": names[k], "algorithmic_bytes_per_launch": bt["stage_bytes"][k],
                     "launch_ms": bt["stage_ms"][k], "window": f"steps 0..{K}, eager steps with hipEvents "
                                                               f"between the launch groups",
                     "stage_ms": bt["stage_ms"], "stage_algorithmic_bytes": bt["stage_bytes"],
                     "mean_pairs": bt["pairs"], "mean_points": bt["points"],
                     "whole_step": {"algorithmic_bytes": B, "ms": dev_step_ms,
                                    "achieved_gbs": B / (dev_step_ms * 1e-3) / 1e9}},
        "settled": bt["settled"],
    }
    if not args.no_plain:
        c1 = scenes.box_rain_scene(10000, seed=1001)
        dt1, cnt1 = run_plain(c1, args.plain_steps)
        plain_c1 = c1.n_dynamic() * args.plain_steps / dt1
        plain_c3 = run_plain_sample(worlds[:64], K)
        out["plain"] = {"box_rain_10k_body_steps_per_s": plain_c1, "box_rain_10k_steps": args.plain_steps,
                        "box_rain_10k_ms_per_step": 1000 * dt1 / args.plain_steps, "box_rain_10k_final_counts": cnt1,
                        "box_rain_10k_roofline": plain_stage_roofline(c1, 0, args.plain_steps),
                        "batch_worlds_one_at_a_time_body_steps_per_s": plain_c3,
                        "batch_worlds_one_at_a_time_sample": f"first 64 worlds, steps 0..{K} each",
                        "batched_over_plain_same_worlds": value / plain_c3,
                        "batched_over_plain_box_rain_10k": value / plain_c1}
        log(f"plain: c1 {plain_c1:.3e} body-steps/s ({1000 * dt1 / args.plain_steps:.2f} ms/step), "
            f"c3 worlds one at a time {plain_c3:.3e}")
    if not args.no_mix:
        mx = run_shape_mix(args)
        out["shape_mix"] = mx
        log(f"shape mix: {mx}")
    if not args.no_serve:
        pool = served_pool(args)
        runs = [run_served(args, pool, float(r), 1004 + i) for i, r in enumerate(args.serve_rates.split(","))]
        plain_same = run_plain_sample(pool[:32], 20, settle=150)
        out["served"] = {"workload": f"Poisson arrivals for {args.serve_seconds:g} s at each rate; body counts "
                                     f"log-uniform 16-4096, horizons U(30, 600); mixed shapes as the 50k shape mix "
                                     f"(boxes, spheres, cylinders, hulls over a 16x16 heightfield), {len(pool)} "
                                     f"distinct request worlds cycled; HBM budget {args.serve_budget_gb:g} GB",
                         "runs": runs, "plain_same_worlds_body_steps_per_s": plain_same,
                         "plain_same_worlds_sample": "first 32 request worlds, 150 untimed then 20 timed steps each",
                         "served_over_plain_same_worlds": runs[0]["body_steps_per_s"] / plain_same}
        log(f"served: {out['served']}")
    if not args.no_cpu:
        th = cpu_threads(args)
        cb = cpu_baseline(worlds, K, th, args.cpu_seconds, 32)
        cb["cpu_model"] = cpu_model()
        cb["host_cpus_available"] = len(os.sched_getaffinity(0))
        from pyoracle import ref_available
        if ref_available(64):
            cb["fp64"] = cpu_baseline(worlds, K, th, args.cpu_seconds, 64)
        out["cpu_baseline"] = cb
        log(f"cpu baseline: {cb}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
