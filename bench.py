"""Benchmark: body-steps/s of the batched+served rigid-body step vs the plain single-world
GPU path, p50/p99 step latency of the served loop, roofline of the fused step kernel, and
the reference engine's multithreaded CPU step on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps 120] [--warmup 10]

Prints one JSON line.  Workload (value): 4096 independent worlds, body counts log-uniform in
8..512 (boxes with U(0.2,1.0) m edges and spheres with r U(0.1,0.6) m, p = 1/2, ~496k bodies),
each on its own kinematic ground box, dt = 1/60, 10 solver iterations, FP32, one
hipGraph-replayed batched step per timed step.
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


def run_batched(args, worlds):
    from pe_amd.batch import Batch
    t0 = time.perf_counter()
    b = Batch(worlds)
    log(f"batch created: {len(worlds)} worlds, {b.n_bodies} bodies in {time.perf_counter() - t0:.1f}s")
    # the worlds settle for args.settle steps before tuning, so the launch configuration is
    # chosen for the contact load the timed steps see; the tuner times every candidate from a
    # snapshot of that state and restores it, so the timed steps follow the settled state
    b.step(args.settle)
    tuned = b.autotune(steps=2) if not args.no_tune else {}
    log(f"autotuned (block, solver group bodies) per class: {tuned}; config {b.config()}")
    b.capture_graph()
    b.step(args.warmup)
    b.sync()
    bt_start = b.steps_done
    # the worlds' states at the first timed step, for the CPU baseline
    start_states = [b.world_state(i) for i in range(len(worlds))] if not args.no_cpu else []
    c0 = b.counts()
    b.sync()
    w0 = time.perf_counter()
    dev_ms = b.step_timed(args.steps)
    b.sync()
    wall = time.perf_counter() - w0
    c1 = b.counts()
    pairs = 0.5 * (c0["pairs"] + c1["pairs"])
    points = 0.5 * (c0["points"] + c1["points"])
    nst = 10
    st_ms = [x / nst for x in b.stage_times(nst)]
    c2 = b.counts()
    return dict(wall=wall, dev_ms=dev_ms, counts0=c0, counts1=c1, pairs=pairs, points=points, start_step=bt_start,
                start_states=start_states,
                n=b.n_bodies, n_dyn=b.n_dynamic, tuned=tuned, batch=b, stage_ms=st_ms, stage_counts=c2)


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
    return dt, w.counts()


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


def run_served(args):
    from pe_amd import scenes
    from pe_amd.server import Server
    rng = np.random.default_rng(1004)
    lam, T = args.serve_rate, args.serve_seconds
    n_req = int(lam * T * 1.2) + 8
    gaps = rng.exponential(1.0 / lam, n_req)
    arrivals = np.cumsum(gaps)
    sizes = scenes.log_uniform_counts(rng, n_req, 16, 4096)
    horizons = rng.integers(30, 601, n_req)
    t0 = time.perf_counter()
    # mixed shapes as the 50k shape mix (boxes, spheres, cylinders, hulls over a heightfield); a pool of
    # distinct request scenes, sized from the same distribution, is cycled through
    pool = [scenes.served_shape_world(int(s), scenes.world_seed(1004, i))
            for i, s in enumerate(sizes[:args.serve_pool])]
    reqs = [pool[i % len(pool)] for i in range(n_req)]
    sizes = np.array([r.n - 1 for r in reqs])
    log(f"served: {len(pool)} distinct mixed-shape request worlds generated in {time.perf_counter() - t0:.1f}s for "
        f"{n_req} arrivals (mean {sizes.mean():.0f} bodies, mean horizon {horizons.mean():.0f})")
    srv = Server(hbm_budget_bytes=int(args.serve_budget_gb * (1 << 30)), max_worlds=16384, max_world_bodies=4096)
    nxt = 0
    done = 0
    start = time.perf_counter()
    ticks = 0
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
    srv.close()
    # the plain single-world path on the same request worlds, one at a time from their first step
    plain_same = run_plain_sample(pool[:32], 20, settle=150)
    return dict(plain_same_worlds_body_steps_per_s=plain_same,
                plain_same_worlds_sample="first 32 request worlds, 150 untimed then 20 timed steps each",
                served_over_plain_same_worlds=st["body_steps"] / elapsed / plain_same,
                workload=f"Poisson arrivals at {lam:g}/s for {T:g} s; body counts log-uniform 16-4096, horizons "
                         f"U(30, 600); mixed shapes as the 50k shape mix (boxes, spheres, cylinders, hulls over a "
                         f"16x16 heightfield), {len(pool)} distinct request worlds cycled",
                rate=lam, seconds=elapsed, submitted=nxt, completed=done, ticks=ticks,
                body_steps=st["body_steps"], body_steps_per_s=st["body_steps"] / elapsed,
                p50_ms=st["p50_ms"], p99_ms=st["p99_ms"], mean_ms=st["mean_ms"], budget_gb=args.serve_budget_gb)


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
    w.set_timing(True)       # stage breakdown over 10 further steps (events between stages)
    w.step(10)
    w.sync()
    st = [x / 10 for x in w.stage_times()]
    out = {"workload": f"{sc.n - 1} mixed bodies (40% boxes, 25% spheres, 15% cylinders, 20% hulls of 32 Gaussian "
                       f"points) over a 64x64 heightfield, steps {args.mix_warm}..{args.mix_warm + args.mix_steps}",
           "body_steps_per_s": sc.n_dynamic() * args.mix_steps / dt, "ms_per_step": 1000 * dt / args.mix_steps,
           "stage_ms": st, "counts": w.counts()}
    if not args.no_cpu:
        from pyoracle import RefWorld, ref_available
        if ref_available(32):
            with StdoutToStderr():
                r = RefWorld(sc, 32, threads=args.cpu_threads)
                r.set_state(*[np.asarray(x, np.float64) for x in w.state()])
                cdt = r.step_timed(args.mix_cpu_steps)
                del r
            out["cpu_reference"] = {"body_steps_per_s": sc.n_dynamic() * args.mix_cpu_steps / cdt,
                                    "ms_per_step": 1000 * cdt / args.mix_cpu_steps, "cores": args.cpu_threads,
                                    "sample": f"{args.mix_cpu_steps} steps from the GPU state at step {args.mix_warm + args.mix_steps + 10}"}
    w.close()
    return out


def cpu_baseline(worlds, states, start_step, threads, budget_s):
    """The reference engine itself (oracle/_ref/libperef32.so, compiled from its sources) when
    present, else the oracle restatement; worlds stepped one after another, each with the
    reference's batched-loop thread pool, until about budget_s seconds of CPU work.  Each world
    starts from its state at the batched leg's first timed step (its contact caches empty), so
    the CPU sees the contact load the timed GPU steps see."""
    from pyoracle import OracleWorld, RefWorld, ref_available
    kind = "reference" if ref_available(32) else "port"
    steps = 10
    tot_t, body_steps, nw = 0.0, 0, 0
    for sc, st in zip(worlds, states):
        st64 = [np.asarray(x, np.float64) for x in st]
        if kind == "reference":
            with StdoutToStderr():
                w = RefWorld(sc, 32, threads=threads)
                w.set_state(*st64)
                dt = w.step_timed(steps)
                del w
        else:
            w = OracleWorld(sc, 32)
            w.set_state(*st64)
            t0 = time.perf_counter()
            w.step(steps)
            dt = time.perf_counter() - t0
        tot_t += dt
        body_steps += sc.n_dynamic() * steps
        nw += 1
        if tot_t >= budget_s:
            break
    return dict(value=body_steps / tot_t, unit="body-steps/s", cores=threads if kind == "reference" else 1,
                kind=kind, sample=f"first {nw} of the 4096 worlds, {steps} steps each from their state at step "
                                  f"{start_step}, sequential, FP32, "
                                  f"{'ThreadPool of ' + str(threads) + ' threads' if kind == 'reference' else 'single thread'}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--settle", type=int, default=150,
                    help="batched worlds are stepped this many steps before tuning and the warmup")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=150.0)
    ap.add_argument("--serve-seconds", type=float, default=15.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-pool", type=int, default=192, help="distinct request worlds the arrivals cycle through")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--plain-steps", type=int, default=600)
    ap.add_argument("--no-mix", action="store_true")
    ap.add_argument("--mix-bodies", type=int, default=50000)
    ap.add_argument("--mix-warm", type=int, default=240)
    ap.add_argument("--mix-steps", type=int, default=60)
    ap.add_argument("--mix-cpu-steps", type=int, default=2)
    ap.add_argument("--profile-leg", choices=["box_rain_10k", "shape_mix"], default=None,
                    help="run only this plain-path leg on the GPU (for a profiler) and print its own line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device-to-device with torch (a kernel of known traffic, "
                         "for scaling rocprof FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build steps one GPU; --gpus {args.gpus} is not supported "
              f"(worlds shard across GPUs with no step-path traffic; see the design notes)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log("calibration copy: 3 x 1 GiB read + 1 GiB written")
        del x, y

    from pe_amd import scenes
    if args.profile_leg == "box_rain_10k":
        c1 = scenes.box_rain_scene(10000, seed=1001)
        dt1, cnt1 = run_plain(c1, args.plain_steps)
        print(json.dumps({"profile_leg": args.profile_leg, "steps": args.plain_steps,
                          "ms_per_step": 1000 * dt1 / args.plain_steps, "final_counts": cnt1}), flush=True)
        return
    if args.profile_leg == "shape_mix":
        args.no_cpu = True
        print(json.dumps({"profile_leg": args.profile_leg, **run_shape_mix(args)}), flush=True)
        return
    t0 = time.perf_counter()
    worlds = scenes.batch_worlds(4096, 8, 512, seed=1003)
    log(f"generated 4096 worlds in {time.perf_counter() - t0:.1f}s")

    bt = run_batched(args, worlds)
    K = args.steps
    value = bt["n_dyn"] * K / bt["wall"]
    ms_step = 1000.0 * bt["wall"] / K
    dev_step_ms = bt["dev_ms"] / K
    B = step_bytes(bt["n"], bt["n_dyn"], bt["pairs"], bt["points"])
    step_gbs = B / (dev_step_ms * 1e-3) / 1e9
    sc = bt["stage_counts"]
    sb = stage_bytes(sc["bodies"], sc["dynamic"], sc["pairs"], sc["points"])
    names = ["k_world_broad (integrate + sweep-and-prune, all size classes)", "k_narrow_list (narrowphase)",
             "k_group_solve (solver setup + colouring + Gauss-Seidel, all size classes)"]
    k = max(range(3), key=lambda i: bt["stage_ms"][i])
    achieved = sb[k] / (bt["stage_ms"][k] * 1e-3) / 1e9
    log(f"batched: {value:.3e} body-steps/s, {ms_step:.3f} ms/step wall, {dev_step_ms:.3f} ms/step device, "
        f"pairs {bt['pairs']:.0f} points {bt['points']:.0f} bytes/step {B / 1e6:.1f} MB -> {step_gbs:.1f} GB/s; "
        f"stage ms {['%.3f' % x for x in bt['stage_ms']]}, dominant {names[k]} {achieved:.1f} GB/s")
    bt["batch"].close()

    # HBM bytes per launch group measured by rocprofv3 FETCH_SIZE / WRITE_SIZE passes over this same
    # command, scaled by the counters' reading for a 1 GiB copy (profiles/batched_traffic.json).
    traffic = None
    tp = os.path.join(ROOT, "profiles", "batched_traffic.json")
    if os.path.exists(tp):
        with open(tp) as f:
            traffic = json.load(f).get("stage_bytes_per_launch", [None, None, None])[k]

    out = {
        "metric": "body-steps/s batched+served vs plain GPU at matched accuracy; p50/p99 step latency",
        "value": value, "unit": "body-steps/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched: 4096 worlds, log-uniform 8-512 bodies (boxes + spheres), "
                               f"~496k bodies, dt 1/60, 10 iterations, hipGraph step, timed from step "
                               f"{bt['start_step']}",
                   "worlds": len(worlds), "bodies": bt["n"], "dynamic_bodies": bt["n_dyn"],
                   "autotuned_block_and_solver_group": {str(k): list(v) for k, v in bt["tuned"].items()}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": names[k], "algorithmic_bytes_per_launch": sb[k], "launch_ms": bt["stage_ms"][k],
                     "stage_ms": bt["stage_ms"], "stage_algorithmic_bytes": sb,
                     "whole_step": {"algorithmic_bytes": B, "ms": dev_step_ms, "achieved_gbs": step_gbs}},
    }
    if not args.no_plain:
        c1 = scenes.box_rain_scene(10000, seed=1001)
        dt1, cnt1 = run_plain(c1, args.plain_steps)
        plain_c1 = c1.n_dynamic() * args.plain_steps / dt1
        plain_c3 = run_plain_sample(worlds[:64], 20, settle=args.settle)
        out["plain"] = {"box_rain_10k_body_steps_per_s": plain_c1, "box_rain_10k_steps": args.plain_steps,
                        "box_rain_10k_ms_per_step": 1000 * dt1 / args.plain_steps, "box_rain_10k_final_counts": cnt1,
                        "batch_worlds_one_at_a_time_body_steps_per_s": plain_c3,
                        "batch_worlds_one_at_a_time_sample": f"first 64 worlds, {args.settle} untimed then 20 timed "
                                                             f"steps each",
                        "batched_over_plain_same_worlds": value / plain_c3,
                        "batched_over_plain_box_rain_10k": value / plain_c1}
        log(f"plain: c1 {plain_c1:.3e} body-steps/s ({1000 * dt1 / args.plain_steps:.2f} ms/step), "
            f"c3 worlds one at a time {plain_c3:.3e}")
    if not args.no_mix:
        mx = run_shape_mix(args)
        out["shape_mix"] = mx
        log(f"shape mix: {mx}")
    if not args.no_serve:
        sv = run_served(args)
        out["served"] = sv
        log(f"served: {sv}")
    if not args.no_cpu:
        cb = cpu_baseline(worlds, bt["start_states"], bt["start_step"], args.cpu_threads, args.cpu_seconds)
        out["cpu_baseline"] = cb
        log(f"cpu baseline: {cb}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
