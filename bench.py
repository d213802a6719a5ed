#!/usr/bin/env python
"""Benchmark of the PPF registration hot path on one MI355X.

Workload (the dataset-like configuration): four seeded superquadric models of
3k-8k points, one 50,000-point scene holding all four on a ground plane,
ref_point_df 5 (10,000 scene reference points), vote threshold 0.3, GPU
spatial-hash clustering.  Models and scene are built/uploaded before the timed
region, so inputs are resident in HBM.  One step = the plain per-scene
pipeline registering all four models in the scene (four Model::ppf_lookup
calls, each ending with its pose on the host).

Prints one JSON line: value = scene pair-lookups/s over the timed steps, with
registrations/s and p50/p99 per-registration latency beside it, the roofline
of the vote kernel (algorithmic bytes 8*L + 4*V per launch over its HIP-event
duration) and a CPU baseline (the oracle's OpenMP vote loop on a bounded
sample of the same scene's reference points, extrapolated per lookup).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ppf-registration-spatial-hashing_amd"))

import ppf_amd  # noqa: E402
import synth  # noqa: E402

METRIC = "vote pair-lookups/s + registrations/s; p50/p99 latency, batched vs plain"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def hbm_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on this workload from the committed
    FETCH_SIZE/WRITE_SIZE counter runs (scaled by the copy-kernel calibration
    recorded beside them), or None when no such run is recorded."""
    path = os.path.join(ROOT, "profiles", "dataset_like_hbm_traffic.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return None, None
    return float(k["bytes_per_dispatch"]), "profiles/dataset_like_hbm_traffic.md"


def valu_insts(kernel: str):
    """Wave64 VALU instructions per launch of `kernel` on this workload, from
    the committed SQ_INSTS_VALU counter run, or None when none is recorded."""
    path = os.path.join(ROOT, "profiles", "dataset_like_sq_counters.json")
    try:
        with open(path) as f:
            return float(json.load(f)["kernels"][kernel]["valu_insts_per_dispatch"])
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(w, budget_s: float):
    """Oracle (hoisted-frame OpenMP port of the reference vote loop) on a
    stride sample of the scene's reference points, all four models."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ppf_oracle as O
    scene = w.scenes[0]
    n_s = len(scene[0])
    refs_all = np.arange(0, n_s, w.rpdf, dtype=np.int64)
    tot_l = tot_v = 0
    tot_s = 0.0
    sampled = 0
    per_model = budget_s / len(w.models)
    for m, d in zip(w.models, w.d_dists):
        om = O.OracleModel(m, d)
        # one stride-sampled reference point first, then as many as fit the
        # per-model budget (the two samples are disjoint)
        k = 1
        refs = refs_all[len(refs_all) // 2: len(refs_all) // 2 + 1]
        r = om.vote_throughput(scene, refs)
        left = per_model - r["seconds"]
        if r["seconds"] > 0 and left > r["seconds"]:
            k2 = int(min(len(refs_all) - 1, left / r["seconds"]))
            if k2 >= 1:
                more = refs_all[:: max(1, len(refs_all) // k2)][:k2]
                more = more[more != refs[0]]
                r2 = om.vote_throughput(scene, more)
                r = {key: r[key] + r2[key] for key in r}
                refs = np.concatenate([refs, more])
        tot_l += r["lookups"]
        tot_v += r["votes"]
        tot_s += r["seconds"]
        sampled += len(refs)
        del om
    return dict(value=tot_l / tot_s, unit="pair-lookups/s", cores=O.num_threads(), kind="port",
                sample=f"oracle vote loop (OpenMP, hoisted frames) on {sampled} of "
                       f"{len(refs_all) * len(w.models)} (model, reference point) rows, "
                       f"{tot_s:.1f} s, extrapolated per lookup",
                votes_per_s=tot_v / tot_s)


def run_large_hash(ctx, rpdf=100):
    """Large-hash stress workload: one 30,000-point superquadric with 8 bumps
    (9e8 model pairs) and a 200,000-point scene with three instances and 40%
    clutter.  The model build (keys, hash-array sort, vote order, records) is
    timed in full, twice; voting runs on every 100th scene point (2,000
    reference points, 4e8 lookups) rather than every 5th (40,000 reference
    points, 8e9 lookups, about four minutes of voting), so that the bench
    stays within minutes; throughput per lookup is what it reports."""
    w = synth.large_hash()
    builds = []
    m = None
    for _ in range(2):
        if m is not None:
            m.close()
        t = time.perf_counter()
        m = ppf_amd.Model(w.models[0], w.d_dists[0], w.vote_count_threshold, ctx=ctx)
        builds.append((time.perf_counter() - t) * 1e3)
    info = m.info()
    log(f"large-hash model built: {builds[-1]:.0f} ms")
    scene = ppf_amd.Scene(w.scenes[0], 0.0, rpdf, ctx=ctx)
    t = time.perf_counter()
    r = m.ppf_lookup(scene)
    r.pose()
    dt = time.perf_counter() - t
    i = dict(r.info)
    r.close()
    scene.close()
    m.close()
    log(f"large-hash sample lookup: {dt:.1f}s")
    return {"model_points": len(w.models[0][0]), "model_pairs": info["n_pairs"],
            "unique_keys": info["n_unique_keys"], "model_build_ms": builds[-1],
            "model_pairs_per_s": info["n_pairs"] / (builds[-1] * 1e-3),
            "scene_points": len(w.scenes[0][0]), "ref_point_df": rpdf,
            "lookups": i["lookups"], "votes": i["votes"], "vote_ms": i["vote_ms"],
            "lookups_per_s": i["lookups"] / (i["vote_ms"] * 1e-3),
            "votes_per_s": i["votes"] / (i["vote_ms"] * 1e-3),
            "n_kept": i["n_kept"], "cluster_ms": i["cluster_ms"], "lookup_s": dt}


def run_forms(ctx, plain_models, plain_scene, w, log_every=True):
    """Batched and served forms beside the plain pipeline.

    1. The dataset-like jobs (4 models x the 50,000-point scene) as one
       ppf_batch_lookup call, against the plain step.
    2. A request set shaped like the batched/served configurations but sized
       to finish within seconds: 8 seeded superquadric models (1k-6k points)
       built once, 48 scenes of 2k-20k points (log-uniform) each holding 1-4
       of them.  Plain = each request's scene uploaded and looked up target by
       target; batched = all (scene, model) jobs in one ppf_batch_lookup;
       served = ppf_server with Poisson arrivals at 50, 200 and 2000 req/s.
       The seed gives models with at least 0.45 unique keys per point, as in
       the dataset-like models; flat-faced superquadrics with ~0.1 (huge
       buckets) keep 10^7-10^8 votes per job and take minutes per job in
       clustering, in the reference's algorithm as here.
    3. The large-hash stress workload (run_large_hash).
    """
    out = {}
    # ---- 1. dataset-like jobs batched
    t = time.perf_counter()
    res, info = ppf_amd.batch_lookup(plain_models, [plain_scene] * len(plain_models),
                                     w.vote_count_threshold, cpu_clustering=w.cpu_clustering, ctx=ctx)
    for r in res:
        r.pose()
        r.close()
    dt = time.perf_counter() - t
    out["dataset_like_batched"] = {"jobs": len(plain_models), "s": dt,
                                   "registrations_per_s": len(plain_models) / dt,
                                   "lookups_per_s": info["lookups"] / dt,
                                   "vote_launches": info["vote_launches"], "vote_ms": info["vote_ms"],
                                   "finish_ms": info["finish_ms"]}
    log(f"batched dataset-like: {dt:.2f}s")

    # ---- 2. request set
    rw = synth.batched(seed=11, n_models=8, n_scenes=48, model_range=(1000, 6000),
                       scene_range=(2000, 20000), models_per_scene=(1, 4))
    models = [ppf_amd.Model(m, d, rw.vote_count_threshold, ctx=ctx) for m, d in zip(rw.models, rw.d_dists)]
    reqs = list(zip(rw.scenes, rw.scene_models))
    n_targets = sum(len(ids) for _, ids in reqs)
    desc = {"models": len(models), "model_points": [len(m[0]) for m in rw.models],
            "requests": len(reqs), "targets": n_targets,
            "scene_points": [len(s[0]) for s, _ in reqs], "ref_point_df": rw.rpdf,
            "threshold": rw.vote_count_threshold, "clustering": "gpu"}

    def plain_request(sc, ids):
        t0 = time.perf_counter()
        scene = ppf_amd.Scene(sc, 0.0, rw.rpdf, ctx=ctx)
        L = 0
        for mi in ids:
            r = models[mi].ppf_lookup(scene)
            r.pose()
            L += r.info["lookups"]
            r.close()
        scene.close()
        return time.perf_counter() - t0, L

    plain_request(*reqs[0])  # warm-up
    service, lookups = [], 0
    t0 = time.perf_counter()
    for sc, ids in reqs:
        dt, L = plain_request(sc, ids)
        service.append(dt)
        lookups += L
    plain_s = time.perf_counter() - t0
    log(f"plain request set: {plain_s:.2f}s")

    # batched: every (scene, model) job in one call (scenes resident)
    scenes = [ppf_amd.Scene(sc, 0.0, rw.rpdf, ctx=ctx) for sc, _ in reqs]
    jm = [models[mi] for _, ids in reqs for mi in ids]
    js = [scenes[k] for k, (_, ids) in enumerate(reqs) for _ in ids]
    res, _ = ppf_amd.batch_lookup(jm, js, rw.vote_count_threshold, ctx=ctx)  # warm-up
    for r in res:
        r.close()
    t0 = time.perf_counter()
    res, binfo = ppf_amd.batch_lookup(jm, js, rw.vote_count_threshold, ctx=ctx)
    for r in res:
        r.pose()
        r.close()
    batch_s = time.perf_counter() - t0
    for sc in scenes:
        sc.close()
    log(f"batched request set: {batch_s:.2f}s")

    # served: Poisson arrivals; the plain comparison is the same arrivals
    # queued FIFO in front of the plain per-request service times above
    served = []
    srv = ppf_amd.Server(models, round_ms=20.0, ctx=ctx)
    rng = np.random.default_rng(11)
    for rate in (50.0, 200.0, 2000.0):
        arr = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
        arr -= arr[0]
        tickets = [None] * len(reqs)

        def submitter():
            t_start = time.perf_counter()
            for k, (sc, ids) in enumerate(reqs):
                d = t_start + arr[k] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[k] = srv.submit(sc, ids, rw.rpdf, rw.vote_count_threshold)

        th = threading.Thread(target=submitter)
        t0 = time.perf_counter()
        th.start()
        lat = []
        for k in range(len(reqs)):
            while tickets[k] is None:
                time.sleep(0.0005)
            _, st, info, _ = srv.wait(tickets[k], 600.0)
            if info["rc"] != ppf_amd.PPF_OK:
                raise RuntimeError(f"served request failed: {info}")
            lat.append(info["latency_ms"])
        th.join()
        wall = time.perf_counter() - t0
        fin, plat = 0.0, []
        for a, sv in zip(arr, service):
            fin = max(fin, a) + sv
            plat.append((fin - a) * 1e3)
        served.append({"rate_req_per_s": rate, "registrations_per_s": n_targets / wall,
                       "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                       "plain_registrations_per_s": n_targets / fin,
                       "plain_p50_ms": float(np.percentile(plat, 50)),
                       "plain_p99_ms": float(np.percentile(plat, 99))})
        log(f"served at {rate} req/s: {wall:.2f}s")
    st = srv.stats()
    srv.close()
    for m in models:
        m.close()
    out["request_set"] = {
        "workload": desc,
        "plain": {"s": plain_s, "registrations_per_s": n_targets / plain_s, "lookups_per_s": lookups / plain_s},
        "batched": {"s": batch_s, "registrations_per_s": n_targets / batch_s,
                    "lookups_per_s": binfo["lookups"] / batch_s,
                    "speedup_vs_plain": plain_s / batch_s, "vote_launches": binfo["vote_launches"],
                    "vote_ms": binfo["vote_ms"], "finish_ms": binfo["finish_ms"]},
        "served": served,
        "server": {"rounds": st["rounds"], "max_requests_in_round": st["max_requests_in_round"],
                   "peak_inflight_bytes": st["peak_inflight_bytes"]},
    }
    out["large_hash_sample"] = run_large_hash(ctx)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--exact-votes", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the batched and served forms")
    ap.add_argument("--forms-only", action="store_true",
                    help="only the batched and served forms (no plain steps, no CPU baseline)")
    ap.add_argument("--large-hash-only", action="store_true",
                    help="only the large-hash stress workload (30,000-point model, 200,000-point scene)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 256 MiB device copy kernel 3x (reference for traffic counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        ctx = ppf_amd.Context(0)
        a = np.arange(1 << 26, dtype=np.uint32)
        for _ in range(3):
            b = ctx.debug_copy(a)
        assert np.array_equal(a, b)
        log(f"copied {a.nbytes} bytes x3")
        return
    if args.large_hash_only:
        print(json.dumps({"large_hash_sample": run_large_hash(ppf_amd.Context(0))}), flush=True)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(voting would shard by scene reference point)", file=sys.stderr)
        sys.exit(2)

    t0 = time.time()
    w = synth.dataset_like()
    log(f"workload generated in {time.time() - t0:.1f}s: models "
        f"{[len(m[0]) for m in w.models]}, scene {len(w.scenes[0][0])}")
    ctx = ppf_amd.Context(0)
    t0 = time.time()
    models = [ppf_amd.Model(m, d, w.vote_count_threshold, cpu_clustering=w.cpu_clustering, ctx=ctx)
              for m, d in zip(w.models, w.d_dists)]
    log(f"models built in {time.time() - t0:.2f}s")
    scene = ppf_amd.Scene(w.scenes[0], 0.0, w.rpdf, ctx=ctx)

    if args.forms_only:
        print(json.dumps({"forms": run_forms(ctx, models, scene, w)}), flush=True)
        return

    def step(record):
        for mdl in models:
            t = time.perf_counter()
            res = mdl.ppf_lookup(scene, exact_votes=args.exact_votes)
            res.pose()
            dt = time.perf_counter() - t
            if record is not None:
                record.append((dt, dict(res.info)))
            res.close()

    for i in range(args.warmup):
        step(None)
        log(f"warmup {i + 1}/{args.warmup} done")
    recs = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        step(recs)
        log(f"step {i + 1}/{args.steps} done")
    elapsed = time.perf_counter() - t_start

    lookups = sum(r["lookups"] for _, r in recs)
    votes = sum(r["votes"] for _, r in recs)
    vote_ms = sum(r["vote_ms"] for _, r in recs)
    lat = np.array([dt for dt, _ in recs]) * 1e3
    n_launch = len(recs)
    alg_bytes = 8.0 * lookups + 4.0 * votes
    achieved = alg_bytes / n_launch / (vote_ms / n_launch * 1e-3) / 1e9
    vote_traffic, traffic_src = hbm_traffic("k_vote<false>" if not args.exact_votes else "k_vote<true>")
    valu = valu_insts("k_vote<false>") if not args.exact_votes else None
    out = {
        "metric": METRIC,
        "value": lookups / elapsed,
        "unit": "pair-lookups/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "dataset_like: 4 seeded superquadric models (3k-8k points), one 50,000-point "
                        "scene with all four on a ground plane, ref_point_df 5, threshold 0.3, "
                        "GPU spatial-hash clustering, plain per-scene pipeline",
            "model_points": [len(m[0]) for m in w.models],
            "scene_points": len(w.scenes[0][0]),
            "ref_point_df": w.rpdf,
            "d_dist": w.d_dists,
            "form": "plain",
            "exact_votes": bool(args.exact_votes),
        },
        "registrations_per_s": n_launch / elapsed,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "votes_per_s": votes / elapsed,
        "lookups_per_step": lookups / args.steps,
        "votes_per_step": votes / args.steps,
        "vote_kernel_ms_per_launch": vote_ms / n_launch,
        "roofline": {
            "kernel": "k_vote",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": vote_traffic,
            "alg_bytes_per_launch": alg_bytes / n_launch,
            "traffic_source": traffic_src,
            # the counters' bytes over the same launch time: what HBM really moved
            "measured_gbs": (vote_traffic / (vote_ms / n_launch * 1e-3) / 1e9) if vote_traffic else None,
            "measured_frac": (vote_traffic / (vote_ms / n_launch * 1e-3) / 1e9 / HBM_PEAK_GBS)
                             if vote_traffic else None,
            # what bounds the kernel: its share of the chip's VALU issue slots
            # (recorded instruction count x 4 cycles per wave64 instruction
            # over 1,024 SIMDs at 2.4 GHz for the launch time measured here)
            "valu_issue_share": (valu * 4.0 / (1024 * 2.4e9 * vote_ms / n_launch * 1e-3)) if valu else None,
            "valu_source": "profiles/dataset_like_sq_counters.md" if valu else None,
        },
    }
    if not args.no_forms:
        out["forms"] = run_forms(ctx, models, scene, w)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
