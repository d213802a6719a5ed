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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--exact-votes", action="store_true")
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
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(voting would shard by scene reference point; see DESIGN.md)", file=sys.stderr)
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
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
