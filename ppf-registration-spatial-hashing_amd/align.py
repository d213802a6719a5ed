"""Registration driver with the reference's command-line flags
(src/alignment.cpp:119-158, flow :196-331): load PLY scenes and models,
d_dist = tau_d * the model's largest bounding-box extent, voxel-grid
downsample each scene with scene_leaf_size and each model with its d_dist,
register every model in every scene on the device, and, given .xf ground
truth, print 1 or 0 per (scene, model) for a pose within
validation_translation_threshold * diameter and
validation_rotation_threshold degrees (ht_dist).

    python align.py --scene_files s.ply --model_files m1.ply,m2.ply \\
        --tau_d 0.05,0.05 --scene_leaf_size 0.01 --ref_point_df 5 \\
        --validation_files s_m1.xf,s_m2.xf

Boolean flags take a value (--cpu_clustering 1), as boost::program_options
reads them; --show_normals and --visualize are accepted and ignored (no
viewer), --logfile/--loglevel set Python logging.
"""
from __future__ import annotations

import argparse
import logging
import math
import os
import sys
from typing import List, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import ppf_amd  # noqa: E402

LOG = logging.getLogger("align")


def _bool(v: str) -> bool:
    s = str(v).strip().lower()
    if s in ("1", "true", "yes", "on"):
        return True
    if s in ("0", "false", "no", "off"):
        return False
    raise argparse.ArgumentTypeError(f"not a boolean: {v}")


def _csv(v: str) -> List[str]:
    return [x for x in v.split(",") if x]


def parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0],
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--dev", type=int, default=1, help="device to use (min(n-1, dev))")
    ap.add_argument("--logfile", type=str, default=None)
    ap.add_argument("--loglevel", type=str, default="info")
    ap.add_argument("--tau_d", type=_csv, required=True, help="d_dist factors, one per model")
    ap.add_argument("--scene_leaf_size", type=float, default=10.0)
    ap.add_argument("--ref_point_df", type=int, default=1)
    ap.add_argument("--vote_count_threshold", type=float, default=0.4)
    ap.add_argument("--cpu_clustering", type=_bool, default=False)
    ap.add_argument("--use_l1_norm", type=_bool, default=False)
    ap.add_argument("--use_averaged_clusters", type=_bool, default=False)
    ap.add_argument("--validation_translation_threshold", type=float, default=0.1)
    ap.add_argument("--validation_rotation_threshold", type=float, default=12.0)
    ap.add_argument("--scene_files", type=_csv, required=True)
    ap.add_argument("--model_files", type=_csv, required=True)
    ap.add_argument("--validation_files", type=_csv, default=None)
    ap.add_argument("--show_normals", type=_bool, default=True)
    ap.add_argument("--visualize", type=_bool, default=True)
    return ap


def run(argv: Sequence[str], out=sys.stdout) -> dict:
    a = parser().parse_args(list(argv))
    logging.basicConfig(filename=a.logfile, level=getattr(logging, a.loglevel.upper(), logging.INFO),
                        format="%(asctime)s %(levelname)s %(message)s")
    LOG.info(" ".join(argv))
    if len(a.tau_d) != len(a.model_files):
        raise SystemExit("Each model must have an associated tau_d.")
    tau_d = [float(x) for x in a.tau_d]
    ctx = ppf_amd.Context(a.dev)
    scenes = []
    for f in a.scene_files:
        LOG.info("Loading scene point cloud: %s", f)
        scenes.append(ppf_amd.read_ply(f))
    models, d_dists = [], []
    for f, t in zip(a.model_files, tau_d):
        LOG.info("Loading model point cloud: %s", f)
        m = ppf_amd.read_ply(f)
        ext = ppf_amd.bbox_extent(m[0])
        d_dists.append(float(np.float32(t) * np.float32(ext.max())))
        models.append(m)
        LOG.debug("model_diam (%f, %f, %f), d_dist %f", *ext, d_dists[-1])
    LOG.info("Downsampling...")
    scenes = [ppf_amd.voxel_downsample(s, a.scene_leaf_size, ctx=ctx) for s in scenes]
    models = [ppf_amd.voxel_downsample(m, d, ctx=ctx) for m, d in zip(models, d_dists)]
    for s in scenes:
        LOG.info("Scene size after filtering: %u", len(s[0]))
    poses, status = ppf_amd.ppf_registration(scenes, models, d_dists, a.ref_point_df,
                                             a.vote_count_threshold, a.cpu_clustering, a.use_l1_norm,
                                             a.use_averaged_clusters, a.dev, ctx=ctx, return_status=True)
    result = {"poses": poses, "status": status, "d_dists": d_dists, "matches": None}
    if a.validation_files:
        if len(a.validation_files) != len(scenes) * len(models):
            raise SystemExit("need one validation file per (scene, model)")
        matches = []
        for i in range(len(scenes)):
            for j in range(len(models)):
                truth = ppf_amd.read_xf(a.validation_files[i * len(models) + j])
                dt, dr = ppf_amd.ht_dist(poses[i][j], truth)
                diam = d_dists[j] / tau_d[j]
                tt = a.validation_translation_threshold * diam
                rt = math.radians(a.validation_rotation_threshold)
                match = dt < tt and dr < rt
                LOG.info("Transformations for %s in %s: estimated %s, ground truth %s",
                         a.model_files[j], a.scene_files[i], poses[i][j].tolist(), truth.tolist())
                LOG.info("Distance (trans, rot): %f, %f; thresholds %f, %f", dt, dr, tt, rt)
                print(int(match), file=out)
                matches.append(bool(match))
        result["matches"] = matches
    ctx.close()
    return result


if __name__ == "__main__":
    run(sys.argv[1:])
