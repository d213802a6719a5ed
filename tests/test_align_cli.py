"""The registration driver with the reference's flags (src/alignment.cpp:
119-158, 196-331): flag parsing on the CPU; on the GPU, PLY files in,
downsampling, registration and the ht_dist validation against .xf files."""
import io
import os

import numpy as np
import pytest

import align
import ppf_amd
import synth


def test_reference_flags_parse():
    a = align.parser().parse_args(
        ["--scene_files", "a.ply,b.ply", "--model_files", "m.ply", "--tau_d", "0.05",
         "--scene_leaf_size", "0.02", "--ref_point_df", "5", "--vote_count_threshold", "0.3",
         "--cpu_clustering", "1", "--use_l1_norm", "false", "--use_averaged_clusters", "0",
         "--validation_files", "a.xf,b.xf", "--validation_translation_threshold", "0.1",
         "--validation_rotation_threshold", "12", "--dev", "0", "--show_normals", "0",
         "--visualize", "0"])
    assert a.scene_files == ["a.ply", "b.ply"] and a.model_files == ["m.ply"]
    assert a.tau_d == ["0.05"] and a.ref_point_df == 5 and a.cpu_clustering is True
    assert a.use_l1_norm is False and a.validation_files == ["a.xf", "b.xf"]
    d = align.parser().parse_args(["--scene_files", "s", "--model_files", "m", "--tau_d", "1"])
    # the reference's defaults
    assert (d.dev, d.scene_leaf_size, d.ref_point_df, d.vote_count_threshold) == (1, 10.0, 1, 0.4)
    assert (d.validation_translation_threshold, d.validation_rotation_threshold) == (0.1, 12.0)
    with pytest.raises(SystemExit):
        align.parser().parse_args(["--scene_files", "s"])  # model_files and tau_d are required


def _bumpy_workload(seed):
    """A superquadric with 8 Gaussian bumps (no rotational symmetry) under a
    random pose, plus 2,000 clutter points."""
    rng = np.random.default_rng(seed)
    sq = synth.random_superquadric(rng, bumps=8)
    mp, mn = synth.as_cloud(*sq.sample(2000, rng))
    R = synth.random_rotation(rng)
    t = rng.uniform(-1, 1, 3)
    sp, sn = synth.transform(mp.astype(np.float64), mn.astype(np.float64), R, t)
    D = synth.bbox_diameter(mp)
    cp = rng.uniform(-1.5 * D, 1.5 * D, (2000, 3))
    cn = rng.normal(size=(2000, 3))
    cn /= np.linalg.norm(cn, axis=1, keepdims=True)
    scene = (np.vstack([sp, cp]).astype(np.float32), np.vstack([sn, cn]).astype(np.float32))
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = t
    return (mp, mn), scene, T


@pytest.mark.gpu
@pytest.mark.parametrize("cpu_clustering", [False, True])
def test_driver_registers_and_validates(tmp_path, cpu_clustering):
    # seed 3: the oracle recovers this instance with both clustering modes
    # (3.1 and 2.5 degrees); not every instance is recovered by the algorithm
    model, scene, T = _bumpy_workload(3)
    sf, mf, xf = str(tmp_path / "s.ply"), str(tmp_path / "m.ply"), str(tmp_path / "s_m.xf")
    ppf_amd.write_ply(sf, *scene)
    ppf_amd.write_ply(mf, *model)
    with open(xf, "w") as f:
        f.write("\n".join(" ".join(f"{v:.9g}" for v in row) for row in T) + "\n")
    D = float(ppf_amd.bbox_extent(model[0]).max())
    out = io.StringIO()
    r = align.run(["--scene_files", sf, "--model_files", mf, "--tau_d", "0.05",
                   "--scene_leaf_size", repr(0.01 * D), "--ref_point_df", "5",
                   "--vote_count_threshold", "0.3", "--cpu_clustering", str(int(cpu_clustering)),
                   "--validation_files", xf, "--dev", "0"], out=out)
    assert r["status"][0][0] == ppf_amd.PPF_OK
    assert out.getvalue().strip() == "1" and r["matches"] == [True]
