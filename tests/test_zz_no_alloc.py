"""The allocation-free lookup (ppf_lookup_noalloc, through the C ABI) gives
the pose and statistics of ppf_lookup on the same inputs, on consecutive
calls whose inputs shrink (the scratch array keeps the previous call's
contents), for GPU and CPU clustering; oversized scenes are rejected on the
host with a clear error."""
import numpy as np
import pytest

import ppf_amd
import synth

KEYS = ("status", "lookups", "votes", "max_count", "n_kept", "max_idx", "n_clusters")


def _case(seed, n_model, n_clutter):
    w = synth.plain_small(seed=seed, n_model=n_model, n_clutter=n_clutter)
    return w.models[0], w.scenes[0], w.d_dists[0]


@pytest.mark.gpu
@pytest.mark.parametrize("cpu", [False, True])
def test_noalloc_matches_lookup_on_shrinking_calls(cpu):
    ctx = ppf_amd.Context(0)
    try:
        cases = [_case(41, 900, 2500), _case(42, 400, 700), _case(43, 700, 1500)]
        for k, (m, s, d) in enumerate(cases):
            print(f"noalloc case {k}: model {len(m[0])} scene {len(s[0])} cpu {cpu}", flush=True)
            model = ppf_amd.Model(m, d, 0.3, cpu_clustering=cpu, ctx=ctx)
            scene = ppf_amd.Scene(s, d, 5, ctx=ctx)
            ref = model.ppf_lookup(scene)
            want_pose, want = ref.pose(), ref.info
            pose, info = model.ppf_lookup_noalloc(scene)
            assert pose.tobytes() == want_pose.astype(np.float32).tobytes(), k
            for key in KEYS:
                assert info[key] == want[key], (k, key, info[key], want[key])
            assert info["votes"] > 0 and info["n_kept"] > 0
            # and again: the same bits from a scratch array that now holds
            # this call's own leftovers
            pose2, info2 = model.ppf_lookup_noalloc(scene)
            assert pose2.tobytes() == pose.tobytes() and info2["n_kept"] == info["n_kept"]
            ref.close()
            scene.close()
            model.close()
    finally:
        ctx.close()


@pytest.mark.gpu
def test_noalloc_rejects_oversized_scene():
    ctx = ppf_amd.Context(0)
    try:
        m, _, d = _case(44, 300, 100)
        rng = np.random.default_rng(5)
        n = ppf_amd.NOALLOC_MAX_SCENE_POINTS + 1
        pts = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
        nrm = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
        model = ppf_amd.Model(m, d, 0.3, ctx=ctx)
        scene = ppf_amd.Scene((pts, nrm), d, 5, ctx=ctx)
        with pytest.raises(ppf_amd.PPFError) as e:
            model.ppf_lookup_noalloc(scene)
        assert e.value.code == ppf_amd.PPF_ERR_ARG and "exceeds the limit" in str(e.value)
    finally:
        ctx.close()
