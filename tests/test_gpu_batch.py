"""Batched form (ppf_batch_lookup): every job's result is bit-identical to
the plain Model::ppf_lookup of the same (model, scene) pair (the reference's
per-pair loop, src/cuda/ppf.cu:63-98), for uneven scene and model sizes."""
import numpy as np
import pytest

import ppf_amd
import synth

pytestmark = pytest.mark.gpu


def _small_batch(seed=21):
    return synth.batched(seed=seed, n_models=4, n_scenes=5, model_range=(300, 1200),
                         scene_range=(800, 3000), models_per_scene=(1, 3))


def _same(a: ppf_amd.LookupResult, b: ppf_amd.LookupResult, gpu_clustering: bool):
    for k in ("status", "lookups", "votes", "max_count", "n_kept", "max_idx", "n_clusters"):
        assert a.info[k] == b.info[k], k
    ca, na = a.votes()
    cb, nb = b.votes()
    assert np.array_equal(ca, cb) and np.array_equal(na, nb)
    assert a.transformations().tobytes() == b.transformations().tobytes()
    if gpu_clustering and a.n_kept:
        for x, y in zip(a.clusters(), b.clusters()):
            assert x.tobytes() == y.tobytes()
    assert a.pose().tobytes() == b.pose().tobytes()


@pytest.mark.parametrize("cpu_clustering", [False, True])
def test_batch_equals_plain(ctx, cpu_clustering):
    w = _small_batch()
    models = [ppf_amd.Model(m, d, w.vote_count_threshold, cpu_clustering=cpu_clustering, ctx=ctx)
              for m, d in zip(w.models, w.d_dists)]
    scenes = [ppf_amd.Scene(s, 0.0, w.rpdf, ctx=ctx) for s in w.scenes]
    jm, js = [], []
    for si, ids in enumerate(w.scene_models):
        for mi in ids:
            jm.append(models[mi])
            js.append(scenes[si])
    # every model also against the first scene (models absent from it too)
    for m in models:
        jm.append(m)
        js.append(scenes[0])
    res, info = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, cpu_clustering=cpu_clustering,
                                     ctx=ctx)
    assert info["n_jobs"] == len(jm) and info["vote_launches"] >= 1
    assert info["lookups"] == sum(r.info["lookups"] for r in res)
    for m, s, r in zip(jm, js, res):
        plain = m.ppf_lookup(s)
        _same(r, plain, not cpu_clustering)
        plain.close()
        r.close()
    for o in models + scenes:
        o.close()


def test_batch_single_job_and_duplicates(ctx):
    w = synth.plain_small(seed=3, n_model=300, n_clutter=400)
    m = ppf_amd.Model(w.models[0], w.d_dists[0], w.vote_count_threshold, ctx=ctx)
    s = ppf_amd.Scene(w.scenes[0], 0.0, w.rpdf, ctx=ctx)
    one, _ = ppf_amd.batch_lookup([m], [s], w.vote_count_threshold, ctx=ctx)
    three, info = ppf_amd.batch_lookup([m, m, m], [s, s, s], w.vote_count_threshold, ctx=ctx)
    assert info["n_jobs"] == 3
    for r in three:
        _same(r, one[0], True)
    for r in one + three:
        r.close()
    m.close()
    s.close()


def test_batch_exact_votes_equal_fast(ctx):
    w = _small_batch(seed=22)
    models = [ppf_amd.Model(m, d, w.vote_count_threshold, ctx=ctx) for m, d in zip(w.models, w.d_dists)]
    scenes = [ppf_amd.Scene(s, 0.0, w.rpdf, ctx=ctx) for s in w.scenes]
    jm = [models[mi] for si, ids in enumerate(w.scene_models) for mi in ids]
    js = [scenes[si] for si, ids in enumerate(w.scene_models) for _ in ids]
    fast, _ = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, ctx=ctx)
    exact, _ = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, exact_votes=True, ctx=ctx)
    for a, b in zip(fast, exact):
        _same(a, b, True)
    for o in fast + exact + models + scenes:
        o.close()


def test_batch_per_reference_argmax_equals_plain(ctx):
    w = _small_batch(seed=23)
    models = [ppf_amd.Model(m, d, w.vote_count_threshold, ctx=ctx) for m, d in zip(w.models, w.d_dists)]
    scenes = [ppf_amd.Scene(s, 0.0, w.rpdf, ctx=ctx) for s in w.scenes]
    jm = [models[mi] for si, ids in enumerate(w.scene_models) for mi in ids]
    js = [scenes[si] for si, ids in enumerate(w.scene_models) for _ in ids]
    res, _ = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, per_reference_argmax=True, ctx=ctx)
    for m, s, r in zip(jm, js, res):
        plain = m.ppf_lookup(s, per_reference_argmax=True)
        _same(r, plain, True)
        full = m.ppf_lookup(s)
        # one kept bin per reference row at most, each one of the full list's
        # bins with the same count
        c, n = r.votes()
        fc, fn = full.votes()
        assert len(np.unique(c >> np.uint64(32))) == len(c)
        assert len(c) <= len(fc)
        assert set(zip(c.tolist(), n.tolist())) <= set(zip(fc.tolist(), fn.tolist()))
        for o in (plain, full, r):
            o.close()
    for o in models + scenes:
        o.close()


def test_batch_rejects_foreign_objects(ctx):
    other = ppf_amd.Context(0)
    try:
        w = synth.plain_small(seed=5, n_model=200, n_clutter=100)
        m = ppf_amd.Model(w.models[0], w.d_dists[0], 0.3, ctx=ctx)
        s = ppf_amd.Scene(w.scenes[0], 0.0, 5, ctx=other)
        with pytest.raises(ppf_amd.PPFError) as e:
            ppf_amd.batch_lookup([m], [s], 0.3, ctx=ctx)
        assert e.value.code == ppf_amd.PPF_ERR_ARG
        m.close()
    finally:
        other.close()


def test_autotuner_choice_keeps_results(ctx):
    """ppf_batch_tune picks launch shapes by timing them on the batch; the
    batch's results afterwards equal the default shapes' and the plain
    lookups', bit for bit."""
    w = synth.batched(seed=21, n_models=3, n_scenes=4, model_range=(300, 900),
                      scene_range=(1500, 4000), models_per_scene=(1, 2))
    c = ppf_amd.Context(0)
    try:
        models = [ppf_amd.Model(m, d, w.vote_count_threshold, ctx=c) for m, d in zip(w.models, w.d_dists)]
        scenes = [ppf_amd.Scene(s, 0.0, w.rpdf, ctx=c) for s in w.scenes]
        jm = [models[mi] for ids in w.scene_models for mi in ids]
        js = [scenes[k] for k, ids in enumerate(w.scene_models) for _ in ids]
        base, _ = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, ctx=c)
        choice = ppf_amd.batch_tune(jm, js, w.vote_count_threshold, ctx=c)
        print("tuned:", choice, flush=True)
        assert choice["vote_blocks_per_cu"] in (1, 2) and choice["finish_lanes"] in (1, 2, 4)
        assert choice["sort_tile_keys"] in (2048, 4096) and choice["sort_tile_threads"] in (256, 512)
        assert choice["tune_ms"] > 0
        tuned, info = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, ctx=c)
        assert info["lanes"] == min(choice["finish_lanes"], len(jm))
        for a, b, m, s in zip(base, tuned, jm, js):
            ca, na = a.votes()
            cb, nb = b.votes()
            assert np.array_equal(ca, cb) and np.array_equal(na, nb)
            assert a.pose().tobytes() == b.pose().tobytes()
            p = m.ppf_lookup(s)
            assert p.pose().tobytes() == b.pose().tobytes()
        ppf_amd.tune_reset(c)
    finally:
        c.close()
