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
        assert choice["vote_blocks_per_cu"] in (1, 2) and choice["finish_lanes"] in (0, 1, 2, 4)
        assert choice["sort_tile_keys"] in (2048, 4096, 8192) and choice["sort_tile_threads"] in (256, 512)
        assert choice["tune_ms"] > 0
        tuned, info = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, ctx=c)
        if choice["finish_lanes"] == 0:
            assert info["fused"] == 1 and info["lanes"] == 0
        else:
            assert info["fused"] == 0 and info["lanes"] == min(choice["finish_lanes"], len(jm))
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


def _jobs(w, models, scenes):
    jm = [models[mi] for ids in w.scene_models for mi in ids]
    js = [scenes[k] for k, ids in enumerate(w.scene_models) for _ in ids]
    return jm, js


def _stats_same(a, b):
    for k in ("cluster_tests", "cluster_mixed_chunks"):
        assert a.info[k] == b.info[k], k


@pytest.mark.parametrize("use_l1,use_avg", [(False, False), (True, False), (False, True), (True, True)])
def test_fused_finishing_equals_lanes(ctx, use_l1, use_avg):
    """The batch finished together (one threshold/order sort, one pose launch,
    one clustering launch over every job's chunks, one argmax) gives each job
    the same kept votes, poses, clustering and final pose, bit for bit, as
    the job-by-job finishing on lanes and as the plain lookup."""
    w = _small_batch(seed=24)
    models = [ppf_amd.Model(m, d, w.vote_count_threshold, use_l1_norm=use_l1, use_averaged_clusters=use_avg,
                            ctx=ctx) for m, d in zip(w.models, w.d_dists)]
    scenes = [ppf_amd.Scene(s, 0.0, w.rpdf, ctx=ctx) for s in w.scenes]
    jm, js = _jobs(w, models, scenes)
    kw = dict(use_l1_norm=use_l1, use_averaged_clusters=use_avg, ctx=ctx)
    fused, fi = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, **kw)
    ctx.debug_set(3, 1)
    try:
        lanes, li = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, **kw)
    finally:
        ctx.debug_set(3, 0)
    assert fi["fused"] == 1 and fi["lanes"] == 0 and li["fused"] == 0 and li["lanes"] >= 1
    print(f"fused finishing {fi['finish_ms']:.1f} ms, lanes {li['finish_ms']:.1f} ms", flush=True)
    for m, s, a, b in zip(jm, js, fused, lanes):
        _same(a, b, True)
        _stats_same(a, b)
        plain = m.ppf_lookup(s)
        _same(a, plain, True)
        plain.close()
    for o in fused + lanes + models + scenes:
        o.close()


def test_vote_graph_replays_equal_direct_launches():
    """The batched vote stage captured as HIP graphs: the first call captures,
    repeated calls replay (with plain lookups and another batch, which
    re-captures, in between), and every replay gives the results of the
    directly launched vote stage and of the plain lookups, bit for bit."""
    c = ppf_amd.Context(0)
    try:
        w = _small_batch(seed=26)
        models = [ppf_amd.Model(m, d, w.vote_count_threshold, ctx=c) for m, d in zip(w.models, w.d_dists)]
        scenes = [ppf_amd.Scene(s, 0.0, w.rpdf, ctx=c) for s in w.scenes]
        jm, js = _jobs(w, models, scenes)
        c.debug_set(4, 0)
        direct, di = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, ctx=c)
        c.debug_set(4, 1)
        assert di["vote_graph"] == 0
        plain = [m.ppf_lookup(s) for m, s in zip(jm, js)]
        for m, s, a, p in zip(jm, js, direct, plain):
            _same(a, p, True)
        calls = []
        for rep in range(4):
            if rep == 2:  # another batch in between: a different launch signature
                other, oi = ppf_amd.batch_lookup(jm[:2], js[:2], w.vote_count_threshold, ctx=c)
                assert oi["vote_graph"] == 1
                for m, s, r in zip(jm[:2], js[:2], other):
                    q = m.ppf_lookup(s)
                    _same(r, q, True)
                    q.close()
                for o in other:
                    o.close()
            res, info = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, ctx=c)
            assert info["vote_graph"] == 1
            for a, b in zip(res, direct):
                _same(a, b, True)
            # eager work between replays
            q = jm[0].ppf_lookup(js[0])
            _same(q, plain[0], True)
            q.close()
            calls.append(res)
        for o in direct + plain + [r for res in calls for r in res] + models + scenes:
            o.close()
    finally:
        c.close()


def test_batch_vote_progress_slices_equal_one_launch():
    """With progress reports, ppf_batch_lookup votes its jobs' rows in
    launches of about rows_per_launch rows (a job's slices share its
    candidates and running maximum): the same results as one launch, and a
    report after each launch up to all rows."""
    c = ppf_amd.Context(0)
    try:
        w = _small_batch(seed=27)
        models = [ppf_amd.Model(m, d, w.vote_count_threshold, ctx=c) for m, d in zip(w.models, w.d_dists)]
        scenes = [ppf_amd.Scene(s, 0.0, w.rpdf, ctx=c) for s in w.scenes]
        jm, js = _jobs(w, models, scenes)
        one, oi = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, ctx=c)
        reports = []
        c.set_progress(lambda st, done, tot: reports.append((st, done, tot)), 37)
        try:
            sliced, si = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, ctx=c)
        finally:
            c.set_progress(None)
        votes = [r for r in reports if r[0] == "vote"]
        assert si["vote_launches"] > 1 and oi["vote_launches"] == 1
        assert len(votes) == si["vote_launches"]
        assert all(a[1] < b[1] for a, b in zip(votes, votes[1:]))
        assert votes[-1][1] == votes[-1][2]
        for a, b in zip(sliced, one):
            _same(a, b, True)
        for o in one + sliced + models + scenes:
            o.close()
    finally:
        c.close()


def test_fused_finishing_with_empty_jobs(ctx):
    """Jobs without votes (a scene whose pairs match no model key) among jobs
    with votes, and a batch where no job keeps a vote: same status and
    results as the plain lookups."""
    w = _small_batch(seed=25)
    models = [ppf_amd.Model(m, d, w.vote_count_threshold, ctx=ctx) for m, d in zip(w.models, w.d_dists)]
    scenes = [ppf_amd.Scene(s, 0.0, w.rpdf, ctx=ctx) for s in w.scenes]
    far = (np.array([[0, 0, 0], [1e4, 0, 0]], np.float32), np.array([[0, 0, 1], [0, 1, 0]], np.float32))
    empty = ppf_amd.Scene(far, 0.0, 1, ctx=ctx)
    jm, js = _jobs(w, models, scenes)
    jm = [models[0]] + jm[:3] + [models[1]]
    js = [empty] + js[:3] + [empty]
    res, info = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, ctx=ctx)
    assert info["fused"] == 1
    assert res[0].info["n_kept"] == 0 and res[-1].info["n_kept"] == 0
    for m, s, r in zip(jm, js, res):
        plain = m.ppf_lookup(s)
        _same(r, plain, True)
        plain.close()
    # a threshold no count exceeds: every job keeps nothing
    none, info = ppf_amd.batch_lookup(jm[1:4], js[1:4], 1.0, ctx=ctx)
    assert info["fused"] == 1
    for m, s, r in zip(jm[1:4], js[1:4], none):
        m.vote_count_threshold = 1.0
        plain = m.ppf_lookup(s)
        m.vote_count_threshold = w.vote_count_threshold
        assert r.info["n_kept"] == 0
        _same(r, plain, True)
        plain.close()
    for o in res + none + models + scenes + [empty]:
        o.close()


def test_batched_configuration_extremes_equal_plain(ctx, batched_workload):
    """The batched configuration's largest scene (about 120,000 points) with
    its smallest model, its largest model (about 20,000 points) with the
    smallest scene, and the smallest pair: batched (finished together)
    equals plain, bit for bit."""
    w = batched_workload
    ns = [len(s[0]) for s in w.scenes]
    nm = [len(m[0]) for m in w.models]
    big_s, small_s = int(np.argmax(ns)), int(np.argmin(ns))
    big_m, small_m = int(np.argmax(nm)), int(np.argmin(nm))
    pairs = [(big_s, small_m), (small_s, big_m), (small_s, small_m)]
    print("extreme jobs:", [(ns[s], nm[m]) for s, m in pairs], flush=True)
    models = {m: ppf_amd.Model(w.models[m], w.d_dists[m], w.vote_count_threshold, ctx=ctx)
              for m in {m for _, m in pairs}}
    scenes = {s: ppf_amd.Scene(w.scenes[s], 0.0, w.rpdf, ctx=ctx) for s in {s for s, _ in pairs}}
    jm = [models[m] for _, m in pairs]
    js = [scenes[s] for s, _ in pairs]
    res, info = ppf_amd.batch_lookup(jm, js, w.vote_count_threshold, ctx=ctx)
    print(f"batched: vote {info['vote_ms']:.0f} ms, finish {info['finish_ms']:.0f} ms, fused {info['fused']}",
          flush=True)
    for m, s, r in zip(jm, js, res):
        plain = m.ppf_lookup(s)
        _same(r, plain, True)
        _stats_same(r, plain)
        plain.close()
        r.close()
    for o in list(models.values()) + list(scenes.values()):
        o.close()
