"""Served form (ppf_server_*): each request's result is the same alone and in
a mix of concurrent requests, and equal to the plain Model::ppf_lookup of the
same (model, scene) pair (src/cuda/model.cu:269-306): vote codes and counts,
transformations, clustering outputs and final poses bit for bit."""
import numpy as np
import pytest

import ppf_amd
import synth

pytestmark = pytest.mark.gpu


def _workload(seed=31):
    return synth.batched(seed=seed, n_models=4, n_scenes=6, model_range=(300, 1200),
                         scene_range=(800, 4000), models_per_scene=(1, 3))


def _digest(res: ppf_amd.LookupResult, gpu: bool):
    c, n = res.votes()
    out = [res.info["status"], res.info["votes"], res.info["max_count"], res.info["n_kept"],
           res.info["max_idx"], c.tobytes(), n.tobytes(), res.transformations().tobytes(),
           res.pose().tobytes()]
    if gpu and res.n_kept:
        out += [x.tobytes() for x in res.clusters()]
    return out


@pytest.mark.parametrize("cpu_clustering", [False, True])
def test_alone_equals_mix_equals_plain(ctx, cpu_clustering):
    w = _workload()
    models = [ppf_amd.Model(m, d, w.vote_count_threshold, cpu_clustering=cpu_clustering, ctx=ctx)
              for m, d in zip(w.models, w.d_dists)]
    srv = ppf_amd.Server(models, round_ms=2.0, ctx=ctx)
    try:
        reqs = [(s, ids) for s, ids in zip(w.scenes, w.scene_models)]
        alone = []
        for s, ids in reqs:
            t = srv.submit(s, ids, w.rpdf, w.vote_count_threshold, cpu_clustering=cpu_clustering)
            poses, st, info, res = srv.wait(t, 120.0, results=True)
            assert info["rc"] == ppf_amd.PPF_OK, info
            alone.append((poses, st, [_digest(r, not cpu_clustering) for r in res]))
            for r in res:
                r.close()
        tickets = [srv.submit(s, ids, w.rpdf, w.vote_count_threshold, cpu_clustering=cpu_clustering)
                   for s, ids in reqs]
        for (s, ids), t, (poses_a, st_a, dig_a) in zip(reqs, tickets, alone):
            poses, st, info, res = srv.wait(t, 120.0, results=True)
            assert info["rc"] == ppf_amd.PPF_OK, info
            assert poses.tobytes() == poses_a.tobytes()
            assert np.array_equal(st, st_a)
            assert [_digest(r, not cpu_clustering) for r in res] == dig_a
            for r in res:
                r.close()
        stats = srv.stats()
        assert stats["completed"] == 2 * len(reqs) and stats["failed"] == 0
        assert stats["max_requests_in_round"] >= 2  # the mix shared vote rounds
        assert stats["inflight_bytes"] == 0
        # alone == plain lookups of the same pairs
        for (s, ids), (poses_a, st_a, dig_a) in zip(reqs, alone):
            sc = ppf_amd.Scene(s, 0.0, w.rpdf, ctx=ctx)
            for k, mi in enumerate(ids):
                r = models[mi].ppf_lookup(sc)
                assert _digest(r, not cpu_clustering) == dig_a[k]
                r.close()
            sc.close()
    finally:
        srv.close()
        for m in models:
            m.close()


def test_budget_cancel_and_errors(ctx):
    w = _workload(seed=32)
    models = [ppf_amd.Model(m, d, w.vote_count_threshold, ctx=ctx) for m, d in zip(w.models, w.d_dists)]
    small = synth.plain_small(seed=9, n_model=200, n_clutter=300).scenes[0]
    big = w.scenes[int(np.argmax([len(s[0]) for s in w.scenes]))]
    # budget: one single-target small request fits, the big one never does
    n_small = len(small[0])
    R_small = (n_small + 4) // 5
    need_small = n_small * 28 + R_small * 64 + 512 + max(1 << 23, R_small * 4096) * 12 + 512
    srv = ppf_amd.Server(models, hbm_budget_bytes=need_small + 1024, round_ms=2.0, ctx=ctx)
    try:
        tb = srv.submit(big, [0], 5, w.vote_count_threshold)
        ts = srv.submit(small, [1], 5, w.vote_count_threshold)
        _, _, info_b, _ = srv.wait(tb, 60.0)
        assert info_b["rc"] == ppf_amd.PPF_ERR_NOMEM and "budget" in info_b["error"]
        poses, st, info_s, _ = srv.wait(ts, 60.0)
        assert info_s["rc"] == ppf_amd.PPF_OK and info_s["vote_rounds"] >= 1
        assert set(st.tolist()) <= {ppf_amd.PPF_OK, ppf_amd.PPF_ERR_NO_VOTES}
        # cancel: the second of two small requests waits for budget
        t1 = srv.submit(small, [2], 5, w.vote_count_threshold)
        t2 = srv.submit(small, [3], 5, w.vote_count_threshold)
        rc = srv.cancel(t2)
        _, _, i1, _ = srv.wait(t1, 60.0)
        _, _, i2, _ = srv.wait(t2, 60.0)
        assert i1["rc"] == ppf_amd.PPF_OK
        if rc == ppf_amd.PPF_OK:
            assert i2["rc"] == ppf_amd.PPF_ERR_CANCELLED
        else:  # already admitted when cancel ran
            assert rc == ppf_amd.PPF_ERR_STATE and i2["rc"] == ppf_amd.PPF_OK
        with pytest.raises(ppf_amd.PPFError):
            srv.submit(small, [7], 5, 0.3)  # model id out of range
        st = srv.stats()
        assert st["rejected"] == 1 and st["inflight_bytes"] == 0
        assert st["peak_inflight_bytes"] <= st["budget_bytes"]
    finally:
        srv.close()
        for m in models:
            m.close()


def test_mixed_per_reference_argmax_requests(ctx):
    """Requests with and without the per-reference score pass share vote
    rounds; each one's result equals the plain lookup with its own flag."""
    w = _workload(seed=33)
    models = [ppf_amd.Model(m, d, w.vote_count_threshold, ctx=ctx) for m, d in zip(w.models, w.d_dists)]
    srv = ppf_amd.Server(models, round_ms=2.0, ctx=ctx)
    try:
        reqs = [(s, ids, k % 2 == 0) for k, (s, ids) in enumerate(zip(w.scenes, w.scene_models))]
        tickets = [srv.submit(s, ids, w.rpdf, w.vote_count_threshold, per_reference_argmax=pr)
                   for s, ids, pr in reqs]
        for (s, ids, pr), t in zip(reqs, tickets):
            _, _, info, res = srv.wait(t, 120.0, results=True)
            assert info["rc"] == ppf_amd.PPF_OK, info
            sc = ppf_amd.Scene(s, 0.0, w.rpdf, ctx=ctx)
            for mi, r in zip(ids, res):
                plain = models[mi].ppf_lookup(sc, per_reference_argmax=pr)
                assert _digest(r, True) == _digest(plain, True)
                plain.close()
                r.close()
            sc.close()
        assert srv.stats()["failed"] == 0
    finally:
        srv.close()
        for m in models:
            m.close()


def test_grouped_finishing_equals_plain(ctx):
    """One finishing lane: jobs that leave voting while the lane is busy queue
    up and are finished together (one launch per stage for the group); jobs
    with other clustering flags are never grouped with them.  Every result
    equals the plain lookup of its pair, bit for bit."""
    w = _workload()
    models = {}
    srv_models = [ppf_amd.Model(m, d, w.vote_count_threshold, ctx=ctx) for m, d in zip(w.models, w.d_dists)]
    srv = ppf_amd.Server(srv_models, lanes=1, round_ms=1.0, ctx=ctx)
    try:
        reqs = [(s, ids, k % 3 == 2) for k, (s, ids) in enumerate(zip(w.scenes, w.scene_models))]
        reqs = reqs + reqs
        tickets = [srv.submit(s, ids, w.rpdf, w.vote_count_threshold, use_l1_norm=l1) for s, ids, l1 in reqs]
        got = []
        for t in tickets:
            poses, st, info, res = srv.wait(t, 120.0, results=True)
            assert info["rc"] == ppf_amd.PPF_OK, info
            got.append([_digest(r, True) for r in res])
            for r in res:
                r.close()
        stats = srv.stats()
        print("finish groups:", stats["finish_groups"], "largest:", stats["max_finish_group"], flush=True)
        assert stats["completed"] == len(reqs) and stats["failed"] == 0
        assert stats["max_finish_group"] >= 2
        for (s, ids, l1), dig in zip(reqs, got):
            sc = ppf_amd.Scene(s, 0.0, w.rpdf, ctx=ctx)
            for k, mi in enumerate(ids):
                key = (mi, l1)
                if key not in models:
                    models[key] = ppf_amd.Model(w.models[mi], w.d_dists[mi], w.vote_count_threshold,
                                                use_l1_norm=l1, ctx=ctx)
                r = models[key].ppf_lookup(sc)
                assert _digest(r, True) == dig[k]
                r.close()
            sc.close()
    finally:
        srv.close()
        for m in srv_models + list(models.values()):
            m.close()


def test_served_configuration_requests_alone_mix_plain(ctx, batched_workload):
    """Requests of the served configuration at its sizes (scenes of
    5k-100k points with several targets among the 16 resident models of the
    batched configuration; targets kept when n_scene * n_model <= 6e7, the
    bench's bound): each result is the same alone, in the mix and as a plain
    lookup, bit for bit."""
    w = batched_workload
    reqs = synth.served_requests(w, seed=5, n_requests=6,
                                 keep=lambda n_s, mi: n_s * len(w.models[mi][0]) <= 6e7)
    print("requests:", [(len(s[0]), ids) for s, ids in reqs], flush=True)
    models = [ppf_amd.Model(m, d, w.vote_count_threshold, ctx=ctx) for m, d in zip(w.models, w.d_dists)]
    srv = ppf_amd.Server(models, round_ms=20.0, ctx=ctx)
    try:
        alone = []
        for s, ids in reqs:
            t = srv.submit(s, ids, w.rpdf, w.vote_count_threshold)
            poses, st, info, res = srv.wait(t, 300.0, results=True)
            assert info["rc"] == ppf_amd.PPF_OK, info
            alone.append([_digest(r, True) for r in res])
            for r in res:
                r.close()
        tickets = [srv.submit(s, ids, w.rpdf, w.vote_count_threshold) for s, ids in reqs]
        for t, dig_a in zip(tickets, alone):
            poses, st, info, res = srv.wait(t, 300.0, results=True)
            assert info["rc"] == ppf_amd.PPF_OK, info
            assert [_digest(r, True) for r in res] == dig_a
            for r in res:
                r.close()
        for (s, ids), dig_a in zip(reqs, alone):
            sc = ppf_amd.Scene(s, 0.0, w.rpdf, ctx=ctx)
            for k, mi in enumerate(ids):
                r = models[mi].ppf_lookup(sc)
                assert _digest(r, True) == dig_a[k]
                r.close()
            sc.close()
    finally:
        srv.close()
        for m in models:
            m.close()
