"""Model files (ppf_model_save / ppf_model_load): a loaded model gives the
same hash array and the same lookup results, bit for bit, as the model it
was saved from; damaged or foreign files are rejected with a status."""
import numpy as np
import pytest

import ppf_amd
import synth

pytestmark = pytest.mark.gpu


def test_save_load_round_trip(ctx, tmp_path):
    w = synth.plain_small(seed=17, n_model=700, n_clutter=900)
    m = ppf_amd.Model(w.models[0], w.d_dists[0], w.vote_count_threshold, ctx=ctx)
    f = str(tmp_path / "m.ppfm")
    m.save(f)
    m2 = ppf_amd.Model.load(f, w.vote_count_threshold, ctx=ctx)
    assert m2.info()["n_unique_keys"] == m.info()["n_unique_keys"]
    for a, b in zip(m.hash_array(), m2.hash_array()):
        assert np.array_equal(a, b)
    s = ppf_amd.Scene(w.scenes[0], 0.0, w.rpdf, ctx=ctx)
    r1, r2 = m.ppf_lookup(s), m2.ppf_lookup(s)
    c1, n1 = r1.votes()
    c2, n2 = r2.votes()
    assert np.array_equal(c1, c2) and np.array_equal(n1, n2)
    assert r1.pose().tobytes() == r2.pose().tobytes()
    for o in (r1, r2, s, m, m2):
        o.close()


def test_damaged_files_rejected(ctx, tmp_path):
    w = synth.plain_small(seed=18, n_model=300, n_clutter=100)
    m = ppf_amd.Model(w.models[0], w.d_dists[0], 0.3, ctx=ctx)
    f = tmp_path / "m.ppfm"
    m.save(str(f))
    data = bytearray(f.read_bytes())
    flipped = bytearray(data)
    flipped[200] ^= 0x40
    (tmp_path / "flip.ppfm").write_bytes(bytes(flipped))
    # the low mantissa bit of d_dist (header offset 24): a consistent header
    # that only the checksum over the header catches
    hdr = bytearray(data)
    hdr[24] ^= 0x01
    (tmp_path / "hdr.ppfm").write_bytes(bytes(hdr))
    (tmp_path / "short.ppfm").write_bytes(bytes(data[:len(data) // 2]))
    (tmp_path / "magic.ppfm").write_bytes(b"NOTMODEL" + bytes(data[8:]))
    for name, msg in (("flip", "checksum"), ("hdr", "checksum"), ("short", "ends early"),
                      ("magic", "not a model")):
        with pytest.raises(ppf_amd.PPFError) as e:
            ppf_amd.Model.load(str(tmp_path / f"{name}.ppfm"), ctx=ctx)
        assert e.value.code == ppf_amd.PPF_ERR_ARG and msg in str(e.value)
    m.close()
