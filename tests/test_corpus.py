"""Benchmark corpus (factor-cuda_amd/fc_corpus.py): the loader reads the
reference's own corpus files (tests/golden/corpus_synth_smoke_v1.*, copied
as data from the reference's benchmark_corpus/), checking the file and
per-array SHA-256 the way its loader does (corpus_loader_v1.py:36-79);
tampering is refused; generation is byte-for-byte deterministic; prefix
subsets follow corpus_loader_v1.py:82-96; the forward-return label rule is
checked against a direct loop."""
import pathlib
import shutil

import numpy as np
import pytest

import fc_corpus as fcc

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
REF_ID = "corpus_synth_smoke_v1"


def test_reference_corpus_loads_with_hashes_checked():
    d, m = fcc.load(REF_ID, GOLD)
    assert d["factors"].shape == (40, 100, 4) and d["factors"].dtype == np.float32
    assert m["shapes"] == {"T": 40, "N": 100, "F": 4}
    assert np.array_equal(d["factor_a"], d["factors"][..., 0])
    assert d["mask"].dtype == np.bool_ and not d["factors"].flags.writeable
    assert int(d["h"][0]) == 5 and int(d["lag"][0]) == 1
    # the reference's label rule holds on its own data where both ends are finite
    fwd = fcc.forward_returns(d["price"], np.isfinite(d["price"]))
    both = np.isfinite(fwd) & np.isfinite(d["forward_returns"])
    assert both.sum() > 1000
    np.testing.assert_allclose(fwd[both], d["forward_returns"][both], rtol=0, atol=1e-15)


def test_tampered_or_unhashed_corpus_is_refused(tmp_path):
    for ext in (".npz", ".manifest.json"):
        shutil.copy(GOLD / (REF_ID + ext), tmp_path / (REF_ID + ext))
    p = tmp_path / (REF_ID + ".npz")
    b = bytearray(p.read_bytes())
    b[-200] ^= 1
    p.write_bytes(bytes(b))
    with pytest.raises(fcc.CorpusError):
        fcc.load(REF_ID, tmp_path)
    with pytest.raises(fcc.CorpusError):
        fcc.load("absent_v1", tmp_path)


def test_generation_is_deterministic_and_round_trips(tmp_path):
    a = fcc.generate(30, 70, 5)
    m1 = fcc.write("g_v1", a, tmp_path / "one")
    m2 = fcc.write("g_v1", fcc.generate(30, 70, 5), tmp_path / "two")
    assert m1["hash"]["data_sha256"] == m2["hash"]["data_sha256"]
    assert (tmp_path / "one" / "g_v1.npz").read_bytes() == (tmp_path / "two" / "g_v1.npz").read_bytes()
    d, m = fcc.load("g_v1", tmp_path / "one")
    for k, v in a.items():
        assert np.array_equal(d[k], v, equal_nan=v.dtype.kind == "f"), k
    other = fcc.generate(30, 70, 5, master=fcc.MASTER_SEED + 1)
    assert not np.array_equal(other["factors"], a["factors"])


def test_generated_panel_properties():
    d = fcc.generate(200, 400, 6)
    T, N, F = d["factors"].shape
    assert np.array_equal(d["factor_a"], d["factors"][..., 0])
    # values on the 1e-3 grid; a few percent of exact ties per date
    x = d["factors"].astype(np.float64)
    assert np.allclose(x * 1000, np.round(x * 1000), atol=1e-3)
    ties = np.mean([1 - np.unique(r).size / r.size for r in d["factors"][..., 0]])
    assert 0.005 < ties < 0.1
    frac = d["mask"].mean(axis=1)
    assert frac.min() >= 0.0 and frac.max() <= 1.0 and 0.3 < frac.mean() < 0.95
    # the masked cells keep finite prices: the mask, not the value, marks them
    assert np.isfinite(d["price"]).all()
    assert np.isnan(d["forward_returns"][-(fcc.H + fcc.LAG):]).all()


def test_forward_returns_against_a_loop():
    rng = np.random.default_rng(3)
    price = 100 * np.cumprod(1 + 0.01 * rng.standard_normal((15, 9)), axis=0)
    mask = rng.random((15, 9)) < 0.8
    got = fcc.forward_returns(price, mask)
    for t in range(15):
        for i in range(9):
            if t + 6 < 15 and mask[t + 1, i] and mask[t + 6, i]:
                assert got[t, i] == price[t + 6, i] / price[t + 1, i] - 1.0
            else:
                assert np.isnan(got[t, i])


def test_subset_prefix():
    d, _ = fcc.load(REF_ID, GOLD)
    s = fcc.subset_prefix(d, t=10, n=20, f=2)
    assert s["factors"].shape == (10, 20, 2) and s["mask"].shape == (10, 20)
    assert np.array_equal(s["factors"], d["factors"][:10, :20, :2])
    assert not s["factors"].flags.writeable
    with pytest.raises(ValueError):
        fcc.subset_prefix(d, t=41)
    with pytest.raises(ValueError):
        fcc.subset_prefix(d, f=0)
