"""Pins the oracle before it is trusted: every function is checked against
the reference's frozen golden vectors and against what the mathematics
fixes (contract definitions, brute force, exact rationals), so a dropped
term, a wrong tie order or a transposed operand fails here."""
import json
import math
import pathlib
from fractions import Fraction

import numpy as np
import pytest

from bench import make_ic_cpu_inputs
from oracle import c_oracle, exact, np_arm

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def anchors():
    z = np.load(GOLD / "parity_anchors_v1.npz", allow_pickle=False)
    m = json.loads((GOLD / "parity_anchors_v1.manifest.json").read_text())
    return z, m


def test_anchor_npz_hash(anchors):
    import hashlib
    _, m = anchors
    got = hashlib.sha256((GOLD / "parity_anchors_v1.npz").read_bytes()).hexdigest()
    assert got.lower() == m["hash"]["data_sha256"].lower()


def test_anchor_cases_through_oracle(anchors):
    """Every frozen case of parity_anchors_v1 through the oracle."""
    z, m = anchors
    n_checked = 0
    for c in m["case_schema"]:
        n_checked += 1
        inputs = [z[k] for k in c["inputs"]]
        mask = z[c["masks"]["mask"]] if (c.get("masks") or {}).get("mask") else None
        op, exp = c["operation"], c["expected"]
        if op == "cs_rank":
            x = inputs[0].reshape(1, -1)
            m2 = None if mask is None else mask.reshape(1, -1)
            for got in (np_arm.np_cs_rank(x, m2, c["params"]["descending"]),
                        c_oracle.rank_rows(x, m2, c["params"]["descending"])):
                e = np.array([np.nan if v == "nan" else v for v in exp], dtype=np.float32)
                g = got.reshape(-1)
                assert np.array_equal(np.isnan(g), np.isnan(e)), c["id"]
                assert np.array_equal(g[~np.isnan(e)], e[~np.isnan(e)]), c["id"]
                if "nan_payload" in c:
                    assert np.all(g[np.isnan(g)].view(np.uint32) == 0x7FC00000), c["id"]
        elif op == "rolling_ic":
            f = inputs[0].reshape(1, -1)
            r = inputs[1].reshape(1, -1)
            mv = c["params"]["min_valid"]
            for got in (np_arm.np_rolling_ic(f, r, None, None, mv)[0],
                        c_oracle.rolling_ic(f.astype(np.float64), r.astype(np.float64),
                                            None, None, mv)[0]):
                if exp == "nan":
                    assert np.isnan(got), c["id"]
                else:
                    assert abs(got - float(exp)) <= 1e-12, c["id"]
        elif op == "factor_corr":
            if len(inputs) == 1:
                with pytest.raises(ValueError):
                    np_arm.np_factor_corr(inputs[0], None)
                continue
            F3 = np.stack([inputs[0], inputs[1]], axis=1).reshape(1, -1, 2)
            if exp == "ValueError":
                with pytest.raises(ValueError):
                    np_arm.np_factor_corr(F3, None)
                continue
            got = np_arm.np_factor_corr(F3, mask.reshape(1, -1) if mask is not None else None)[0, 1]
            if exp == "nan":
                assert np.isnan(got), c["id"]
            else:
                assert abs(got - float(exp)) <= 1e-12, c["id"]
        elif op == "parameter_scan":
            groups = np_arm.np_parameter_scan(inputs[0], mask)
            assert len(groups) == 4 and all(g.shape == inputs[0].shape for g in groups)
    assert n_checked == 27


def test_corr_corpus_oracle(anchors):
    """corr_corpus_v1: the per-pair oracle reproduces every frozen expected
    value (they were produced by np.corrcoef); the HG-2 serial-Kahan
    reference matches the arm on the bias cases."""
    z = np.load(GOLD / "corr_corpus_v1.npz", allow_pickle=False)
    m = json.loads((GOLD / "corr_corpus_v1.manifest.json").read_text())
    for c in m["cases"]:
        nm = c["id"]
        a, b = z[nm + "_a"], z[nm + "_b"]
        v = np.isfinite(a) & np.isfinite(b)
        for mk in ("_mask_a", "_mask_b"):
            if nm + mk in z.files:
                v &= z[nm + mk]
        got = exact.pair_corrcoef(a, b, v)
        if c["expected"] is None:
            assert np.isnan(got), nm
        else:
            assert got == pytest.approx(c["expected"], abs=1e-15), nm
        if c["branch"] == "kahan" and v.sum() >= 2:
            X = np.column_stack([a, b]).astype(np.float64)
            arm = np_arm.np_stock_corr(X, np.column_stack([v, v]))[0, 1]
            assert abs(arm - exact.serial_kahan_corr(a[v], b[v])) <= 1e-12, nm


def test_rank_definitions_agree():
    """np arm == C oracle == stable argsort of canonical keys == O(N^2)
    contract definition, with dense ties, +-0, inf and NaN."""
    rng = np.random.default_rng(0)
    X = np.round(rng.standard_normal((6, 97)), 1).astype(np.float32)
    X[0, :5] = [0.0, -0.0, np.inf, -np.inf, np.nan]
    X[1] = 3.0
    M = rng.random(X.shape) < 0.7
    for desc in (False, True):
        for m in (None, M):
            a = np_arm.np_cs_rank(X, m, desc)
            b = c_oracle.rank_rows(X, m, desc)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
            v = np.isfinite(X) & (True if m is None else m)
            for t in range(X.shape[0]):
                br = exact.brute_rank_row(X[t], v[t], desc)
                sr = exact.stable_rank_row(X[t], v[t], desc)
                assert np.array_equal(np.nan_to_num(br, nan=-1), np.nan_to_num(a[t], nan=-1))
                assert np.array_equal(np.nan_to_num(sr, nan=-1), np.nan_to_num(a[t], nan=-1))


def test_planted_mistakes_fail():
    """A tie-order or sign mistake planted in a copy of the oracle is caught
    by the brute-force definition."""
    x = np.array([3.0, 3.0, 1.0], dtype=np.float32)
    v = np.ones(3, bool)
    bad_ties = np.array([3.0, 2.0, 1.0], dtype=np.float32)  # reversed tie order
    assert not np.array_equal(exact.brute_rank_row(x, v), bad_ties)
    bad_desc = exact.brute_rank_row(x, v, False)            # forgot to negate
    assert not np.array_equal(exact.brute_rank_row(x, v, True), bad_desc)


def test_ic_exact_closed_form():
    """IC from the arm equals the exact rational closed form
    3(4S - m(m+1)^2)/(m(m^2-1)) within 1e-12, and the C oracle's integer
    finalize equals the arm bitwise."""
    rng = np.random.default_rng(1)
    F = np.round(rng.standard_normal((25, 300)), 2)
    R = 0.2 * F + rng.standard_normal((25, 300))
    F[rng.random(F.shape) < 0.03] = np.nan
    M = rng.random(F.shape) < 0.9
    arm = np_arm.np_rolling_ic(F, R, M, None, 30)
    c = c_oracle.rolling_ic(F, R, M, None, 30)
    assert np.array_equal(arm, c, equal_nan=True)
    ok = np.isfinite(F) & np.isfinite(R) & M
    for t in range(F.shape[0]):
        e = exact.spearman_exact(F[t], R[t], ok[t], 30)
        cc = exact.spearman_corrcoef(F[t], R[t], ok[t], 30)
        assert (np.isnan(e) and np.isnan(arm[t])) or abs(e - arm[t]) <= 1e-12
        assert (np.isnan(cc) and np.isnan(arm[t])) or abs(cc - arm[t]) <= 1e-12


def test_ic_cpu_config_exact():
    """On the rolling_ic CPU-arm workload (bench.py make_ic_cpu_inputs:
    T=250, N=500, f32, min_valid=30) the arm and the C oracle agree bitwise and
    both sit within 1e-12 of the exact rational IC and of np.corrcoef of the
    ordinal ranks, date by date."""
    X, Y, M = make_ic_cpu_inputs()
    arm = np_arm.np_rolling_ic(X, Y, M, None, 30)
    c = c_oracle.rolling_ic(X, Y, M, None, 30)
    assert np.array_equal(arm, c, equal_nan=True)
    assert not np.isnan(arm).any()
    ok = np.isfinite(X) & np.isfinite(Y) & M
    for t in range(X.shape[0]):
        assert abs(exact.spearman_exact(X[t], Y[t], ok[t], 30) - arm[t]) <= 1e-12
        assert abs(exact.spearman_corrcoef(X[t], Y[t], ok[t], 30) - arm[t]) <= 1e-12


def test_ic_anchor_semantics():
    # constant section: ranks exist but IC is NaN (contract)
    f = np.full((1, 40), 0.5)
    r = np.linspace(0, 1, 40)[None, :]
    assert np.isnan(np_arm.np_rolling_ic(f, r, None, None, 2)[0])
    # ties: f = [1,1,2], r = [2,1,3] -> ordinal ranks give 0.5 (average would give 0.866)
    f = np.array([[1.0, 1.0, 2.0]])
    r = np.array([[2.0, 1.0, 3.0]])
    assert np_arm.np_rolling_ic(f, r, None, None, 2)[0] == pytest.approx(0.5, abs=1e-15)


def test_label_fixture_ic():
    """rolling_ic on the reference's label fixture agrees across the arm, the
    C oracle and the exact form."""
    z = np.load(GOLD / "rolling_ic_labels_v1.npz", allow_pickle=False)
    price, fwd = z["price"], z["forward_returns"]
    factor = -np.log(price)
    arm = np_arm.np_rolling_ic(factor, fwd, None, None, 3)
    c = c_oracle.rolling_ic(factor, fwd, None, None, 3)
    assert np.array_equal(arm, c, equal_nan=True)
    assert np.isfinite(arm).sum() > 0


def test_corr_arm_vs_pairwise_oracle():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((80, 25)) + rng.uniform(-1, 1, 25)
    X[rng.random(X.shape) < 0.02] = np.nan
    M = rng.random(X.shape) < 0.85
    X[:, 4] = 1.5
    a = np_arm.np_stock_corr(X, M)
    b = exact.corr_matrix_pairwise(X, M)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(b)
    assert np.max(np.abs(a[ok] - b[ok])) <= 1e-12
    # one value against the Fraction-exact Pearson
    o = np.isfinite(X[:, 0]) & np.isfinite(X[:, 1]) & M[:, 0] & M[:, 1]
    assert abs(a[0, 1] - exact.pearson_fraction(X[o, 0], X[o, 1])) <= 1e-14


def test_dd_mean_equals_kahan_on_hg2_fixtures():
    """The exact pass of factor_corr takes each mean as the correctly rounded
    sum / n; on the HG-2 bias fixtures that equals the serial Kahan mean the
    HG-2 clause is defined on."""
    z = np.load(GOLD / "corr_corpus_v1.npz", allow_pickle=False)
    for nm in ("bias_1e12", "bias_1e15", "f64_ulp_bias"):
        for side in ("_a", "_b"):
            x = z[nm + side].astype(np.float64)
            assert math.fsum(x) / x.size == np_arm.kahan_mean(x), nm + side
            assert float(sum(Fraction(v) for v in x.tolist())) == math.fsum(x)
    # near_zero_var: serial Kahan's sum is 1 ulp off the correctly rounded one,
    # and the correlation from either mean still agrees within 1e-12
    a, b = z["near_zero_var_a"], z["near_zero_var_b"]
    def corr_with(ma, mb):
        ac, bc = a - ma, b - mb
        return (np.sum(ac * bc) / np.sqrt(np.sum(ac * ac))) / np.sqrt(np.sum(bc * bc))
    r_k = corr_with(np_arm.kahan_mean(a), np_arm.kahan_mean(b))
    r_f = corr_with(math.fsum(a) / a.size, math.fsum(b) / b.size)
    assert abs(r_k - r_f) <= 1e-12
