"""The fc operators on corpus data, against the oracle: the reference's own
smoke corpus (tests/golden/corpus_synth_smoke_v1.*, with its halted cells
masked but finite, as its corpus design has them) and a generated corpus,
through the host path (numpy inputs) and the device-resident path (GPU
tensors). Ranks bitwise; IC and contract-tier correlations within 1e-12."""
import pathlib

import numpy as np
import pytest

import fc
import fc_corpus
from fc import _util as u
from oracle import c_oracle, exact

pytestmark = pytest.mark.gpu
GOLD = pathlib.Path(__file__).resolve().parent / "golden"
TOL = 1e-12


def _corpora():
    d, _ = fc_corpus.load("corpus_synth_smoke_v1", GOLD)
    yield "reference_smoke", d
    yield "generated", fc_corpus.generate(120, 300, 6)


def _close(got, exp, tol=TOL):
    got, exp = np.asarray(got, np.float64), np.asarray(exp, np.float64)
    assert np.array_equal(np.isnan(got), np.isnan(exp))
    ok = ~np.isnan(exp)
    if ok.any():
        assert np.abs(got[ok] - exp[ok]).max() <= tol


@pytest.mark.parametrize("device", [False, True])
def test_corpus_ops_match_oracle(device):
    t = u._torch
    for name, d in _corpora():
        fa, m = d["factor_a"], d["mask"]
        fwd = d["forward_returns"]
        conv = (lambda a: t.from_numpy(np.array(a)).to("cuda:0")) if device else \
            (lambda a: a)
        back = (lambda a: a.cpu().numpy()) if device else np.asarray
        # cross_sectional_rank (masked ascending): bitwise
        r = back(fc.cross_sectional_rank(conv(fa), conv(m)))
        ref = c_oracle.rank_rows(fa, m)
        assert np.array_equal(r.view(np.uint32), ref.view(np.uint32)), name
        # rolling_ic of factor_a (f32) vs the f64 labels: both sides upcast to f64
        ic = back(fc.rolling_ic(conv(fa), conv(fwd), conv(m), conv(m), min_valid=30,
                                device="cuda"))
        _close(ic, c_oracle.rolling_ic(fa.astype(np.float64), fwd, m, m, 30))
        # factor_corr over the (T,N,F) panel, pooled over (t,i)
        F3 = d["factors"]
        c = back(fc.factor_corr(conv(F3), conv(m), backend="cuda"))
        T, N, F = F3.shape
        _close(c, exact.corr_matrix_pairwise(F3.reshape(T * N, F), m.reshape(T * N, 1)))
        # stock_corr of the returns over t
        s = back(fc.stock_corr(conv(d["returns"]), conv(m), backend="cuda"))
        _close(s, exact.corr_matrix_pairwise(d["returns"], m))
