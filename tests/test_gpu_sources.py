"""tgtp_score_sources (source subsets) on the GPU, through the C ABI.

FAITHFUL rows equal the oracle's rows of the same sources bit for bit;
scores of edges whose both endpoints are listed equal the full run's in
every mode (per-source work is independent: aesc.h:456-527); unlisted
entries are left untouched; invalid ranges are refused.
"""
import numpy as np
import pytest

import oracle
from tgtp_amd import MODE_FAITHFUL, MODE_FAST_BF16, MODE_FAST_FP32, TgtpError
from tgtp_amd import eigenprep as E
from tgtp_amd import graph as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cl2000():
    g = G.chung_lu(2000, 2.4, 14.0, seed=11)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 48))
    return g, lam, X


RANGES = [(0, 5), (64, 200), (700, 701), (1500, 2000)]


def _listed(g, ranges):
    mask = np.zeros(g.n, dtype=bool)
    for b, e in ranges:
        mask[b:e] = True
    rows_of = np.repeat(np.arange(g.n), np.diff(g.idx.astype(np.int64)))
    return mask, mask[rows_of], mask[rows_of] & mask[g.adj]


@pytest.mark.parametrize("eps", [0.05, 0.01])
def test_faithful_rows_equal_oracle(gpu_ctx, cl2000, eps):
    g, lam, X = cl2000
    dg = gpu_ctx.upload(g)
    rows, scores, st = gpu_ctx.score_sources(dg, lam, X, eps, RANGES, mode=MODE_FAITHFUL)
    full, _ = gpu_ctx.score(dg, lam, X, eps, 10, mode=MODE_FAITHFUL)
    dg.free()
    mask, row_pos, both = _listed(g, RANGES)
    src = np.nonzero(mask)[0].astype(np.uint32)
    taus = oracle.taus(g.idx, g.adj, lam, X, np.float32(eps / 2))
    ref = oracle.tgtp_sources(g.idx, g.adj, taus, eps, 10, src)
    assert st["work_items"] == src.shape[0]
    assert np.array_equal(rows[row_pos].view(np.uint32), ref[row_pos].view(np.uint32))
    assert np.isnan(rows[~row_pos]).all()
    assert np.array_equal(scores[both].view(np.uint32), full[both].view(np.uint32))
    assert np.isnan(scores[~both]).all()


@pytest.mark.parametrize("mode", [MODE_FAST_FP32, MODE_FAST_BF16])
def test_fast_subset_scores_equal_full_run(gpu_ctx, cl2000, mode):
    g, lam, X = cl2000
    dg = gpu_ctx.upload(g)
    for eps in (0.05, 0.01):
        rows, scores, _ = gpu_ctx.score_sources(dg, lam, X, eps, RANGES, mode=mode)
        full, tr, _ = gpu_ctx.score_trace(dg, lam, X, eps, 10, mode=mode)
        _, row_pos, both = _listed(g, RANGES)
        assert np.array_equal(scores[both].view(np.uint32), full[both].view(np.uint32))
        assert np.array_equal(rows[row_pos].view(np.uint32),
                              tr["preds_row"][row_pos].view(np.uint32))
    dg.free()


def test_invalid_ranges_refused(gpu_ctx, cl2000):
    g, lam, X = cl2000
    dg = gpu_ctx.upload(g)
    for bad in ([(5, 5)], [(10, 20), (15, 30)], [(20, 30), (0, 10)], [(0, g.n + 1)]):
        with pytest.raises(TgtpError):
            gpu_ctx.score_sources(dg, lam, X, 0.05, bad, mode=MODE_FAST_FP32)
    dg.free()
