"""The counts of sqmi_align_counts are the exact (sub)gradient of the optimal
score with respect to the scoring parameters.

AlignScore wraps the library as a torch.autograd.Function (test harness
only, CPU tensors: the library drives the GPU itself).  The score of a pair
is max over alignments of counts(alignment) . params, so wherever the +1 and
-1 finite differences of a parameter agree the score is linear there and the
difference must equal that parameter's count on the reported path."""
import numpy as np
import pytest

import seqan_amd as sa

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


class AlignScore(torch.autograd.Function):
    """scores = align(H, V; table, gap_open, gap_extend), global alignment."""

    @staticmethod
    def forward(fctx, table, gap_open, gap_extend, H, V, lib_ctx):
        A = table.shape[0]
        score = sa.Score(table=table.detach().round().to(torch.int32).numpy(),
                         gap_extend=int(round(gap_extend.item())), gap_open=int(round(gap_open.item())))
        s, c = lib_ctx.align_counts(H, V, score, 0)
        fctx.save_for_backward(torch.from_numpy(c.astype(np.float64)))
        fctx.A = A
        return torch.from_numpy(s.astype(np.float64))

    @staticmethod
    def backward(fctx, grad):
        (c,) = fctx.saved_tensors
        A = fctx.A
        g = grad.unsqueeze(1) * c
        return g[:, :A * A].sum(0).reshape(A, A), g[:, A * A].sum(), g[:, A * A + 1].sum(), None, None, None


def _pairs(seed, count, lo, hi):
    rng = np.random.default_rng(seed)
    hs, vs = [], []
    for _ in range(count):
        h = rng.integers(0, 4, int(rng.integers(lo, hi + 1))).astype(np.uint8)
        v = h.copy()
        mut = rng.random(len(v)) < 0.1
        v[mut] = rng.integers(0, 4, int(mut.sum()))
        cut = int(rng.integers(0, len(v)))
        v = np.concatenate([v[:cut], v[cut + int(rng.integers(0, 4)):]])
        if len(v) == 0:
            v = h[:1].copy()
        hs.append(h)
        vs.append(v)
    return sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)


def test_autograd_gradient_is_summed_counts(ctx):
    H, V = _pairs(1, 128, 50, 400)
    table = torch.tensor(sa.simple_score(2, -3, -1, -5).table, dtype=torch.float64, requires_grad=True)
    o = torch.tensor(-5.0, dtype=torch.float64, requires_grad=True)
    e = torch.tensor(-1.0, dtype=torch.float64, requires_grad=True)
    w = torch.arange(1, 129, dtype=torch.float64)
    scores = AlignScore.apply(table, o, e, H, V, ctx)
    (w * scores).sum().backward()
    s, c = ctx.align_counts(H, V, sa.simple_score(2, -3, -1, -5))
    assert np.array_equal(scores.detach().numpy(), s)
    wc = (w.numpy()[:, None] * c).sum(0)
    assert np.allclose(table.grad.numpy().ravel(), wc[:16])
    assert o.grad.item() == wc[16] and e.grad.item() == wc[17]


@pytest.mark.parametrize("base", [(2, -3, -1, -5), (16, -24, -8, -40), (5, -4, -1, -11)])
def test_counts_equal_finite_differences(ctx, base):
    H, V = _pairs(sum(abs(x) for x in base), 256, 30, 300)
    sc = sa.simple_score(*base)
    s0, c = ctx.align_counts(H, V, sc)
    params = [("t", k) for k in range(16)] + [("o", None), ("e", None)]
    agree = np.ones(len(s0), bool)
    checked = 0
    for idx, (kind, k) in enumerate(params):
        diffs = []
        for d in (+1, -1):
            t = sc.table.copy()
            o, e = sc.gap_open, sc.gap_extend
            if kind == "t":
                t.ravel()[k] += d
            elif kind == "o":
                o += d
            else:
                e += d
            if o > 0 or e > 0:
                diffs.append(None)
                continue
            diffs.append(ctx.align_score(H, V, sa.Score(table=t, gap_extend=e, gap_open=o), 0).astype(np.int64) - s0)
        if diffs[0] is None or diffs[1] is None:
            continue
        lin = diffs[0] == -diffs[1]
        # Where the score is linear in this parameter, the count is its slope.
        assert np.array_equal(c[lin, idx], diffs[0][lin]), (kind, k)
        agree &= lin
        checked += int(lin.sum())
    print("tie-free pairs", int(agree.sum()), "of", len(s0), "checked slopes", checked, flush=True)
    assert agree.sum() >= len(s0) // 4
