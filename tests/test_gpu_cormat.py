"""GPU parity: the dense correlation matrix, subject averaging (plain and
Fisher) and thresholds on the averaged matrix, against the FP64 oracle.

Bit-exact: the averaged matrix without Fisher (every pair's FP32 r is the
oracle's), and every CSR (edges and weights) given the matrix.  Fisher mode:
within 1 FP32 ulp per entry (the device and host FP64 log/exp may differ in the
last bit, which moves the FP32 rounding in rare cases)."""
import numpy as np
import pytest

from conftest import say
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def subjects():
    from pagani_amd import synth
    n = 700
    subs = [synth.bold(300 + s, n, t, 9) for s, t in enumerate((120, 95, 150))]
    # exact duplicate voxels (r = 1, clamped) and a constant voxel (r = 0)
    for x in subs:
        x[:, 11] = x[:, 3]
        x[:, 500] = 2.5
    return n, subs


def _gpu_average(ctx, n, subs, fisher):
    ctx.cormat_begin(n, fisher)
    for x in subs:
        ctx.cormat_add(np.ascontiguousarray(x), x.shape[0])
    ctx.cormat_finish()
    return ctx.cormat_read(n)


def _ulp_diff(a, b):
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7fffffff), ia)
    ib = np.where(ib < 0, -(ib & 0x7fffffff), ib)
    return np.abs(ia - ib)


def test_single_subject_matrix_bit_exact(ctx, subjects):
    n, subs = subjects
    say("cormat one subject")
    got = _gpu_average(ctx, n, subs[:1], False)
    np.testing.assert_array_equal(got.view(np.uint32), O.cormat_average(subs[:1]).view(np.uint32))


def test_average_plain_bit_exact(ctx, subjects):
    n, subs = subjects
    say("cormat average plain")
    got = _gpu_average(ctx, n, subs, False)
    np.testing.assert_array_equal(got.view(np.uint32), O.cormat_average(subs).view(np.uint32))


def test_average_inline_recheck_path(ctx, subjects, monkeypatch):
    # no room in the ambiguous-pair list: every such pair is recomputed inline
    n, subs = subjects
    monkeypatch.setenv("PAGANI_CM_AMBCAP", "0")
    say("cormat average, inline rechecks")
    got = _gpu_average(ctx, n, subs[:2], False)
    np.testing.assert_array_equal(got.view(np.uint32), O.cormat_average(subs[:2]).view(np.uint32))


def test_average_fisher_within_one_ulp(ctx, subjects):
    n, subs = subjects
    say("cormat average fisher")
    got = _gpu_average(ctx, n, subs, True)
    want = O.cormat_average(subs, True)
    d = _ulp_diff(got, want)
    assert d.max() <= 1, int(d.max())
    assert (d > 0).mean() < 1e-3


@pytest.mark.parametrize("fisher", [False, True])
@pytest.mark.parametrize("mode,thr", [("r", 0.3), ("r", 0.55), ("s", 1.0), ("s", 12.5)])
@pytest.mark.parametrize("weighted", [False, True])
def test_thresholds_on_average(ctx, subjects, fisher, mode, thr, weighted):
    n, subs = subjects
    say(f"cormat threshold fisher={fisher} {mode} {thr} weighted={weighted}")
    a = _gpu_average(ctx, n, subs, fisher)
    info = ctx.cormat_threshold(thr, sparsity=mode == "s", weighted=weighted)
    row, col, val = ctx.corr_csr(n, info.nnz, weighted)
    if mode == "r":
        orow, ocol, oval = O.dense_csr_r(a, n, thr, weighted)
        m1 = n * (n - 1) // 2
        assert info.s_thresh == np.float32(100.0 * orow[-1] / m1 / 2)
    else:
        orow, ocol, oval, t = O.dense_csr_s(a, n, thr, weighted)
        assert info.r_thresh == t
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    if weighted:
        np.testing.assert_array_equal(val.view(np.uint32), oval.view(np.uint32))


def test_state_errors(ctx):
    from pagani_amd.lib import PgError
    ctx.cormat_begin(50, False)
    with pytest.raises(PgError, match="finish"):
        ctx.cormat_threshold(0.5)
    with pytest.raises(PgError, match="no subject"):
        ctx.cormat_finish()


def test_smallest_matrix(ctx):
    # n = 2, t = 2: one pair, r = +-1 exactly; 100% sparsity keeps it
    x = np.array([[0.0, 1.0], [2.0, 3.0]], np.float32)
    say("cormat n=2 t=2")
    got = _gpu_average(ctx, 2, [x], False)
    np.testing.assert_array_equal(got.view(np.uint32), O.cormat_average([x]).view(np.uint32))
    info = ctx.cormat_threshold(100.0, sparsity=True)
    row, col, _ = ctx.corr_csr(2, info.nnz)
    assert info.nnz == 2 and list(row) == [0, 1, 2] and list(col) == [1, 0]
