"""GPU parity: voxel series -> CSR (libpagani) against the FP64 oracle.

Edge sets, row offsets and (weighted) values must be bit-identical to the
oracle for both the FP32 and the BF16x3 Gram.
"""
import numpy as np
import pytest

from conftest import hip_d2h, say
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PRECS = [0, 1]  # FP32, BF16X3


def _data(n, t, k, seed):
    from pagani_amd import synth
    return synth.bold(seed, n, t, k)


def _build(ctx, x, thr, sparsity=False, weighted=False, prec=1, layout=0):
    t, n = x.shape
    if layout == 1:
        x = np.ascontiguousarray(x.T)
    info = ctx.corr_build(x, n, t, thr, sparsity=sparsity, weighted=weighted, prec=prec, layout=layout)
    row, col, val = ctx.corr_csr(n, info.nnz, weighted)
    return info, row, col, val


def _assert_same(row, col, val, orow, ocol, oval):
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    if oval is not None:
        np.testing.assert_array_equal(val.view(np.uint32), oval.view(np.uint32))


@pytest.mark.parametrize("t,layout", [(173, 0), (173, 1), (420, 0), (420, 1)])
def test_zscore_matches_oracle_bitwise(ctx, t, layout):
    # T <= 360 takes the LDS-staged z-score kernel, longer series the two-pass one
    x = _data(300, t, 7, 11)
    say(f"zscore n=300 t={t} layout={layout}")
    info, *_ = _build(ctx, x, 0.5, layout=layout)
    z = np.frombuffer(hip_d2h(ctx.corr_z64_view(), 300 * t * 8), np.float64).reshape(300, t)
    z64, _ = O.zscore(x)
    np.testing.assert_array_equal(z.view(np.uint64), z64.view(np.uint64))


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("thr", [0.4, 0.15])
def test_rmode_bit_exact(ctx, prec, weighted, thr):
    x = _data(700, 200, 10, 3)
    z64, _ = O.zscore(x)
    say(f"rmode prec={prec} weighted={weighted} thr={thr}")
    info, row, col, val = _build(ctx, x, thr, weighted=weighted, prec=prec)
    orow, ocol, oval = O.corr_csr_r(z64, thr, weighted)
    _assert_same(row, col, val, orow, ocol, oval)
    assert info.nnz == orow[-1]
    m1 = 700 * 699 // 2
    assert info.s_thresh == np.float32(100.0 * info.nnz / m1 / 2)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("s", [1.0, 7.5])
def test_smode_bit_exact(ctx, prec, weighted, s):
    x = _data(640, 150, 8, 5)
    z64, _ = O.zscore(x)
    say(f"smode prec={prec} weighted={weighted} s={s}")
    info, row, col, val = _build(ctx, x, s, sparsity=True, weighted=weighted, prec=prec)
    orow, ocol, oval, tstar = O.corr_csr_s(z64, s, weighted)
    _assert_same(row, col, val, orow, ocol, oval)
    assert info.t_star == tstar
    assert info.nnz == 2 * O.sparsity_k(640, s)


@pytest.mark.parametrize("t", [30, 40, 48])
def test_bf16x3_chunk_counts(ctx, t):
    # T = 30 fills the two 16-deep chunks the ring starts with (no copy is
    # issued inside a tile), T = 40 and 48 three (one copy in flight across
    # the tile boundary): r-mode and sparsity against the oracle
    x = _data(900, t, 9, 17)
    z64, _ = O.zscore(x)
    say(f"chunks t={t}")
    info, row, col, val = _build(ctx, x, 0.35, prec=1)
    orow, ocol, oval = O.corr_csr_r(z64, 0.35, False)
    _assert_same(row, col, val, orow, ocol, None)
    info, row, col, _ = _build(ctx, x, 2.0, sparsity=True, prec=1)
    orow, ocol, _, tstar = O.corr_csr_s(z64, 2.0)
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    assert info.t_star == tstar


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("s", [1.0, 3.0])
def test_smode_single_pass_large(ctx, prec, s):
    """N large enough for the sampled single-pass sparsity path."""
    x = _data(6000, 100, 12, 13)
    z64, _ = O.zscore(x)
    say(f"smode large prec={prec} s={s}")
    info, row, col, _ = _build(ctx, x, s, sparsity=True, prec=prec)
    orow, ocol, _, tstar = O.corr_csr_s(z64, s)
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    assert info.t_star == tstar
    assert info.gram_passes == 1  # no fallback to the two-pass path


@pytest.mark.parametrize("prec", PRECS)
def test_smode_ties_row_major(ctx, prec):
    # duplicated voxels give bit-identical r against every other voxel: the
    # k-th largest value is tied many times and the earliest ties must be kept
    x = _data(400, 120, 6, 9)
    for dst, src in [(50, 10), (51, 10), (300, 10), (201, 77), (399, 77)]:
        x[:, dst] = x[:, src]
    z64, _ = O.zscore(x)
    n = 400
    s = _sparsity_splitting_a_tie(z64, n)
    orow, ocol, _, tstar = O.corr_csr_s(z64, s)
    say(f"ties prec={prec} s={s}")
    info, row, col, _ = _build(ctx, x, s, sparsity=True, prec=prec)
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    assert info.t_star == tstar


@pytest.mark.parametrize("prec", PRECS)
def test_smode_tie_group_larger_than_a_bin(ctx, prec):
    # two series, 200 copies each: 39,800 pairs at r = 1 and 40,000 tied at
    # r(A, B); 75% sparsity cuts the tied group, whose members all share one
    # selection bin (more than the selection gathers into LDS)
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal((2, 120)).astype(np.float32)
    x = np.empty((120, 400), np.float32)
    x[:, :200] = a[:, None]
    x[:, 200:] = b[:, None]
    z64, _ = O.zscore(x)
    orow, ocol, _, tstar = O.corr_csr_s(z64, 75.0)
    say(f"tie group prec={prec} nnz={orow[-1]}")
    info, row, col, _ = _build(ctx, x, 75.0, sparsity=True, prec=prec)
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    assert info.t_star == tstar


def test_smode_dense_cluster_overflows_block_staging(ctx):
    # 0.5% sparsity expects about 1% of the pairs kept, so the band-free
    # emit stages hits in its small per-wave area (255 records); a 300-voxel
    # cluster at r ~ 0.9 fills whole 128 x 64 blocks with hits, which then
    # take the direct-write path.  The edge set must not change.
    rng = np.random.default_rng(11)
    x = rng.standard_normal((120, 3000)).astype(np.float32)
    x[:, :300] += 3.0 * rng.standard_normal((120, 1)).astype(np.float32)
    z64, _ = O.zscore(x)
    orow, ocol, _, tstar = O.corr_csr_s(z64, 0.5)
    say(f"dense cluster nnz={orow[-1]}")
    info, row, col, _ = _build(ctx, x, 0.5, sparsity=True, prec=1)
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    assert info.t_star == tstar


def _sparsity_splitting_a_tie(z64, n):
    """A sparsity percent whose k-th largest r is tied and the cut splits the tie."""
    iu = np.triu_indices(n, 1)
    r = np.clip((z64 @ z64.T)[iu], -1, 1)  # ordering only; exact r via the oracle below
    order = np.argsort(-r, kind="stable")
    exact = np.array([O.r_pair(z64, int(iu[0][q]), int(iu[1][q])) for q in order[: r.size // 4]])
    for s in np.linspace(0.3, 24.0, 800, dtype=np.float32):
        k = O.sparsity_k(n, float(s))
        if k < 2 or k > exact.size:
            continue
        t = exact[k - 1]
        g, e = int((exact > t).sum()), int((exact == t).sum())
        if e > 1 and g < k < g + e:
            return float(s)
    raise AssertionError("no sparsity splits a tie group")


@pytest.mark.parametrize("prec", PRECS)
def test_voxel_major_layout_and_zero_variance(ctx, prec):
    x = _data(333, 97, 5, 21)
    x[:, 17] = 3.25  # constant series: r = 0 with everything
    z64, _ = O.zscore(x)
    say(f"layout1 prec={prec}")
    info, row, col, _ = _build(ctx, x, 0.3, prec=prec, layout=1)
    orow, ocol, _ = O.corr_csr_r(z64, 0.3)
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    assert row[18] == row[17]
    info, row, col, _ = _build(ctx, x, -0.2, prec=prec)  # negative threshold: r=0 > -0.2
    orow, ocol, _ = O.corr_csr_r(z64, -0.2)
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)


def test_timing_levels_change_no_result(ctx):
    # level 1 keeps only the emitting Gram kernel's events, level 0 none; the
    # edge set is the same at every level and untimed stages report 0
    x = _data(3000, 120, 8, 31)
    base, row0, col0, _ = _build(ctx, x, 1.0, sparsity=True)
    assert base.ms_zscore > 0 and base.ms_csr > 0 and base.ms_gram_emit > 0
    try:
        ctx.set_timing(1)
        i1, row1, col1, _ = _build(ctx, x, 1.0, sparsity=True)
        assert i1.ms_zscore == 0 and i1.ms_select == 0 and i1.ms_csr == 0 and i1.ms_gram == 0
        assert i1.ms_gram_hist == 0
        assert i1.ms_gram_emit > 0
        ctx.set_timing(0)
        i0, row2, col2, _ = _build(ctx, x, 1.0, sparsity=True)
        assert i0.ms_gram_emit == 0 and i0.ms_gram_hist == 0 and i0.ms_zscore == 0
        for r, c in ((row1, col1), (row2, col2)):
            np.testing.assert_array_equal(r, row0)
            np.testing.assert_array_equal(c, col0)
        with pytest.raises(Exception):
            ctx.set_timing(3)
    finally:
        ctx.set_timing(2)


def test_tiny_inputs(ctx):
    for n in (1, 2, 3, 129):
        x = _data(n, 40, 2, 4)
        z64, _ = O.zscore(x)
        say(f"tiny n={n}")
        info, row, col, _ = _build(ctx, x, 0.0)
        orow, ocol, _ = O.corr_csr_r(z64, 0.0)
        np.testing.assert_array_equal(row, orow)
        np.testing.assert_array_equal(col, ocol)


def test_bad_arguments(ctx):
    import pagani_amd
    x = _data(10, 20, 2, 1)
    with pytest.raises(pagani_amd.PgError):
        ctx.corr_build(x, 10, 20, 1.5)  # |r threshold| > 1
    with pytest.raises(pagani_amd.PgError):
        ctx.corr_build(x, 10, 20, 0.0, sparsity=True)  # sparsity must be > 0
    with pytest.raises(pagani_amd.PgError):
        ctx.corr_build(x, 10, 1, 0.2)  # T < 2


@pytest.mark.parametrize("prec", PRECS)
def test_plain_workload_n5000(ctx, prec):
    """The CPU-runnable plain workload: N=5000, T=200, K=10, r > 0.4."""
    from pagani_amd import synth
    w = synth.WORKLOADS["plain_n5000_t200"]
    x = synth.bold(w["seed"], w["n"], w["t"], w["k"])
    z64, _ = O.zscore(x)
    say(f"n5000 prec={prec}")
    info, row, col, _ = _build(ctx, x, 0.4, prec=prec)
    orow, ocol, _ = O.corr_csr_r(z64, 0.4)
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    say(f"  nnz={info.nnz} band_pairs={info.band_pairs} flips={info.band_flips} "
        f"gram_ms={info.ms_gram:.3f} csr_ms={info.ms_csr:.3f}")


@pytest.mark.parametrize("n", [65_000, 66_000])
def test_unweighted_rows_match_weighted_rows(ctx, n):
    """Both unweighted row sorts against the weighted one, at full row length.

    The unweighted rows are sorted in LDS bitmaps: at n = 65,000 the bitmaps
    and one base per 128 columns fit 160 KB and every entry is placed by its
    rank; at n = 66,000 only the bitmaps fit and the columns are read back
    lane by lane.  The weighted build sorts its rows with another kernel, so
    the same selection must come out with identical rows and columns, each
    row strictly ascending.
    """
    x = _data(n, 24, 40, 9)
    say(f"row sorts n={n}")
    info, row, col, _ = _build(ctx, x, 0.2, sparsity=True)
    winfo, wrow, wcol, _ = _build(ctx, x, 0.2, sparsity=True, weighted=True)
    assert info.nnz == winfo.nnz == 2 * O.sparsity_k(n, 0.2)
    np.testing.assert_array_equal(row, wrow)
    np.testing.assert_array_equal(col, wcol)
    nnz = int(row[-1])
    inner = np.ones(nnz - 1, dtype=bool)
    b = row[1:-1].astype(np.int64)
    inner[b[(b > 0) & (b < nnz)] - 1] = False
    assert (np.diff(col.astype(np.int64))[inner] > 0).all()


_MULTI = {}


def _multi_oracle(key, fn):
    """Oracle results shared by the multi-threshold cases (one per threshold)."""
    if key not in _MULTI:
        _MULTI[key] = fn()
    return _MULTI[key]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("weighted", [False, True])
def test_multi_threshold_one_gram_pass(ctx, prec, weighted):
    """pg_corr_emit + pg_corr_select: three r thresholds and three sparsities
    from one emitting Gram pass each (N large enough for the sampled emission
    bound), every CSR identical to the oracle's for that threshold alone."""
    import pagani_amd
    n, t = 6000, 100
    x = _multi_oracle("x", lambda: _data(n, t, 12, 13))
    z64 = _multi_oracle("z", lambda: O.zscore(x)[0])
    say(f"multi r prec={prec} weighted={weighted}")
    e = ctx.corr_emit(x, n, t, 0.25, weighted=weighted, prec=prec)
    assert e.gram_passes == 1 and e.band_pairs > 0
    for thr in (0.45, 0.25, 0.3):
        info = ctx.corr_select(thr)
        row, col, val = ctx.corr_csr(n, info.nnz, weighted)
        orow, ocol, oval = _multi_oracle(("r", thr, weighted), lambda: O.corr_csr_r(z64, thr, weighted))
        _assert_same(row, col, val, orow, ocol, oval)
        assert info.r_thresh == np.float32(thr)
        assert info.band_max_err <= info.band / 1.5
    with pytest.raises(pagani_amd.PgError) as ex:
        ctx.corr_select(0.2)  # looser than the emitted threshold
    assert ex.value.status == 5
    say(f"multi s prec={prec} weighted={weighted}")
    e = ctx.corr_emit(x, n, t, 3.0, sparsity=True, weighted=weighted, prec=prec)
    assert e.gram_passes == 1  # the sampled bound held: no histogram of every pair
    for s_ in (3.0, 0.5, 1.0):
        info = ctx.corr_select(s_)
        row, col, val = ctx.corr_csr(n, info.nnz, weighted)
        orow, ocol, oval, tstar = _multi_oracle(("s", s_, weighted), lambda: O.corr_csr_s(z64, s_, weighted))
        _assert_same(row, col, val, orow, ocol, oval)
        assert info.t_star == tstar
        assert info.nnz == 2 * O.sparsity_k(n, s_)
    with pytest.raises(pagani_amd.PgError):
        ctx.corr_select(4.0)  # more pairs than the emission kept
    # a build in between drops the kept pairs
    ctx.corr_build(x, n, t, 0.5, prec=prec)
    with pytest.raises(pagani_amd.PgError) as ex:
        ctx.corr_select(1.0)
    assert ex.value.status == 7


@pytest.mark.parametrize("prec", PRECS)
def test_multi_threshold_small_and_tied(ctx, prec):
    """Emission paths without the sampled bound (few supertiles) and a
    sparsity whose cut splits a tie group."""
    x = _data(400, 120, 6, 9)
    for dst, src in [(50, 10), (51, 10), (300, 10), (201, 77), (399, 77)]:
        x[:, dst] = x[:, src]
    z64, _ = O.zscore(x)
    s_tie = _sparsity_splitting_a_tie(z64, 400)
    say(f"multi small prec={prec} s_tie={s_tie}")
    ctx.corr_emit(x, 400, 120, max(20.0, s_tie), sparsity=True, prec=prec)
    for s in (s_tie, 20.0, 1.0):
        if s > max(20.0, s_tie):
            continue
        info = ctx.corr_select(s)
        row, col, _ = ctx.corr_csr(400, info.nnz)
        orow, ocol, _, tstar = O.corr_csr_s(z64, s)
        np.testing.assert_array_equal(row, orow)
        np.testing.assert_array_equal(col, ocol)
        assert info.t_star == tstar
    for n in (1, 2):
        xs = _data(n, 40, 2, 4)
        ctx.corr_emit(xs, n, 40, 10.0, sparsity=True, prec=prec)
        info = ctx.corr_select(5.0)
        assert info.nnz == 0
        ctx.corr_emit(xs, n, 40, -0.5, prec=prec)
        info = ctx.corr_select(0.0)
        row, col, _ = ctx.corr_csr(n, info.nnz)
        orow, ocol, _ = O.corr_csr_r(O.zscore(xs)[0], 0.0)
        np.testing.assert_array_equal(row, orow)
        np.testing.assert_array_equal(col, ocol)
