"""Autotuner: every launch shape gives the reference's results (edge sets,
BFS levels, FW distances bit-exact; BC to FP64 summation order), searches
store their choice in the cache file and later calls reuse it, and
PAGANI_AUTOTUNE=1 searches on first use."""
import numpy as np
import pytest

from conftest import say
from helpers import random_graph
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def x6000():
    from pagani_amd import synth
    x = synth.bold(13, 6000, 100, 12)
    return x, O.zscore(x)[0]


@pytest.mark.parametrize("sup", ["4", "16"])
@pytest.mark.parametrize("prec", [0, 1])
def test_gram_supertile_shapes_keep_the_edge_set(ctx, x6000, monkeypatch, sup, prec):
    x, z64 = x6000
    monkeypatch.setenv("PAGANI_GRAM_SUPER", sup)
    say(f"gram super={sup} prec={prec}")
    info = ctx.corr_build(x, 6000, 100, 1.0, sparsity=True, prec=prec)
    row, col, _ = ctx.corr_csr(6000, info.nnz)
    orow, ocol, _, tstar = O.corr_csr_s(z64, 1.0)
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    assert info.t_star == tstar
    info = ctx.corr_build(x, 6000, 100, 0.35, prec=prec)
    row, col, _ = ctx.corr_csr(6000, info.nnz)
    orow, ocol, _ = O.corr_csr_r(z64, 0.35)
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)


@pytest.mark.parametrize("nb", ["1", "3", "16"])
def test_bc_batch_shapes(ctx, monkeypatch, nb):
    row, col, _ = random_graph(700, 0.03, 17)
    monkeypatch.setenv("PAGANI_BC_BATCHES", nb)
    src = np.arange(0, 700, 3, dtype=np.int32)  # 234 sources: a ragged last group
    say(f"bc batches={nb}")
    np.testing.assert_allclose(ctx.bc(row, col, src), O.bc(row, col, src), rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(ctx.bc(row, col), O.bc(row, col), rtol=1e-9, atol=1e-9)
    lev = ctx.bfs(row, col, src[:130])
    for k in (0, 63, 64, 129):
        np.testing.assert_array_equal(lev[k], O.bfs(row, col, int(src[k])))


def test_autotuner_searches_stores_and_reuses(ctx, x6000, monkeypatch, tmp_path):
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("PAGANI_TUNE_CACHE", str(cache))
    x, z64 = x6000
    say("autotune corr")
    r = ctx.autotune_corr(x, 6000, 100, 1.0, sparsity=True)
    assert r.value in (4, 8, 16) and r.candidates == 3 and r.ms > 0
    r2 = ctx.autotune_corr(x, 6000, 100, 1.0, sparsity=True)
    assert r2.value == r.value and r2.candidates == 0
    row, col, _ = random_graph(900, 0.02, 5)
    src = np.arange(0, 900, 2, dtype=np.int32)
    say("autotune bc")
    r = ctx.autotune_bc(row, col, src)
    assert 1 <= r.value <= 16 and r.candidates >= 2
    assert ctx.autotune_bc(row, col, src).candidates == 0
    np.testing.assert_allclose(ctx.bc(row, col, src), O.bc(row, col, src), rtol=1e-9, atol=1e-9)
    d = ctx.fw_init(row, col, None, 2)
    say("autotune fw")
    keep = d.copy()
    r = ctx.autotune_fw(d)
    assert r.value in (64, 128) and r.candidates == 2
    np.testing.assert_array_equal(d, keep)  # the search works on a copy
    got = ctx.fw(d.copy(), 0)
    np.testing.assert_array_equal(got.view(np.uint32), O.fw_blocked(keep, r.value).view(np.uint32))
    lines = cache.read_text().splitlines()
    assert {ln.split()[1] for ln in lines} == {"gram_super", "bc_batches", "fw_block"}
    # PAGANI_TUNE=off: no cache, defaults
    monkeypatch.setenv("PAGANI_TUNE", "off")
    assert ctx.autotune_corr(x, 6000, 100, 1.0, sparsity=True).candidates == 3


def test_autotune_on_first_use(ctx, monkeypatch, tmp_path):
    from pagani_amd import synth
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("PAGANI_TUNE_CACHE", str(cache))
    monkeypatch.setenv("PAGANI_AUTOTUNE", "1")
    x = synth.bold(3, 3000, 90, 8)
    say("autotune first use")
    info = ctx.corr_build(x, 3000, 90, 0.3)
    row, col, _ = ctx.corr_csr(3000, info.nnz)
    orow, ocol, _ = O.corr_csr_r(O.zscore(x)[0], 0.3)
    np.testing.assert_array_equal(col, ocol)
    ctx.bc(row, col, np.arange(200, dtype=np.int32))
    knobs = {ln.split()[1] for ln in cache.read_text().splitlines()}
    assert knobs == {"gram_super", "bc_batches"}
