"""GPU degree-preserving rewiring (pg_rewire) against the sequential oracle.

A random rewiring has no bit-exact counterpart: the reference seeds rand()
with the wall clock, and the GPU makes its attempts in rounds of disjoint edge
pairs.  The bar here: every invariant of a degree-preserving rewiring holds
exactly (helpers.check_rewired), graphs that admit no swap come back
unchanged, the result is a function of the seed alone, and the statistics of
the rewired networks agree with the sequential oracle within the tolerances
below.
"""
import numpy as np
import pytest

from conftest import say
from helpers import check_rewired, complete_graph, csr_from_edges, random_graph, ring_lattice, star_graph
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CP_ATOL = 0.01    # mean binary clustering over 6 rewired lattices, GPU vs oracle
ACC_RTOL = 0.03   # swaps / attempts, GPU vs oracle


def _sparse_random(n, m, seed):
    """About m undirected edges among n vertices, without forming n^2 pairs."""
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    keep = a != b
    lo, hi = np.minimum(a[keep], b[keep]), np.maximum(a[keep], b[keep])
    key = np.unique(lo.astype(np.int64) * n + hi)
    lo, hi = key // n, key % n
    r = np.concatenate([lo, hi])
    c = np.concatenate([hi, lo])
    o = np.lexsort((c, r))
    row = np.zeros(n + 1, np.int32)
    np.cumsum(np.bincount(r, minlength=n), out=row[1:])
    return row, c[o].astype(np.int32)


@pytest.mark.parametrize("weighted", [False, True])
def test_invariants(ctx, weighted):
    row, col, val = random_graph(2000, 0.02, 11, weighted=weighted)
    v = val if weighted else None
    say(f"rewire n=2000 nnz={col.size} weighted={weighted}")
    oc, ov, sw = ctx.rewire(row, col, v, seed=5)
    check_rewired(row, col, v, oc, ov)
    m = col.size // 2
    assert 0.5 * m < sw <= 2 * m
    assert not np.array_equal(oc, col)


def test_correlation_networks(ctx):
    from pagani_amd import synth
    w = synth.WORKLOADS["plain_n5000_t200"]
    x = synth.bold(w["seed"], w["n"], w["t"], w["k"])
    info = ctx.corr_build(x, w["n"], w["t"], 0.4)
    row, col, _ = ctx.corr_csr(w["n"], info.nnz)
    say(f"rewire correlation network n={w['n']} nnz={col.size}")
    oc, _, sw = ctx.rewire(row, col, seed=1)
    check_rewired(row, col, None, oc, None)
    assert sw > 0
    x = synth.bold(7, 1500, 180, 12)
    info = ctx.corr_build(x, 1500, 180, 0.3, weighted=True)
    row, col, val = ctx.corr_csr(1500, info.nnz, True)
    oc, ov, sw = ctx.rewire(row, col, val, seed=2)
    check_rewired(row, col, val, oc, ov)


def test_result_is_a_function_of_the_seed(ctx):
    row, col, val = random_graph(800, 0.03, 4, weighted=True)
    a = ctx.rewire(row, col, val, seed=9)
    b = ctx.rewire(row, col, val, seed=9)
    c = ctx.rewire(row, col, val, seed=10)
    assert a[2] == b[2]
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert not np.array_equal(a[0], c[0])


def test_no_swap_possible(ctx):
    for row, col, _ in (star_graph(9), complete_graph(7), star_graph(2)):
        oc, _, sw = ctx.rewire(row, col, seed=3)
        assert sw == 0
        np.testing.assert_array_equal(oc, col)


def test_two_edges_reach_every_matching(ctx):
    row, col, _ = csr_from_edges(4, [(0, 1), (2, 3)])
    seen = {}
    for s in range(120):
        oc, _, _ = ctx.rewire(row, col, seed=s)
        check_rewired(row, col, None, oc, None)
        seen[tuple(oc.tolist())] = seen.get(tuple(oc.tolist()), 0) + 1
    assert len(seen) == 3 and min(seen.values()) > 15, seen


def test_statistics_match_oracle(ctx):
    # ring lattice n = 2000, k = 20 (1% density, Cp = 0.711): at the default
    # 2M attempts both rewirings must accept swaps at the same rate and leave
    # the same (random-level) clustering.  At smaller budgets the two differ by
    # design: a GPU round puts every edge in exactly one attempt, the
    # sequential draw with replacement leaves some edges untouched.
    row, col, _ = ring_lattice(2000, 20)
    m = col.size // 2
    g_cp, o_cp, g_acc, o_acc = [], [], [], []
    for s in range(6):
        oc, _, sw = ctx.rewire(row, col, seed=100 + s)
        g_cp.append(O.cp_binary(row, oc)[1])
        g_acc.append(sw / (2 * m))
        oc, _, sw = O.rewire(row, col, None, seed=100 + s)
        o_cp.append(O.cp_binary(row, oc)[1])
        o_acc.append(sw / (2 * m))
    say(f"lattice: Cp gpu {np.mean(g_cp):.4f} oracle {np.mean(o_cp):.4f}; "
        f"acceptance gpu {np.mean(g_acc):.4f} oracle {np.mean(o_acc):.4f}")
    assert abs(np.mean(g_cp) - np.mean(o_cp)) < CP_ATOL
    assert abs(np.mean(g_acc) / np.mean(o_acc) - 1) < ACC_RTOL


def test_long_rows(ctx):
    # a hub joined to every vertex (row of 2999 > 1024: the LDS-bitmap sort)
    # over a dense random graph (rows of 500-700: the 16-per-lane sort)
    n = 3000
    row0, col0, _ = random_graph(n, 0.2, 8)
    edges = [(i, int(j)) for i in range(n) for j in col0[row0[i]:row0[i + 1]] if j > i]
    edges += [(0, j) for j in range(1, n)]
    row, col, val = csr_from_edges(n, edges, np.linspace(0.1, 1.0, len(edges)))
    assert row[1] - row[0] == n - 1
    say(f"rewire hub graph nnz={col.size}")
    oc, ov, sw = ctx.rewire(row, col, val, seed=4)
    check_rewired(row, col, val, oc, ov)
    np.testing.assert_array_equal(oc[row[0]:row[1]], np.arange(1, n))  # the hub keeps every neighbour
    assert sw > 0


def test_large_sparse_network(ctx):
    row, col = _sparse_random(20000, 1_000_000, 3)
    say(f"rewire n=20000 nnz={col.size}")
    oc, _, sw = ctx.rewire(row, col, seed=8)
    check_rewired(row, col, None, oc, None)
    assert sw > 0.9 * col.size  # 0.5% dense: nearly all of the 2M attempts succeed


def test_attempt_count_and_device_tensors(ctx):
    import torch
    row, col, val = random_graph(500, 0.05, 6, weighted=True)
    oc0, _, sw0 = ctx.rewire(row, col, val, seed=1, attempts=0)
    oc1, _, sw1 = ctx.rewire(row, col, val, seed=1, attempts=col.size)  # = 2M, the default
    assert sw0 == sw1
    np.testing.assert_array_equal(oc0, oc1)
    _, _, few = ctx.rewire(row, col, val, seed=1, attempts=10)
    assert 0 <= few <= 10
    d = [torch.from_numpy(a).cuda() for a in (row, col, val)]
    dc = torch.empty_like(d[1])
    dv = torch.empty_like(d[2])
    _, _, sw2 = ctx.rewire(d[0], d[1], d[2], seed=1, out_col=dc, out_val=dv)
    ctx.synchronize()
    assert sw2 == sw0
    np.testing.assert_array_equal(dc.cpu().numpy(), oc0)


def test_rejects_bad_input(ctx):
    from pagani_amd.lib import PgError
    row, col, _ = random_graph(50, 0.2, 1)
    # drop one mirror entry: not symmetric
    r2 = row.copy()
    c2 = np.delete(col, row[0])
    r2[1:] -= 1
    with pytest.raises(PgError, match="symmetric"):
        ctx.rewire(r2, c2, seed=1)
    # a self-loop on vertex 0
    c3 = np.insert(col, row[0], 0).astype(np.int32)
    r3 = row.copy()
    r3[1:] += 1
    c3[row[0]:r3[1]] = np.sort(c3[row[0]:r3[1]])
    with pytest.raises(PgError, match="self-loop"):
        ctx.rewire(r3, c3, seed=1)
    # a row out of order
    c4 = col.copy()
    c4[row[0]], c4[row[0] + 1] = c4[row[0] + 1], c4[row[0]]
    with pytest.raises(PgError, match="ascending"):
        ctx.rewire(row, c4, seed=1)


def test_full_size_network(ctx):
    # the N=58,000, T=240, 1%-sparsity network (33.6M entries): every invariant
    # at full size, and nearly all of the 2M attempts succeed at this density
    from pagani_amd import synth
    w = synth.WORKLOADS["corr_n58000_t240"]
    x = synth.bold(w["seed"], w["n"], w["t"], w["k"])
    info = ctx.corr_build(x, w["n"], w["t"], 1.0, sparsity=True)
    row, col, _ = ctx.corr_csr(w["n"], info.nnz)
    say(f"rewire full-size network n={w['n']} nnz={col.size}")
    oc, _, sw = ctx.rewire(row, col, seed=12)
    check_rewired(row, col, None, oc, None)
    assert sw > 0.9 * col.size
