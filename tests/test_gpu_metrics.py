"""GPU parity: per-node metrics (libpagani) against the FP64 oracle.

Bit-exact: degree (int and the reference's float), strength, triangles, binary
Cp, BFS hop levels, Floyd-Warshall.  Tolerance (relative, stated per test):
betweenness, weighted clustering, modularity.
"""
import numpy as np
import pytest

from conftest import say
from helpers import cliques, complete_graph, path_graph, random_graph, star_graph
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BC_RTOL = 1e-9      # FP64 dependencies, different summation order
CPW_RTOL = 1e-9     # FP64 sums in a different order, pow on device vs host
Q_ATOL = 1e-9


@pytest.fixture(scope="module")
def plain_graph(ctx):
    """CSR of the N=5000, T=200, K=10, r > 0.4 workload (built on the GPU)."""
    from pagani_amd import synth
    w = synth.WORKLOADS["plain_n5000_t200"]
    x = synth.bold(w["seed"], w["n"], w["t"], w["k"])
    info = ctx.corr_build(x, w["n"], w["t"], 0.4)
    row, col, _ = ctx.corr_csr(w["n"], info.nnz)
    return row, col


@pytest.fixture(scope="module")
def weighted_graph(ctx):
    from pagani_amd import synth
    x = synth.bold(7, 1500, 180, 12)
    info = ctx.corr_build(x, 1500, 180, 0.3, weighted=True)
    return ctx.corr_csr(1500, info.nnz, True)


def _graphs():
    return {
        "path": path_graph(40),
        "star": star_graph(33),
        "complete": complete_graph(20),
        "random": random_graph(300, 0.05, 1)[:2],
        "sparse_disconnected": random_graph(500, 0.004, 2)[:2],
        "cliques": cliques([12, 9, 15, 7]),
    }


def test_degree_and_reference_float(ctx, plain_graph):
    row, col = plain_graph
    say("degree")
    deg, dref = ctx.degree(row)
    od, odr = O.degree(row)
    np.testing.assert_array_equal(deg, od)
    np.testing.assert_array_equal(dref.view(np.uint32), odr.view(np.uint32))
    # offsets past 2^24: the reference's float subtraction rounds
    big = np.array([0, 16777217, 33554439, 33554440], np.int32)
    d2, r2 = ctx.degree(big)
    o2, or2 = O.degree(big)
    np.testing.assert_array_equal(d2, o2)
    np.testing.assert_array_equal(r2.view(np.uint32), or2.view(np.uint32))


def test_strength_bit_exact(ctx, weighted_graph):
    row, col, val = weighted_graph
    say("strength")
    s = ctx.strength(row, val)
    np.testing.assert_array_equal(s.view(np.uint32), O.strength(row, val).view(np.uint32))


@pytest.mark.parametrize("name", list(_graphs()))
def test_triangles_cp_small(ctx, name):
    row, col = _graphs()[name][:2]
    say(f"triangles {name}")
    np.testing.assert_array_equal(ctx.triangles(row, col), O.triangles(row, col))
    cp, mean = ctx.cp_binary(row, col)
    ocp, omean = O.cp_binary(row, col)
    np.testing.assert_array_equal(cp.view(np.uint32), ocp.view(np.uint32))
    assert abs(mean - omean) <= 1e-12 * max(1.0, abs(omean))


def test_triangles_cp_plain(ctx, plain_graph):
    row, col = plain_graph
    say("triangles plain")
    np.testing.assert_array_equal(ctx.triangles(row, col), O.triangles(row, col))
    cp, mean = ctx.cp_binary(row, col)
    ocp, omean = O.cp_binary(row, col)
    np.testing.assert_array_equal(cp.view(np.uint32), ocp.view(np.uint32))


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("wmax_mode", [0, 1])
def test_cp_weighted(ctx, weighted_graph, kind, wmax_mode):
    row, col, val = weighted_graph
    say(f"cpw kind={kind} wmax={wmax_mode}")
    cp, mean = ctx.cp_weighted(row, col, val, kind, wmax_mode)
    ocp, omean = O.cp_weighted(row, col, val, kind, wmax_mode)
    np.testing.assert_allclose(cp, ocp, rtol=CPW_RTOL, atol=1e-14)
    assert abs(mean - omean) <= CPW_RTOL * abs(omean)


@pytest.mark.parametrize("name", list(_graphs()))
def test_bc_small_all_sources(ctx, name):
    row, col = _graphs()[name][:2]
    say(f"bc {name}")
    b = ctx.bc(row, col)
    np.testing.assert_allclose(b, O.bc(row, col), rtol=BC_RTOL, atol=1e-9)


def test_bc_closed_forms(ctx):
    n = 41
    row, col = path_graph(n)[:2]
    b = ctx.bc(row, col)
    i = np.arange(n)
    np.testing.assert_allclose(b, 2.0 * i * (n - 1 - i), rtol=1e-12)
    row, col = star_graph(n)[:2]
    b = ctx.bc(row, col)
    assert b[0] == (n - 1) * (n - 2)
    assert np.all(b[1:] == 0)


def test_bc_plain_sampled(ctx, plain_graph):
    from pagani_amd import synth
    row, col = plain_graph
    src = synth.sample_sources(row.size - 1, 200)
    say("bc plain 200 sources")
    b = ctx.bc(row, col, src)
    np.testing.assert_allclose(b, O.bc(row, col, src), rtol=BC_RTOL, atol=1e-9)


def test_bfs_levels(ctx, plain_graph):
    row, col = plain_graph
    src = np.array([0, 5, 17, 4999, 2500] + list(range(100, 170)), np.int32)
    say("bfs plain")
    lev = ctx.bfs(row, col, src)
    for k, s in enumerate(src):
        np.testing.assert_array_equal(lev[k], O.bfs(row, col, int(s)))
    row, col = _graphs()["sparse_disconnected"]
    lev = ctx.bfs(row, col, np.array([0, 1, 2], np.int32))
    for k in range(3):
        np.testing.assert_array_equal(lev[k], O.bfs(row, col, k))


@pytest.mark.parametrize("block", [64, 128])
@pytest.mark.parametrize("n", [100, 300])
def test_fw_bit_exact(ctx, weighted_graph, block, n):
    row, col, val = weighted_graph
    # induced subgraph on the first n voxels, d = 1/w as the reference
    sub_r = [0]
    sub_c, sub_v = [], []
    for i in range(n):
        a, b = row[i], row[i + 1]
        m = col[a:b] < n
        sub_c += col[a:b][m].tolist()
        sub_v += val[a:b][m].tolist()
        sub_r.append(len(sub_c))
    r = np.array(sub_r, np.int32)
    c = np.array(sub_c, np.int32)
    v = np.array(sub_v, np.float32)
    say(f"fw n={n} block={block}")
    d = ctx.fw_init(r, c, v, 0)
    np.testing.assert_array_equal(d.view(np.uint32), O.fw_init(r, c, v, 0).view(np.uint32))
    got = ctx.fw(d.copy(), block)
    want = O.fw_blocked(d, block)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    # nodal efficiency / Lp (W CUBFW_Lp.cpp:70-113): skip d < 0.1, e < 1/n -> 0
    eff, lp, iso = ctx.fw_efficiency(got)
    with np.errstate(divide="ignore"):
        q = (np.float32(1.0) / want).astype(np.float64)  # 1.0f / d, summed in FP64 for Lp
    inv = np.where((want >= 0.1) & ~np.eye(n, dtype=bool), q, 0.0)
    e_ref = (inv.sum(1) / (n - 1)).astype(np.float32)
    small = e_ref < np.float32(1.0 / n)
    e_ref[small] = 0
    np.testing.assert_allclose(eff, e_ref, rtol=1e-5, atol=1e-7)
    assert iso == int(small.sum())
    assert abs(lp - 1.0 / (inv.sum() / n / (n - 1))) <= 1e-9 * lp


def _check_eff(got, want):
    e, lp, iso, unc = got
    oe, olp, oiso, ounc = want
    np.testing.assert_array_equal(e.view(np.uint32), oe.view(np.uint32))
    # FP64 Lp sums in another grouping (per-source shares, then over sources)
    assert abs(lp - olp) <= 1e-12 * abs(olp)
    assert (iso, unc) == (oiso, ounc)


@pytest.mark.parametrize("name", list(_graphs()))
def test_bfs_efficiency_small(ctx, name):
    row, col = _graphs()[name][:2]
    say(f"bfs efficiency {name}")
    _check_eff(ctx.bfs_efficiency(row, col), O.bfs_efficiency(row, col))


def test_bfs_efficiency_deep_path(ctx):
    row, col = path_graph(300)[:2]  # depth 299: nine level planes
    say("bfs efficiency path 300")
    _check_eff(ctx.bfs_efficiency(row, col), O.bfs_efficiency(row, col))


@pytest.mark.parametrize("batch", [0, 4096])
def test_bfs_efficiency_plain(ctx, plain_graph, batch, monkeypatch):
    row, col = plain_graph
    if batch:
        monkeypatch.setenv("PAGANI_EFF_BATCH", str(batch))  # two batches of sources
    say(f"bfs efficiency plain batch={batch}")
    _check_eff(ctx.bfs_efficiency(row, col), O.bfs_efficiency(row, col))


def test_bfs_efficiency_rejects_bad_csr(ctx):
    row, col = path_graph(10)[:2]
    bad = col.copy()
    bad[3] = 10
    say("bfs efficiency bad csr")
    from pagani_amd.lib import PgError
    with pytest.raises(PgError, match="malformed CSR"):
        ctx.bfs_efficiency(row, bad)
    # rows must list their columns in ascending order
    row, col = star_graph(6)[:2]
    swapped = col.copy()
    swapped[[0, 1]] = swapped[[1, 0]]
    with pytest.raises(PgError, match="ascending"):
        ctx.bfs_efficiency(row, swapped)
    with pytest.raises(PgError, match="ascending"):
        ctx.eigenvector_centrality(row, swapped)


def test_fw_unit_weights_equal_bfs(ctx):
    row, col = random_graph(200, 0.03, 5)[:2]
    d = ctx.fw_init(row, col, None, 2)
    ctx.fw(d, 64)
    for s in range(0, 200, 37):
        lev = O.bfs(row, col, s).astype(np.float32)
        lev[lev < 0] = np.inf
        np.testing.assert_array_equal(d[s], lev)


def test_modularity_cliques(ctx):
    row, col = cliques([20, 20, 20, 20])[:2]
    say("modularity cliques")
    lab, q, lam, eig, nr = ctx.modularity(row, col)
    olab, oq, olam, oeig, onr = O.modularity(row, col)
    np.testing.assert_array_equal(lab, olab)
    assert abs(q - oq) <= Q_ATOL
    # four cliques come out as four modules
    assert len(set(lab[i * 20] for i in range(4))) == 4
    for i in range(4):
        assert len(set(lab[i * 20:(i + 1) * 20])) == 1


def mod_errors(got, want):
    """Errors of a modularity result against the FP64 oracle's: relative
    error of lambda0 and of the round-0 vector (sign-aligned), |Q - Q_oracle|
    and the fraction of vertices with the oracle's label."""
    lab, q, lam, eig, _ = got
    olab, oq, olam, oeig, _ = want
    sgn = 1.0 if float(eig @ oeig) >= 0 else -1.0
    return dict(lam=abs(lam - olam) / abs(olam), eig=float(np.linalg.norm(sgn * eig - oeig) / np.linalg.norm(oeig)),
                q=abs(q - oq), agree=float((lab == olab).mean()))


@pytest.mark.parametrize("stop", [0, 1])
@pytest.mark.parametrize("kmode", [0, 1])
def test_modularity_weighted(ctx, weighted_graph, kmode, stop):
    row, col, val = weighted_graph
    say(f"modularity weighted kmode={kmode} stop={stop}")
    got = ctx.modularity(row, col, val, kmode, stop=stop)
    want = O.modularity(row, col, val, kmode, stop=stop)
    err = mod_errors(got, want)
    print(f"modularity n={row.size - 1} kmode={kmode} stop={stop} rounds={got[4]}/{want[4]}: "
          f"lambda0 rel err {err['lam']:.2e}, eigenvector rel err {err['eig']:.2e}, "
          f"|dQ| {err['q']:.2e}, label agreement {err['agree']:.4f}")
    assert err["lam"] <= 1e-7
    assert err["eig"] <= 1e-5
    assert err["q"] <= 1e-6
    assert err["agree"] > 0.99, err


@pytest.mark.parametrize("depth", [1, 2])
def test_modularity_depth_limit(ctx, weighted_graph, depth):
    import pagani_amd
    row, col, val = weighted_graph
    say(f"modularity depth {depth}")
    got = ctx.modularity(row, col, val, 0, stop=pagani_amd.MOD_STOP_RAYLEIGH, max_depth=depth)
    want = O.modularity(row, col, val, 0, stop=1, max_depth=depth)
    err = mod_errors(got, want)
    assert got[4] == want[4]  # rounds processed
    assert err["lam"] <= 1e-7 and err["q"] <= 1e-6 and err["agree"] > 0.99, err
    assert len(set(got[0].tolist())) <= 2 ** depth


def _check_ec(got, want):
    vh, av, lam, it = got
    ovh, oav, olam, oit = want
    assert it == oit
    assert abs(lam - olam) <= 1e-12 * abs(olam)
    # FP64, sums in another order: the converged vectors agree far below the
    # 1e-6 stop tolerance
    np.testing.assert_allclose(vh, ovh, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(av, oav, rtol=1e-9, atol=1e-12 * abs(olam))


@pytest.mark.parametrize("name", list(_graphs()))
def test_ec_small(ctx, name):
    row, col = _graphs()[name][:2]
    say(f"ec {name}")
    _check_ec(ctx.eigenvector_centrality(row, col), O.ec(row, col))


def test_ec_plain_and_weighted(ctx, plain_graph, weighted_graph):
    row, col = plain_graph
    say("ec plain")
    _check_ec(ctx.eigenvector_centrality(row, col), O.ec(row, col))
    row, col, val = weighted_graph
    say("ec weighted")
    _check_ec(ctx.eigenvector_centrality(row, col, val), O.ec(row, col, val))


def test_pc_bit_exact(ctx, plain_graph, weighted_graph):
    row, col, val = weighted_graph
    lab = O.modularity(row, col, val, 1)[0]
    say("pc weighted")
    p, q, k = ctx.participation(row, col, lab, val)
    op, oq, ok = O.pc(row, col, lab, val)
    assert k == ok
    np.testing.assert_array_equal(p.view(np.uint32), op.view(np.uint32))
    np.testing.assert_array_equal(q.view(np.uint32), oq.view(np.uint32))
    row, col = plain_graph
    lab = (np.arange(row.size - 1) % 37 + 1).astype(np.int32)
    say("pc plain, 37 modules")
    p, q, k = ctx.participation(row, col, lab)
    op, oq, ok = O.pc(row, col, lab)
    np.testing.assert_array_equal(p.view(np.uint32), op.view(np.uint32))
    np.testing.assert_array_equal(q.view(np.uint32), oq.view(np.uint32))
    lab[5] = 0
    from pagani_amd.lib import PgError
    with pytest.raises(PgError, match="labels"):
        ctx.participation(row, col, lab)


def _same(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


@pytest.mark.parametrize("edges", [[], [(0, 1)]])
@pytest.mark.parametrize("n", [1, 2, 3])
def test_edge_cases_tiny_graphs(ctx, n, edges):
    from helpers import csr_from_edges
    edges = [e for e in edges if max(e) < n]
    row, col = csr_from_edges(n, edges)[:2]
    say(f"tiny graph n={n} edges={edges}")
    e, lp, iso, unc = ctx.bfs_efficiency(row, col)
    oe, olp, oiso, ounc = O.bfs_efficiency(row, col)
    assert _same(e, oe) and _same([lp], [olp]) and (iso, unc) == (oiso, ounc)
    # power iteration: no edges leave u = 0 after one step (x = 0/0 on the next)
    vh, av, lam, it = ctx.eigenvector_centrality(row, col)
    ovh, oav, olam, oit = O.ec(row, col)
    assert it == oit and _same(vh, ovh) and _same(av, oav) and _same([lam], [olam])
    # one module: the normalised coefficient is the coefficient
    lab = np.ones(n, np.int32)
    p, q, k = ctx.participation(row, col, lab)
    op, oq, ok = O.pc(row, col, lab)
    assert k == ok == 1
    np.testing.assert_array_equal(p.view(np.uint32), op.view(np.uint32))
    np.testing.assert_array_equal(q.view(np.uint32), p.view(np.uint32))


def _malformed():
    """(name, row, col, val) cases every CSR entry point must reject with
    PG_EFORMAT before a kernel indexes with them."""
    row, col, val = random_graph(50, 0.2, 9, weighted=True)
    bad_col = col.copy()
    bad_col[7] = 50  # column == n
    neg_col = col.copy()
    neg_col[3] = -1
    bad_rn = row.copy()
    bad_rn[-1] = col.size - 1  # row[n] != nnz
    bad_mono = row.copy()
    bad_mono[10], bad_mono[11] = bad_mono[11], bad_mono[10]  # offsets decrease
    unsorted = col.copy()
    a = row[4]
    unsorted[a], unsorted[a + 1] = unsorted[a + 1], unsorted[a]
    return [("col_ge_n", row, bad_col, val, False), ("col_negative", row, neg_col, val, False),
            ("row_n_ne_nnz", bad_rn, col, val, False), ("row_decreasing", bad_mono, col, val, False),
            ("unsorted_row", row, unsorted, val, True)]


@pytest.mark.parametrize("case", [c[0] for c in _malformed()])
def test_malformed_csr_rejected(ctx, case):
    import pagani_amd
    name, row, col, val, sort_only = next(c for c in _malformed() if c[0] == case)
    say(f"malformed csr {name}")
    calls = {
        "triangles": lambda: ctx.triangles(row, col),
        "cp_binary": lambda: ctx.cp_binary(row, col),
        "cp_weighted": lambda: ctx.cp_weighted(row, col, val),
        "bc": lambda: ctx.bc(row, col, np.array([0, 1], np.int32)),
        "bfs": lambda: ctx.bfs(row, col, np.array([0, 1], np.int32)),
        "modularity": lambda: ctx.modularity(row, col, val),
        "fw_init": lambda: ctx.fw_init(row, col, val, 0),
        "bfs_efficiency": lambda: ctx.bfs_efficiency(row, col),
    }
    # unsorted columns are an error only where the kernels rely on the order
    needs_sorted = {"triangles", "cp_binary", "cp_weighted", "bfs_efficiency"}
    for fn, call in calls.items():
        if sort_only and fn not in needs_sorted:
            continue
        with pytest.raises(pagani_amd.PgError) as e:
            call()
        assert e.value.status == 4, (fn, str(e.value))
    if name == "row_decreasing":  # strength reads the offsets only
        with pytest.raises(pagani_amd.PgError) as e:
            ctx.strength(row, val)
        assert e.value.status == 4
    # the context stays usable
    r, c = path_graph(5)[:2]
    np.testing.assert_allclose(ctx.bc(r, c), O.bc(r, c))
