"""Pin the FP64 oracle: closed forms, numpy FP64, the compiled reference."""
import numpy as np
import pytest

from helpers import (check_rewired, cliques, complete_graph, csr_from_edges, path_graph, random_graph,
                     ring_lattice, star_graph)
from oracle import oracle as O


def _x(n=300, t=120, k=6, seed=2):
    from pagani_amd import synth
    return synth.bold(seed, n, t, k)


def test_zscore_closed_forms():
    x = _x()
    z64, z32 = O.zscore(x)
    assert np.abs(z64.mean(1)).max() < 1e-15
    np.testing.assert_allclose(np.linalg.norm(z64, axis=1), 1.0, atol=1e-14)
    np.testing.assert_array_equal(z32, z64.astype(np.float32))
    # z of an affine image is the same z (up to rounding); of a negation, -z
    y = x.copy()
    y[:, 1] = 3.0 * x[:, 0] + 7.0
    y[:, 2] = -2.0 * x[:, 0]
    zz, _ = O.zscore(y)
    np.testing.assert_allclose(zz[1], zz[0], atol=1e-6)
    np.testing.assert_allclose(zz[2], -zz[0], atol=1e-6)


def test_zscore_matches_numpy_fp64():
    x = _x()
    xd = x.astype(np.float64)
    mu = xd.mean(0)
    d = xd - mu
    zref = (d / np.sqrt((d * d).sum(0))).T
    z64, _ = O.zscore(x)
    np.testing.assert_allclose(z64, zref, atol=1e-15, rtol=1e-13)


def test_zero_variance_voxel_has_no_edges():
    x = _x()
    x[:, 5] = 1.0
    z64, _ = O.zscore(x)
    assert np.all(z64[5] == 0)
    row, col, _ = O.corr_csr_r(z64, 0.0)
    assert row[6] == row[5]


def test_r_matches_corrcoef():
    x = _x()
    z64, _ = O.zscore(x)
    cc = np.corrcoef(x.T.astype(np.float64))
    rng = np.random.default_rng(0)
    for _ in range(200):
        i, j = sorted(rng.choice(300, 2, replace=False))
        assert abs(O.r_pair(z64, i, j) - cc[i, j]) < 1e-12


def test_r_closed_forms():
    rng = np.random.default_rng(1)
    base = rng.standard_normal(64)
    x = np.stack([base, 2.5 * base + 1.0, -0.5 * base + 3.0, np.cos(np.arange(64) * 2 * np.pi / 64),
                  np.sin(np.arange(64) * 2 * np.pi / 64)], 1).astype(np.float32)
    z64, _ = O.zscore(x)
    assert abs(O.r_pair(z64, 0, 1) - 1.0) < 1e-6
    assert abs(O.r_pair(z64, 0, 2) + 1.0) < 1e-6
    assert abs(O.r_pair(z64, 3, 4)) < 1e-6  # orthogonal sinusoids


def test_rmode_threshold_brute_force():
    x = _x(250, 90, 5, 8)
    z64, _ = O.zscore(x)
    for thr, weighted in [(0.3, False), (0.3, True), (0.05, False), (-0.1, False)]:
        row, col, val = O.corr_csr_r(z64, thr, weighted)
        the = O.threshold_eff(thr, weighted)
        dense = np.zeros((250, 250), bool)
        for i in range(250):
            for j in range(i + 1, 250):
                if O.r_pair(z64, i, j) > the:
                    dense[i, j] = dense[j, i] = True
        rr, cc = np.nonzero(dense)
        np.testing.assert_array_equal(col, cc)
        np.testing.assert_array_equal(np.diff(row), dense.sum(1))


def test_rows_alone_equal_whole_csr():
    # the per-row form (used where the whole matrix is too large) gives the
    # same rows, values and order as the whole CSR, which is pinned above
    x = _x(260, 80, 5, 9)
    z64, _ = O.zscore(x)
    rows = np.array([0, 1, 77, 130, 259], np.int32)
    for thr, weighted in [(0.2, False), (0.2, True), (-0.05, True)]:
        row, col, val = O.corr_csr_r(z64, thr, weighted)
        got = O.corr_rows_r(z64, thr, rows, weighted)
        for q, i in enumerate(rows):
            np.testing.assert_array_equal(got[q][0], col[row[i]:row[i + 1]])
            if weighted:
                np.testing.assert_array_equal(got[q][1].view(np.uint32), val[row[i]:row[i + 1]].view(np.uint32))


def test_weighted_cp_at_vertices_equals_whole():
    row, col, val = random_graph(120, 0.2, 5, weighted=True)
    verts = np.array([3, 0, 119, 57, 57], np.int32)
    for kind in (1, 2):
        for wm in (0, 1):
            full, _ = O.cp_weighted(row, col, val, kind, wm)
            np.testing.assert_array_equal(O.cp_weighted_at(row, col, val, verts, kind, wm), full[verts])


def test_weighted_threshold_is_float_rounded():
    # W CorMat.cpp:310 passes r_thresh - 1e-6 into a float parameter
    for thr in [0.2, 0.25, 0.3, 0.504]:
        assert O.threshold_eff(thr, True) == np.float32(np.float64(np.float32(thr)) - 1e-6)
        assert O.threshold_eff(thr, False) == np.float32(thr)


def test_sparsity_k_reference_arithmetic():
    # U CorMat.cpp:311-312: (long long) 2*M1*s/100.0 is (float)(2*M1)*s in float,
    # then /100.0 in double, truncated, then rounded up to even.
    for n, s in [(5000, 1.0), (58000, 1.0), (58523, 0.5), (120000, 2.0), (25218, 1.0), (640, 7.5)]:
        m1 = n * (n - 1) // 2
        prod = np.float32(np.float32(2 * m1) * np.float32(s))
        clen = int(np.float64(prod) / 100.0)
        clen += clen % 2
        assert O.sparsity_k(n, s) == clen // 2


def test_smode_count_and_tie_rule():
    x = _x(200, 80, 4, 3)
    x[:, 20] = x[:, 3]
    x[:, 21] = x[:, 3]
    z64, _ = O.zscore(x)
    iu = np.triu_indices(200, 1)
    r = np.array([O.r_pair(z64, i, j) for i, j in zip(*iu)])
    for s in [0.5, 1.0, 3.3, 10.0]:
        k = O.sparsity_k(200, s)
        row, col, _, tstar = O.corr_csr_s(z64, s)
        assert row[-1] == 2 * k
        srt = np.sort(r)[::-1]
        assert tstar == srt[k - 1]
        # brute force: all r > t*, then ties in row-major order
        keep = r > tstar
        ties = np.nonzero(r == tstar)[0]
        keep[ties[: k - keep.sum()]] = True
        dense = np.zeros((200, 200), bool)
        dense[iu[0][keep], iu[1][keep]] = True
        dense |= dense.T
        np.testing.assert_array_equal(col, np.nonzero(dense)[1])


def test_degree_float_quirk():
    row = np.array([0, 3, 16777217, 16777220, 33554439], np.int32)
    d, dr = O.degree(row)
    np.testing.assert_array_equal(d, np.diff(row))
    ref = np.float32(row[1:].astype(np.float32) - row[:-1].astype(np.float32))
    np.testing.assert_array_equal(dr, ref)
    assert dr[1] != d[1]  # the reference's float subtraction is inexact here


def test_strength_sequential_float():
    row, col, val = random_graph(120, 0.1, 3, weighted=True)
    s = O.strength(row, val)
    for i in range(120):
        acc = np.float32(0)
        for e in range(row[i], row[i + 1]):
            acc = np.float32(acc + val[e])
        assert s[i] == acc


def _dense(row, col):
    n = row.size - 1
    a = np.zeros((n, n))
    for i in range(n):
        a[i, col[row[i]:row[i + 1]]] = 1
    return a


@pytest.mark.parametrize("g", ["path", "star", "complete", "random", "cliques"])
def test_triangles_trace(g):
    row, col = {"path": path_graph(30), "star": star_graph(20), "complete": complete_graph(15),
                "random": random_graph(150, 0.1, 4), "cliques": cliques([8, 6, 9])}[g][:2]
    a = _dense(row, col)
    np.testing.assert_array_equal(O.triangles(row, col), np.round(np.diag(a @ a @ a) / 2).astype(np.int64))
    cp, _ = O.cp_binary(row, col)
    if g == "complete":
        assert np.all(cp == 1.0)
    if g in ("star", "path"):
        assert np.all(cp == 0.0)


def test_bc_closed_forms():
    n = 25
    row, col = path_graph(n)[:2]
    i = np.arange(n)
    np.testing.assert_allclose(O.bc(row, col), 2.0 * i * (n - 1 - i), rtol=1e-14)
    row, col = star_graph(n)[:2]
    b = O.bc(row, col)
    assert b[0] == (n - 1) * (n - 2) and np.all(b[1:] == 0)
    row, col = complete_graph(12)[:2]
    assert np.all(O.bc(row, col) == 0)


def test_bc_sum_identity():
    # sum_v BC(v) = sum over ordered reachable pairs (s != t) of (d(s,t) - 1)
    row, col = random_graph(120, 0.03, 9)[:2]
    b = O.bc(row, col)
    tot = 0
    for s in range(120):
        lev = O.bfs(row, col, s)
        tot += int(np.sum(lev[lev > 0] - 1))
    assert abs(b.sum() - tot) <= 1e-9 * tot


def test_bc_matches_compiled_reference():
    ref = O.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    row, col = random_graph(300, 0.04, 12)[:2]
    b = O.bc(row, col)
    rb = np.empty(300, np.float32)
    ref.ref_bc_all(row, col, 300, rb)
    np.testing.assert_allclose(rb, b, rtol=2e-5, atol=1e-3)


@pytest.mark.parametrize("b", [32, 64])
def test_fw_matches_compiled_reference_bit_exact(b):
    ref = O.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    row, col, val = random_graph(2 * b, 0.05, 13, weighted=True)
    n = 2 * b
    d_ref = np.empty((n, n), np.float32)
    ref.ref_fw_init(row, col, val, n, b, d_ref)
    d = O.fw_init(row, col, val, 0)
    np.testing.assert_array_equal(d.view(np.uint32), d_ref.view(np.uint32))
    ref.ref_fw_dense(d_ref, n, b)
    np.testing.assert_array_equal(O.fw_blocked(d, b).view(np.uint32), d_ref.view(np.uint32))


def test_fw_unit_weights_equal_bfs():
    row, col = random_graph(90, 0.05, 6)[:2]
    d = O.fw_blocked(O.fw_init(row, col, None, 2), 32)
    for s in range(0, 90, 11):
        lev = O.bfs(row, col, s).astype(np.float32)
        lev[lev < 0] = np.inf
        np.testing.assert_array_equal(d[s], lev)


def test_bfs_efficiency_closed_forms():
    n = 37
    row, col = complete_graph(n)[:2]
    eff, lp, iso, unc = O.bfs_efficiency(row, col)
    assert np.all(eff == 1.0) and lp == 1.0 and iso == 0 and unc == 0
    row, col = star_graph(n)[:2]  # centre 0, n - 1 leaves
    eff, lp, iso, unc = O.bfs_efficiency(row, col)
    assert eff[0] == 1.0
    assert np.all(eff[1:] == np.float32((1 + (n - 2) * 0.5) / (n - 1)))
    want = (n * (n - 1) / 2) / ((n - 1) + (n - 1) * (n - 2) / 2 * 0.5)
    assert abs(lp - want) <= 1e-15 * want
    row, col = path_graph(n)[:2]
    eff, lp, iso, unc = O.bfs_efficiency(row, col)
    exact = sum((n - k) / k for k in range(1, n))
    assert abs(lp - (n * (n - 1) / 2) / exact) <= 1e-6 * lp
    # two components and an isolated vertex
    r1, c1 = complete_graph(5)[:2]
    row = np.concatenate([r1, r1[1:] + r1[-1], [2 * r1[-1]]]).astype(np.int32)
    col = np.concatenate([c1, c1 + 5]).astype(np.int32)
    eff, lp, iso, unc = O.bfs_efficiency(row, col)
    assert unc == 1 and iso == 1 and eff[10] == 0
    assert np.all(eff[:10] == np.float32(np.float32(4) / np.float32(10)))


def _dense(row, col, val=None):
    n = row.size - 1
    a = np.zeros((n, n))
    for i in range(n):
        a[i, col[row[i]:row[i + 1]]] = 1.0 if val is None else val[row[i]:row[i + 1]]
    return a


def test_ec_leading_pair_matches_eigh():
    for seed, wt in [(3, False), (4, True)]:
        row, col, val = random_graph(160, 0.07, seed)
        v = val if wt else None
        vh, av, lam, it = O.ec(row, col, v)
        w, vec = np.linalg.eigh(_dense(row, col, v))
        lead = vec[:, -1] * np.sign(vec[:, -1].sum())
        assert abs(lam - w[-1]) <= 1e-9 * w[-1]
        np.testing.assert_allclose(vh, lead, atol=1e-6)
        np.testing.assert_allclose(av, lam * vh, atol=1e-5 * lam)


def test_ec_closed_forms():
    # odd cycle (2-regular, not bipartite): eigenvalue 2, uniform vector
    n = 31
    row = np.array([2 * i for i in range(n + 1)], np.int32)
    col = np.array([c for i in range(n) for c in sorted(((i - 1) % n, (i + 1) % n))], np.int32)
    vh, av, lam, it = O.ec(row, col)
    assert abs(lam - 2.0) < 1e-6 and it < 10000
    np.testing.assert_allclose(vh, 1 / np.sqrt(n), atol=1e-6)
    row, col = complete_graph(12)[:2]
    vh, av, lam, it = O.ec(row, col)
    assert abs(lam - 11.0) < 1e-9
    # a star is bipartite: the iteration oscillates and runs to the cap
    row, col = star_graph(9)[:2]
    assert O.ec(row, col)[3] == 10000


def test_pc_closed_forms_and_fp64():
    # each clique its own module: every edge stays inside, pc = 0
    lab = np.repeat(np.arange(1, 4), [6, 9, 5]).astype(np.int32)
    row, col = cliques([6, 9, 5], bridge=False)[:2]
    p, q, k = O.pc(row, col, lab)
    assert k == 3 and np.all(p == 0) and np.all(q == 0)
    # with bridges 0-6 and 6-15: vertex 0 has 5 + 1 edges, vertex 6 has 1 + 8 + 1
    row, col = cliques([6, 9, 5])[:2]
    p, q, k = O.pc(row, col, lab)
    assert abs(p[0] - (1 - (5 / 6) ** 2 - (1 / 6) ** 2)) < 1e-7
    assert abs(p[6] - (1 - (8 / 10) ** 2 - 2 * (1 / 10) ** 2)) < 1e-7
    # star centre whose 12 leaves are spread evenly over 4 modules
    row, col = star_graph(13)[:2]
    lab = np.array([1] + [1 + (i % 4) for i in range(12)], np.int32)
    p, q, k = O.pc(row, col, lab)
    assert k == 4 and abs(p[0] - 0.75) < 1e-7 and abs(q[0] - 1.0) < 1e-6
    assert np.all(p[1:] == 0)  # a leaf has one neighbour
    # FP64 brute force on a weighted graph
    row, col, val = random_graph(150, 0.08, 9)
    lab = (np.arange(150) % 7 + 1).astype(np.int32)
    p, q, k = O.pc(row, col, lab, val)
    a = _dense(row, col, val)
    s = a.sum(1)
    share = np.stack([a[:, lab == j].sum(1) for j in range(1, 8)], 1) / np.where(s > 0, s, 1)[:, None]
    want = np.where(s > 0, 1 - (share ** 2).sum(1), 0)
    np.testing.assert_allclose(p, want, atol=2e-6)
    np.testing.assert_allclose(q, want / (6 / 7), atol=3e-6)


def test_modularity_cliques_split_exactly():
    row, col = cliques([10, 10, 10, 10])[:2]
    lab, q, lam, eig, nr = O.modularity(row, col)
    for i in range(4):
        assert len(set(lab[i * 10:(i + 1) * 10])) == 1
    assert len(set(lab)) == 4
    assert abs(q - O.modularity_q(row, col, lab)) < 1e-15


def test_modularity_leading_vector_matches_eigh():
    row, col = cliques([12, 9, 15])[:2]
    n = row.size - 1
    a = _dense(row, col)
    k = a.sum(1)
    bmat = a - np.outer(k, k) / k.sum()
    w, v = np.linalg.eigh(bmat)
    lab, q, lam, eig, nr = O.modularity(row, col)
    assert abs(lam - w[-1]) < 1e-5
    vv = v[:, -1] * np.sign(v[:, -1] @ eig)
    np.testing.assert_allclose(eig, vv, atol=1e-4)


def _bipartites(spec):
    """Disjoint complete bipartite graphs K_{a,b}."""
    edges, o = [], 0
    for a, b in spec:
        edges += [(o + i, o + a + j) for i in range(a) for j in range(b)]
        o += a + b
    return csr_from_edges(o, edges)


@pytest.mark.parametrize("graph", ["cliques", "gnp_weighted", "bipartite"])
def test_modularity_rayleigh_stop_matches_eigh(graph):
    # the quotient rule: lambda0 is the final quotient, which converges to the
    # largest eigenvalue of B; the complete bipartite graph's B has a dominant
    # negative eigenvalue, so its round 0 takes the shifted second pass
    if graph == "cliques":
        row, col, val = cliques([12, 9, 15])
        val = None
    elif graph == "gnp_weighted":
        row, col, val = random_graph(90, 0.12, 4, weighted=True)
    else:
        row, col, val = _bipartites([(8, 8), (3, 3)])  # eigenvalues -8 ... 3.616
        val = None
    n = row.size - 1
    a = _dense(row, col, val)
    k = a.sum(1)
    bmat = a - np.outer(k, k) / k.sum()
    w = np.linalg.eigvalsh(bmat)
    lab, q, lam, eig, nr = O.modularity(row, col, val, stop=1, tol=1e-12)
    assert abs(lam - w[-1]) <= 1e-6 * abs(w[-1])
    # the returned vector is an eigenvector: small residual of B v - lambda v
    assert np.linalg.norm(bmat @ eig - lam * eig) <= 1e-3 * abs(lam)
    if graph == "bipartite":
        assert w[0] < -abs(w[-1])  # the shifted pass was needed
    if graph == "cliques":
        lab0, q0, *_ = O.modularity(row, col)
        np.testing.assert_array_equal(lab, lab0)
        assert q == q0


def test_modularity_depth_limit_coarsens():
    # four cliques: one bisection level gives two groups of two cliques each;
    # a depth-limited partition is always a union of the unlimited one's
    # communities, and its labels run 1..K in round order
    row, col, _ = cliques([20, 20, 20, 20])
    full, q_full, *_ = O.modularity(row, col)
    assert len(set(full.tolist())) == 4
    for d, k in [(1, 2), (2, 4), (5, 4)]:
        lab, q, *_ = O.modularity(row, col, max_depth=d)
        assert sorted(set(lab.tolist())) == list(range(1, k + 1))
        for c in set(full.tolist()):
            assert len(set(lab[full == c].tolist())) == 1
        assert q == O.modularity_q(row, col, lab)
    assert O.modularity(row, col, max_depth=1)[1] < q_full


def _dense_csr_numpy(keep_upper, n):
    iu = np.triu_indices(n, 1)
    m = np.zeros((n, n), bool)
    m[iu[0][keep_upper], iu[1][keep_upper]] = True
    m |= m.T
    row = np.concatenate([[0], np.cumsum(m.sum(1))]).astype(np.int32)
    col = np.nonzero(m)[1].astype(np.int32)
    return row, col


def test_cormat_average_matches_numpy():
    from pagani_amd import synth
    n = 90
    subs = [synth.bold(40 + s, n, t, 5) for s, t in enumerate((60, 75, 50))]
    iu = np.triu_indices(n, 1)
    ccs = [np.corrcoef(x.T.astype(np.float64))[iu] for x in subs]
    # one subject: (float) r
    a1 = O.cormat_average(subs[:1])
    np.testing.assert_allclose(a1, ccs[0].astype(np.float32), rtol=0, atol=6e-8)
    # three subjects: float32 running sum of float32 r, then / 3 in float32
    acc = np.zeros(iu[0].size, np.float32)
    for cc in ccs:
        acc += cc.astype(np.float32)
    acc /= np.float32(3)
    np.testing.assert_allclose(O.cormat_average(subs), acc, rtol=0, atol=2e-7)
    # Fisher: tanh(mean(arctanh r))
    want = np.tanh(np.mean([np.arctanh(cc) for cc in ccs], 0))
    np.testing.assert_allclose(O.cormat_average(subs, True), want, rtol=0, atol=2e-6)


def test_dense_thresholds_brute_force():
    rng = np.random.default_rng(3)
    n = 70
    m1 = n * (n - 1) // 2
    a = np.round(rng.uniform(-1, 1, m1), 2).astype(np.float32)  # many exact ties
    row, col, val = O.dense_csr_r(a, n, 0.5, True)
    r2, c2 = _dense_csr_numpy(a > np.float32(0.5), n)
    np.testing.assert_array_equal(row, r2)
    np.testing.assert_array_equal(col, c2)
    for s in (1.0, 7.3, 20.0):
        k = O.sparsity_k(n, s)
        row, col, val, t = O.dense_csr_s(a, n, s, True)
        order = np.argsort(-a, kind="stable")  # descending, ties in row-major order
        keep = np.zeros(m1, bool)
        keep[order[:k]] = True
        assert t == a[order[k - 1]]
        r2, c2 = _dense_csr_numpy(keep, n)
        np.testing.assert_array_equal(row, r2)
        np.testing.assert_array_equal(col, c2)
        # weights are the matrix values of the kept pairs
        full = np.zeros((n, n), np.float32)
        full[np.triu_indices(n, 1)] = a
        full += full.T
        np.testing.assert_array_equal(val, full[np.repeat(np.arange(n), np.diff(row)), col])


# ------------------------------------------------------------------ rewiring

@pytest.mark.parametrize("weighted", [False, True])
def test_rewire_invariants(weighted):
    row, col, val = random_graph(300, 0.05, 3, weighted=weighted)
    v = val if weighted else None
    oc, ov, sw = O.rewire(row, col, v, seed=7)
    check_rewired(row, col, v, oc, ov)
    m = col.size // 2
    assert 0.5 * m < sw <= 2 * m  # sparse: most of the 2M attempts succeed
    assert not np.array_equal(oc, col)
    oc2, ov2, sw2 = O.rewire(row, col, v, seed=7)
    assert sw2 == sw and np.array_equal(oc2, oc)


def test_rewire_no_swap_possible():
    # every pair of star edges shares the centre; in a complete graph every new
    # edge exists already
    for row, col, _ in (star_graph(9), complete_graph(7)):
        oc, _, sw = O.rewire(row, col, None, seed=1)
        assert sw == 0
        np.testing.assert_array_equal(oc, col)


def test_rewire_two_edges_reach_every_matching():
    # {0-1, 2-3}: each swap moves to one of the two other perfect matchings,
    # the coin choosing which; all three must turn up
    row, col, _ = csr_from_edges(4, [(0, 1), (2, 3)])
    seen = {}
    for s in range(300):
        oc, _, _ = O.rewire(row, col, None, seed=s)
        seen[tuple(oc.tolist())] = seen.get(tuple(oc.tolist()), 0) + 1
    assert len(seen) == 3 and min(seen.values()) > 50, seen


def test_rewire_clustering_matches_networkx():
    nx = pytest.importorskip("networkx")
    n, k = 200, 10
    row, col, _ = ring_lattice(n, k)
    assert abs(O.cp_binary(row, col)[1] - 2 / 3) < 1e-7  # 3(k-2) / (4(k-1)), stored float
    ours, theirs = [], []
    for s in range(8):
        oc, _, sw = O.rewire(row, col, None, seed=s)
        ours.append(O.cp_binary(row, oc)[1])
        g = nx.watts_strogatz_graph(n, k, 0.0)
        nx.double_edge_swap(g, nswap=sw, max_tries=100 * sw, seed=s)
        theirs.append(nx.average_clustering(g))
    assert np.mean(ours) < 0.15
    assert abs(np.mean(ours) - np.mean(theirs)) < 0.015, (np.mean(ours), np.mean(theirs))
