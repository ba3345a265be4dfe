"""Pin the FP64 oracle on golden vectors written by tests/golden/make_golden.py
(networkx 3.4.2 and numpy FP64, generated in this repository)."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GRAPHS = sorted(glob.glob(os.path.join(GOLD, "graph_*.npz")))


def load(p):
    with np.load(p, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("path", GRAPHS, ids=[os.path.basename(p) for p in GRAPHS])
def test_graph_metrics_match_networkx(path):
    g = load(path)
    row, col = g["row"], g["col"]
    np.testing.assert_allclose(O.bc(row, col), g["bc_ordered_pairs"], rtol=1e-12, atol=1e-9)
    np.testing.assert_array_equal(O.triangles(row, col), g["triangles"])
    cp, _ = O.cp_binary(row, col)
    np.testing.assert_allclose(cp, g["clustering"], rtol=0, atol=1e-7)
    for k, s in enumerate(g["bfs_sources"]):
        np.testing.assert_array_equal(O.bfs(row, col, int(s)), g["bfs_levels"][k])
    weighted = g["val"].size > 0
    val = g["val"] if weighted else None
    q = O.modularity_q(row, col, g["partition"], val, 0)
    assert abs(q - float(g["partition_q"])) < 1e-12


@pytest.mark.parametrize("path", GRAPHS, ids=[os.path.basename(p) for p in GRAPHS])
def test_bfs_efficiency_matches_networkx(path):
    g = load(path)
    row, col = g["row"], g["col"]
    n = row.size - 1
    eff, lp, iso, unc = O.bfs_efficiency(row, col)
    # FP32 sums of float reciprocals against FP64 networkx values
    np.testing.assert_allclose(eff, g["nodal_eff"], rtol=1e-6, atol=1e-12)
    assert abs(1.0 / lp - float(g["global_eff"])) <= 1e-6 * float(g["global_eff"])
    assert iso == int((g["nodal_eff"] == 0).sum())
    assert unc == int(np.any(g["bfs_levels"] < 0))
    # the FP32 nodal sum is sequential over j != i in increasing j:
    # numpy float32 add.accumulate of float32 reciprocals of the hop counts
    for i in range(0, n, max(1, n // 25)):
        lev = O.bfs(row, col, i)
        m = (lev > 0)
        r = np.float32(1.0) / lev[m].astype(np.float32)
        s = np.add.accumulate(r, dtype=np.float32)[-1] if r.size else np.float32(0)
        assert (s / np.float32(n - 1)).view(np.uint32) == eff[i].view(np.uint32), i


@pytest.mark.parametrize("path", GRAPHS, ids=[os.path.basename(p) for p in GRAPHS])
def test_ec_matches_networkx(path):
    g = load(path)
    if "ec" not in g:
        pytest.skip("disconnected graph: no networkx reference vector")
    row, col = g["row"], g["col"]
    vh, av, lam, it = O.ec(row, col)
    assert it < 10000
    # the stop rule |A x_k - A x_(k-1)| <= 1e-6 leaves the vector within ~1e-6 / gap
    np.testing.assert_allclose(vh, g["ec"], atol=2e-6)
    np.testing.assert_allclose(av / lam, g["ec"], atol=2e-6)
    if "ec_weighted" in g:
        vh, av, lam, it = O.ec(row, col, g["val"])
        np.testing.assert_allclose(vh, g["ec_weighted"], atol=2e-6)


def test_weighted_clustering_and_fw_match_networkx():
    g = load(os.path.join(GOLD, "graph_gnp150_weighted.npz"))
    row, col, val = g["row"], g["col"], g["val"]
    cp, _ = O.cp_weighted(row, col, val, 2, 0)
    np.testing.assert_allclose(cp, g["clustering_onnela"], rtol=1e-12, atol=1e-15)
    d = O.fw_blocked(O.fw_init(row, col, val, 0), 32)
    ref = g["fw_inv_weight"]
    np.testing.assert_array_equal(np.isinf(d), np.isinf(ref))
    fin = np.isfinite(ref)
    np.testing.assert_allclose(d[fin], ref[fin], rtol=2e-6)


def test_barrat_brute_force():
    g = load(os.path.join(GOLD, "graph_gnp150_weighted.npz"))
    row, col, val = g["row"], g["col"], g["val"]
    n = row.size - 1
    w = np.zeros((n, n))
    for i in range(n):
        w[i, col[row[i]:row[i + 1]]] = val[row[i]:row[i + 1]]
    a = w > 0
    cp, _ = O.cp_weighted(row, col, val, 1, 0)
    for k in range(0, n, 7):
        nb = np.nonzero(a[k])[0]
        if nb.size < 2:
            assert cp[k] == 0
            continue
        s = sum(w[k, i] + w[k, j] for x, i in enumerate(nb) for j in nb[x + 1:] if a[i, j])
        assert abs(cp[k] - s / (nb.size - 1) / w[k].sum()) < 1e-12


def test_correlation_matches_numpy_corrcoef():
    g = load(os.path.join(GOLD, "corr_n240_t90.npz"))
    z64, _ = O.zscore(g["bold"])
    iu = np.triu_indices(240, 1)
    mine = np.array([O.r_pair(z64, int(i), int(j)) for i, j in zip(*iu)])
    np.testing.assert_allclose(mine, g["r_upper"], rtol=0, atol=1e-13)
    for thr in (0.2, 0.35):
        row, col, _ = O.corr_csr_r(z64, thr)
        want = (g["r_upper"] > thr)
        assert row[-1] == 2 * int(want.sum())
