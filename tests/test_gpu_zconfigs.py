"""GPU parity at the workloads' own sizes, against the FP64 oracle.

C1  plain method: N=5,000, T=200, K=10, r > 0.4 (both precisions), degree,
    clustering and exact BC from every source.
C2  correlation precision: N=58,000, T=240, K=90, 1% sparsity; the edge set,
    row offsets and t* of both the FP32 and the BF16x3 Gram equal the FP64
    oracle's, in one Gram pass, with every rechecked pair inside the proven
    error bound.
C3  metrics on the C2 graph: triangles and binary Cp (bit-exact, all
    vertices), BC from a seeded 1,024-source sample; blocked Floyd-Warshall
    on the w = 1 - r subgraph of the first 2,048 voxels against the oracle's
    blocked FW (bit-exact, tiles 64 and 128) and on the first 8,192 voxels with
    unit weights against BFS hop counts.
C5  weighted network N=120,000, T=300, K=200, w = r for r > 0.3: CSR
    invariants over all entries, 64 seeded rows (columns and weights) against
    rows the oracle forms alone, Onnela clustering at 256 seeded vertices, and
    leading-eigenvector modularity with the Rayleigh-quotient stop for the
    first three bisection levels against the oracle to the same depth (labels,
    Q, lambda0, the round-0 vector), Q recomputed by the oracle from the GPU
    labels and the round-0 eigenpair residual |B v - lambda0 v| / |lambda0|
    from one FP64 host product.
"""
import numpy as np
import pytest

from conftest import say
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BC_RTOL = 1e-9


def _workload(name):
    from pagani_amd import synth
    w = synth.WORKLOADS[name]
    return w, synth.bold(w["seed"], w["n"], w["t"], w["k"])


def _subgraph(row, col, val, n):
    """Induced subgraph on vertices 0..n-1 of a CSR."""
    r, c, v = [0], [], []
    for i in range(n):
        a, b = row[i], row[i + 1]
        m = col[a:b] < n
        c.append(col[a:b][m])
        if val is not None:
            v.append(val[a:b][m])
        r.append(r[-1] + int(m.sum()))
    return (np.array(r, np.int32), np.concatenate(c).astype(np.int32),
            None if val is None else np.concatenate(v).astype(np.float32))


def check_csr(row, col, val=None):
    """Structural invariants: offsets from 0 to nnz, columns in range and
    strictly ascending per row, no self loops, a symmetric edge set and (when
    weighted) the same weight in both directions."""
    n = row.size - 1
    assert row[0] == 0 and row[-1] == col.size
    assert (np.diff(row) >= 0).all()
    r = np.repeat(np.arange(n, dtype=np.int64), np.diff(row))
    c = col.astype(np.int64)
    assert ((c >= 0) & (c < n)).all()
    assert (c != r).all(), "self loop"
    same = r[1:] == r[:-1]
    assert (np.diff(c)[same] > 0).all(), "row not strictly ascending"
    fwd = r * n + c
    bwd = c * n + r
    o2 = np.argsort(bwd, kind="stable")
    # fwd is already sorted (rows ascending, columns ascending in a row)
    assert np.array_equal(fwd, bwd[o2]), "edge set not symmetric"
    if val is not None:
        assert np.array_equal(val.view(np.uint32), val[o2].view(np.uint32)), "weights not symmetric"


# ---------------------------------------------------------------------- C1

@pytest.fixture(scope="module")
def c1(ctx):
    w, x = _workload("plain_n5000_t200")
    z64, _ = O.zscore(x)
    rows = O.corr_rows_r(z64, 0.4, np.arange(w["n"], dtype=np.int32))
    orow = np.concatenate([[0], np.cumsum([c.size for c, _ in rows])]).astype(np.int32)
    ocol = np.concatenate([c for c, _ in rows]).astype(np.int32)
    return w, x, orow, ocol


@pytest.mark.parametrize("prec", [0, 1])
def test_c1_plain_edge_set_and_metrics(ctx, c1, prec):
    w, x, orow, ocol = c1
    n, t = w["n"], w["t"]
    say(f"C1 n={n} t={t} r>0.4 prec={prec}")
    info = ctx.corr_build(x, n, t, 0.4, prec=prec)
    row, col, _ = ctx.corr_csr(n, info.nnz)
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    assert info.band_max_err <= info.band / 1.5
    deg, deg_ref = ctx.degree(row)
    odeg, odeg_ref = O.degree(orow)
    np.testing.assert_array_equal(deg, odeg)
    np.testing.assert_array_equal(deg_ref.view(np.uint32), odeg_ref.view(np.uint32))
    cp, _ = ctx.cp_binary(row, col)
    ocp, _ = O.cp_binary(orow, ocol)
    np.testing.assert_array_equal(cp.view(np.uint32), ocp.view(np.uint32))


def test_c1_bc_all_sources(ctx, c1):
    w, x, orow, ocol = c1
    say(f"C1 bc all {w['n']} sources, nnz={ocol.size}")
    b = ctx.bc(orow, ocol)
    ob = O.bc(orow, ocol)
    np.testing.assert_allclose(b, ob, rtol=BC_RTOL, atol=1e-9)
    rel = np.max(np.abs(b - ob) / np.maximum(np.abs(ob), 1e-300))
    print(f"C1 BC all sources: max rel err {rel:.2e}")
    bf = ctx.bc_float(orow, ocol)
    np.testing.assert_array_equal(bf.view(np.uint32), b.astype(np.float32).view(np.uint32))


# ---------------------------------------------------------------------- C2

@pytest.fixture(scope="module")
def c2():
    w, x = _workload("corr_n58000_t240")
    z64, _ = O.zscore(x)
    print(f"C2 oracle: n={w['n']} t={w['t']} 1% sparsity", flush=True)
    orow, ocol, _, tstar = O.corr_csr_s(z64, 1.0)
    return w, x, orow, ocol, tstar


@pytest.mark.parametrize("prec", [0, 1])
def test_c2_edge_set_bit_exact(ctx, c2, prec):
    w, x, orow, ocol, tstar = c2
    n, t = w["n"], w["t"]
    say(f"C2 n={n} t={t} 1% sparsity prec={prec}")
    info = ctx.corr_build(x, n, t, 1.0, sparsity=True, prec=prec)
    row, col, _ = ctx.corr_csr(n, info.nnz)
    print(f"C2 prec={prec}: nnz={info.nnz} t*={info.t_star!r} band={info.band:.3e} "
          f"rechecked={info.band_pairs} flips={info.band_flips} max|r~-r64|={info.band_max_err:.3e} "
          f"gram passes={info.gram_passes}")
    assert info.nnz == ocol.size == 2 * O.sparsity_k(n, 1.0)
    assert info.t_star == tstar
    np.testing.assert_array_equal(row, orow)
    np.testing.assert_array_equal(col, ocol)
    assert info.gram_passes == 1
    # the rounding model behind the band holds with its safety factor to spare
    assert info.band_max_err <= info.band / 1.5


# ---------------------------------------------------------------------- C3

def test_c3_triangles_cp_bc_on_c2_graph(ctx, c2):
    from pagani_amd import synth
    w, x, orow, ocol, _ = c2
    n = w["n"]
    say(f"C3 triangles/cp n={n} nnz={ocol.size}")
    tri = ctx.triangles(orow, ocol)
    np.testing.assert_array_equal(tri, O.triangles(orow, ocol))
    cp, _ = ctx.cp_binary(orow, ocol)
    ocp, _ = O.cp_binary(orow, ocol)
    np.testing.assert_array_equal(cp.view(np.uint32), ocp.view(np.uint32))
    src = synth.sample_sources(n, 1024, 4242)
    say("C3 bc 1024 sources")
    b = ctx.bc(orow, ocol, src)
    ob = O.bc(orow, ocol, src)
    np.testing.assert_allclose(b, ob, rtol=BC_RTOL, atol=1e-9)
    print(f"C3 BC 1024 sources: max rel err {np.max(np.abs(b - ob) / np.maximum(np.abs(ob), 1e-300)):.2e}")


@pytest.fixture(scope="module")
def c3_weighted(ctx, c2):
    """w = 1 - r network of the first 8,192 C2 voxels, edges where r is above
    the C2 1% threshold."""
    w, x, _, _, tstar = c2
    m = 8192
    xs = np.ascontiguousarray(x[:, :m])
    info = ctx.corr_build(xs, m, w["t"], float(np.float32(tstar)), weighted=True)
    return ctx.corr_csr(m, info.nnz, True)


@pytest.mark.parametrize("block", [64, 128])
def test_c3_fw_n2048_bit_exact(ctx, c3_weighted, block):
    row, col, val = _subgraph(*c3_weighted, 2048)
    say(f"C3 fw n=2048 block={block} nnz={col.size}")
    d = ctx.fw_init(row, col, val, 1)
    np.testing.assert_array_equal(d.view(np.uint32), O.fw_init(row, col, val, 1).view(np.uint32))
    got = ctx.fw(d.copy(), block)
    want = O.fw_blocked(d, block)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_c3_fw_n8192_unit_weights_equal_bfs(ctx, c3_weighted):
    row, col, val = c3_weighted
    n = row.size - 1
    say(f"C3 fw n={n} unit weights vs bfs")
    d = ctx.fw(ctx.fw_init(row, col, val, 2), 64)
    lev = ctx.bfs(row, col, np.arange(n, dtype=np.int32))
    hops = np.where(lev >= 0, lev.astype(np.float32), np.float32(np.inf))
    np.testing.assert_array_equal(d, hops)
    for s in (0, 1, 4097, n - 1):
        np.testing.assert_array_equal(lev[s], O.bfs(row, col, s))


# ---------------------------------------------------------------------- C5

@pytest.fixture(scope="module")
def c5(ctx):
    w, x = _workload("weighted_n120000_t300")
    n, t = w["n"], w["t"]
    say(f"C5 n={n} t={t} weighted r>0.3")
    info = ctx.corr_build(x, n, t, 0.3, weighted=True)
    row, col, val = ctx.corr_csr(n, info.nnz, True)
    return w, x, info, row, col, val


def test_c5_weighted_construction(ctx, c5):
    w, x, info, row, col, val = c5
    print(f"C5: nnz={info.nnz} rechecked={info.band_pairs} flips={info.band_flips} "
          f"max|r~-r64|={info.band_max_err:.3e} band={info.band:.3e}")
    assert info.band_max_err <= info.band / 1.5
    check_csr(row, col, val)
    z64, _ = O.zscore(x)
    rng = np.random.default_rng(55)
    rows = np.sort(rng.choice(w["n"], 64, replace=False)).astype(np.int32)
    want = O.corr_rows_r(z64, 0.3, rows, weighted=True)
    for q, i in enumerate(rows):
        a, b = row[i], row[i + 1]
        np.testing.assert_array_equal(col[a:b], want[q][0])
        np.testing.assert_array_equal(val[a:b].view(np.uint32), want[q][1].view(np.uint32))


def test_c5_onnela_clustering_sampled(ctx, c5):
    w, _, _, row, col, val = c5
    say("C5 onnela clustering")
    cp, _ = ctx.cp_weighted(row, col, val, 2, 0)
    verts = np.sort(np.random.default_rng(56).choice(w["n"], 256, replace=False)).astype(np.int32)
    np.testing.assert_allclose(cp[verts], O.cp_weighted_at(row, col, val, verts, 2, 0), rtol=1e-9, atol=1e-12)


def test_c5_modularity_rayleigh(ctx, c5):
    """The first three bisection levels (the full network, then its halves) with
    the Rayleigh-quotient stop, against the FP64 oracle run to the same depth."""
    import pagani_amd
    from test_gpu_metrics import mod_errors
    w, _, _, row, col, val = c5
    n = w["n"]
    say("C5 modularity, Rayleigh-quotient stop, 3 levels")
    got = ctx.modularity(row, col, val, pagani_amd.MOD_STRENGTH, stop=pagani_amd.MOD_STOP_RAYLEIGH,
                         progress=True, max_depth=3)
    lab, q, lam, eig, rounds = got
    say("C5 modularity oracle, 3 levels")
    want = O.modularity(row, col, val, 0, stop=1, max_depth=3)
    err = mod_errors(got, want)
    oq = O.modularity_q(row, col, lab, val, 0)
    # round-0 eigenpair residual, B v = A v - k (k . v) / 2m in FP64 on the host
    r64 = np.repeat(np.arange(n), np.diff(row))
    k = np.bincount(r64, weights=val.astype(np.float64), minlength=n)
    av = np.bincount(r64, weights=val.astype(np.float64) * eig[col], minlength=n)
    bv = av - k * (k @ eig) / k.sum()
    res = np.linalg.norm(bv - lam * eig) / abs(lam) / np.linalg.norm(eig)
    print(f"C5 modularity 3 levels: Q={q:.6f} (oracle {want[1]:.6f}; oracle Q of the GPU labels {oq:.6f}), "
          f"{len(set(lab.tolist()))} modules, {rounds} rounds, lambda0={lam:.6f} (rel err {err['lam']:.2e}), "
          f"eigenvector rel err {err['eig']:.2e}, label agreement {err['agree']:.5f}, "
          f"eigenpair residual {res:.2e}")
    assert abs(q - oq) <= 1e-9 * max(1.0, abs(oq))
    assert rounds == want[4]
    assert err["lam"] <= 1e-7 and err["q"] <= 1e-6 and err["agree"] > 0.99, err
    assert res <= 1e-3
