"""Shared test utilities (graphs, CSR helpers)."""
import numpy as np


def csr_from_edges(n, edges, weights=None):
    """Symmetric CSR (sorted columns) from undirected edge pairs."""
    adj = [dict() for _ in range(n)]
    for k, (a, b) in enumerate(edges):
        if a == b:
            continue
        w = 1.0 if weights is None else float(weights[k])
        adj[a][b] = w
        adj[b][a] = w
    row = np.zeros(n + 1, np.int32)
    cols, vals = [], []
    for i in range(n):
        ks = sorted(adj[i])
        cols += ks
        vals += [adj[i][k] for k in ks]
        row[i + 1] = len(cols)
    return row, np.array(cols, np.int32), np.array(vals, np.float32)


def random_graph(n, p, seed, weighted=False):
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    m = rng.random(iu[0].size) < p
    edges = list(zip(iu[0][m].tolist(), iu[1][m].tolist()))
    w = rng.uniform(0.05, 1.0, len(edges)) if weighted else None
    return csr_from_edges(n, edges, w)


def path_graph(n):
    return csr_from_edges(n, [(i, i + 1) for i in range(n - 1)])


def star_graph(n):
    return csr_from_edges(n, [(0, i) for i in range(1, n)])


def complete_graph(n):
    return csr_from_edges(n, [(i, j) for i in range(n) for j in range(i + 1, n)])


def cliques(sizes, bridge=True):
    edges, base = [], 0
    starts = []
    for s in sizes:
        starts.append(base)
        edges += [(base + i, base + j) for i in range(s) for j in range(i + 1, s)]
        base += s
    if bridge:
        for a, b in zip(starts[:-1], starts[1:]):
            edges.append((a, b))
    return csr_from_edges(base, edges)


def ring_lattice(n, k):
    """Each vertex joined to its k/2 nearest neighbours on either side."""
    return csr_from_edges(n, [(i, (i + d) % n) for i in range(n) for d in range(1, k // 2 + 1)])


def check_rewired(row, col_in, val_in, col, val):
    """Invariants of a degree-preserving rewiring under the same row offsets:
    columns in range, strictly ascending per row, no self-loops, a symmetric
    edge set, each edge's weight equal in both directions and the multiset of
    edge weights unchanged."""
    n = row.size - 1
    assert col.shape == col_in.shape
    r = np.repeat(np.arange(n, dtype=np.int64), np.diff(row))
    c = col.astype(np.int64)
    assert ((c >= 0) & (c < n)).all()
    assert (c != r).all(), "self-loop"
    same = r[1:] == r[:-1]
    assert (np.diff(c)[same] > 0).all(), "row not strictly ascending"
    fwd, bwd = r * n + c, c * n + r
    o1, o2 = np.argsort(fwd), np.argsort(bwd)
    assert np.array_equal(fwd[o1], bwd[o2]), "edge set not symmetric"
    if val is not None:
        assert np.array_equal(val[o1].view(np.uint32), val[o2].view(np.uint32)), "weights not symmetric"
        up_in = col_in.astype(np.int64) > r
        assert np.array_equal(np.sort(val[c > r]), np.sort(val_in[up_in])), "weights changed"
