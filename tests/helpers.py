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
