"""Write golden vectors for the oracle from networkx 3.4.2 and numpy FP64.

Run from the repository root:  python tests/golden/make_golden.py
Outputs tests/golden/*.npz (inputs and expected outputs only).
"""
import os
import sys

import networkx as nx
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "pagani-toolkit_amd"))
from pagani_amd import synth  # noqa: E402


def to_csr(g, n, weight=None):
    row = [0]
    col, val = [], []
    for i in range(n):
        nb = sorted(g.adj[i])
        col += nb
        if weight:
            val += [g.adj[i][j][weight] for j in nb]
        row.append(len(col))
    return (np.array(row, np.int32), np.array(col, np.int32),
            np.array(val, np.float32) if weight else np.zeros(0, np.float32))


def graph_fixture(name, g, weighted):
    n = g.number_of_nodes()
    row, col, val = to_csr(g, n, "weight" if weighted else None)
    out = dict(row=row, col=col, val=val)
    bc = nx.betweenness_centrality(g, normalized=False)
    out["bc_ordered_pairs"] = np.array([2.0 * bc[i] for i in range(n)])
    tri = nx.triangles(g)
    out["triangles"] = np.array([tri[i] for i in range(n)], np.int64)
    cl = nx.clustering(g)
    out["clustering"] = np.array([cl[i] for i in range(n)])
    src = np.array(sorted(set([0, n // 3, n - 1])), np.int32)
    lev = np.full((src.size, n), -1, np.int32)
    for k, s in enumerate(src):
        for v, dd in nx.single_source_shortest_path_length(g, int(s)).items():
            lev[k, v] = dd
    out["bfs_sources"], out["bfs_levels"] = src, lev
    # binary nodal efficiency in FP64 from networkx hop counts, and the global
    # efficiency (mean of 1/d over ordered pairs, 0 for unreachable) = 1 / Lp
    eff = np.zeros(n)
    for i, dist in nx.all_pairs_shortest_path_length(g):
        eff[i] = sum(1.0 / dd for j, dd in dist.items() if j != i) / (n - 1)
    out["nodal_eff"] = eff
    out["global_eff"] = np.float64(nx.global_efficiency(g))
    # leading eigenvector, unit 2-norm, positive (binary adjacency; the
    # weighted graph also gets the weighted one); connected graphs only
    if nx.is_connected(g):
        ec = nx.eigenvector_centrality_numpy(g)
        out["ec"] = np.array([ec[i] for i in range(n)])
    if weighted and nx.is_connected(g):
        ecw = nx.eigenvector_centrality_numpy(g, weight="weight")
        out["ec_weighted"] = np.array([ecw[i] for i in range(n)])
    if weighted:
        wcl = nx.clustering(g, weight="weight")
        out["clustering_onnela"] = np.array([wcl[i] for i in range(n)])
        h = nx.Graph()
        h.add_nodes_from(range(n))
        for a, b, d in g.edges(data=True):
            h.add_edge(a, b, dist=1.0 / np.float64(np.float32(d["weight"])))
        out["fw_inv_weight"] = nx.floyd_warshall_numpy(h, nodelist=list(range(n)), weight="dist")
    comms = nx.community.greedy_modularity_communities(g, weight="weight" if weighted else None)
    lab = np.zeros(n, np.int32)
    for c, members in enumerate(comms):
        for v in members:
            lab[v] = c + 1
    out["partition"] = lab
    out["partition_q"] = np.float64(nx.community.modularity(g, comms, weight="weight" if weighted else None))
    np.savez_compressed(os.path.join(HERE, f"graph_{name}.npz"), **out)


def main():
    rng = np.random.default_rng(20261202)
    g = nx.gnp_random_graph(180, 0.06, seed=11)
    graph_fixture("gnp180", g, False)
    g = nx.gnp_random_graph(150, 0.08, seed=12)
    for a, b in g.edges():
        g[a][b]["weight"] = float(np.float32(rng.uniform(0.05, 1.0)))
    graph_fixture("gnp150_weighted", g, True)
    g = nx.disjoint_union_all([nx.complete_graph(12), nx.cycle_graph(9), nx.path_graph(7), nx.star_graph(8)])
    g.add_node(g.number_of_nodes())  # an isolated vertex
    graph_fixture("mixed_components", g, False)
    g = nx.connected_caveman_graph(5, 8)
    graph_fixture("caveman5x8", g, False)

    # correlation thresholds: numpy FP64 Pearson on seeded synthetic series
    x = synth.bold(77, 240, 90, 6)
    cc = np.corrcoef(x.T.astype(np.float64))
    iu = np.triu_indices(240, 1)
    np.savez_compressed(os.path.join(HERE, "corr_n240_t90.npz"), bold=x, r_upper=cc[iu])


if __name__ == "__main__":
    main()
