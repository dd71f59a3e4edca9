"""Writes the known answers for data_test.gr (tests/golden/data_test_answers.json).

data_test.gr is the reference's only shipped input (datasets/data_test.gr, a
20-vertex excerpt of USA-road-d.BAY; copied here as a fixture).  The reference
ships no expected outputs, so the answers are computed with library routines
that share no code with the oracle or the product:

  SSSP distances  scipy.sparse.csgraph.dijkstra (and networkx Dijkstra)
  BFS levels      scipy.sparse.csgraph.shortest_path(unweighted=True)
  colours         networkx topological longest chain in the precedence DAG:
                  colour(v) = 1 + length of the longest chain of
                  predecessors u -> ... -> v where each step is an arc of
                  row(v) to a smaller id (the reference's rule,
                  apps/graph-color/src/graph_color_kernel.cu:28-49)
  out-degree      counted from the 'a' lines

The file is parsed here with plain Python (the 'p sp N M', 's S', 'a u v w'
lines of the DIMACS9 format), not with the product's reader.
Run: python tests/golden/make_golden.py
"""
import json
import os

import networkx as nx
import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra, shortest_path

HERE = os.path.dirname(os.path.abspath(__file__))


def parse(path):
    n = src = None
    arcs = []
    for line in open(path):
        if line.startswith("p"):
            _, _, n, m = line.split()[:4]
            n = int(n)
        elif line.startswith("s"):
            src = int(line.split()[1]) - 1
        elif line.startswith("a"):
            _, u, v, w = line.split()[:4]
            arcs.append((int(u) - 1, int(v) - 1, int(w)))
    return n, src, arcs


def main():
    path = os.path.join(HERE, "data_test.gr")
    n, src, arcs = parse(path)
    # min weight per (u, v) for scipy (it would sum duplicates)
    best = {}
    for u, v, w in arcs:
        best[(u, v)] = min(w, best.get((u, v), w))
    rows = [u for (u, v) in best]
    cols = [v for (u, v) in best]
    data = [best[k] for k in best]
    A = sp.csr_matrix((data, (rows, cols)), shape=(n, n))
    d = dijkstra(A, directed=True, indices=src)
    G = nx.DiGraph()
    G.add_nodes_from(range(n))
    for (u, v), w in best.items():
        G.add_edge(u, v, weight=w)
    dn = nx.single_source_dijkstra_path_length(G, src)
    INF = 1073741824
    dist = [int(x) if np.isfinite(x) else INF for x in d]
    assert all(dist[v] == dn.get(v, INF) for v in range(n))
    lv = shortest_path(A, directed=True, unweighted=True, indices=src)
    level = [int(x) if np.isfinite(x) else 0x7FFFFFFF for x in lv]
    # precedence DAG: u -> v when u in row(v) and u < v
    D = nx.DiGraph()
    D.add_nodes_from(range(n))
    for u, v, _ in arcs:  # arc u->v means v in row(u)
        if v < u:
            D.add_edge(v, u)
    color = [0] * n
    for v in nx.topological_sort(D):
        color[v] = 1 + max((color[p] for p in D.predecessors(v)), default=0)
    outdeg = [0] * n
    for u, _, _ in arcs:
        outdeg[u] += 1
    out = {"n": n, "m": len(arcs), "source": src, "dist": dist, "level": level, "color": color,
           "outdeg": outdeg}
    with open(os.path.join(HERE, "data_test_answers.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
