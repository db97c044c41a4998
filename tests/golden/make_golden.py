"""Writes tests/golden/pervertex.json: small seeded random graphs with their
per-vertex triangle counts from networkx.triangles (a textbook routine, not
the reference).  Run once: python tests/golden/make_golden.py
"""
import json
import os

import networkx as nx
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20260101)
    cases = []
    for i, (n, p) in enumerate([(12, 0.5), (40, 0.2), (64, 0.15), (100, 0.08), (150, 0.1), (200, 0.05)]):
        g = nx.gnp_random_graph(n, p, seed=int(rng.integers(1 << 30)))
        edges = sorted((min(a, b), max(a, b)) for a, b in g.edges())
        tri = nx.triangles(g)
        cases.append({
            "name": f"gnp_{n}_{p}",
            "n": n,
            "edges": edges,
            "per_vertex": [tri[v] for v in range(n)],
            "total": sum(tri.values()) // 3,
        })
    # a power-law graph with hubs
    g = nx.barabasi_albert_graph(300, 6, seed=7)
    tri = nx.triangles(g)
    cases.append({"name": "ba_300_6", "n": 300, "edges": sorted((min(a, b), max(a, b)) for a, b in g.edges()),
                  "per_vertex": [tri[v] for v in range(300)], "total": sum(tri.values()) // 3})
    with open(os.path.join(HERE, "pervertex.json"), "w") as f:
        json.dump({"generator": "networkx " + nx.__version__, "cases": cases}, f)


if __name__ == "__main__":
    main()
