"""CPU oracle (TEST INFRASTRUCTURE ONLY): reverse Cuthill-McKee ordering of
the cell adjacency graph, restated from the published algorithm (Cuthill &
McKee 1969; George & Liu 1981, "reverse" variant):

* nodes are visited component by component; each component starts from its
  lowest-degree unvisited node (ties: lowest index);
* breadth-first search; the unvisited neighbours of each node are appended in
  increasing degree (ties: lowest index);
* the final order is reversed.

The product's C++ ordering (eulersr_amd/csrc/esr_mesh.cpp) must equal this one
element for element. The reference has no renumbering (the new layout is ours).
"""
from __future__ import annotations

import numpy as np


def cell_graph(neighbors):
    """Adjacency lists (sorted) from a (N,3) neighbour table, -1 = none."""
    n = len(neighbors)
    adj = [sorted({int(j) for j in neighbors[i] if j >= 0 and j != i}) for i in range(n)]
    return adj


def rcm(neighbors):
    adj = cell_graph(neighbors)
    n = len(adj)
    deg = np.array([len(a) for a in adj])
    seeds = sorted(range(n), key=lambda i: (deg[i], i))
    visited = np.zeros(n, dtype=bool)
    order = []
    for s in seeds:
        if visited[s]:
            continue
        visited[s] = True
        order.append(s)
        head = len(order) - 1
        while head < len(order):
            i = order[head]
            head += 1
            nb = [j for j in adj[i] if not visited[j]]
            nb.sort(key=lambda j: (deg[j], j))
            for j in nb:
                visited[j] = True
                order.append(j)
    return np.array(order[::-1], dtype=np.int32)


def bandwidth(neighbors, perm):
    """max |pos(i) - pos(j)| over adjacent cells under ordering ``perm``
    (perm[new] = old)."""
    pos = np.empty(len(perm), dtype=np.int64)
    pos[perm] = np.arange(len(perm))
    nb = np.asarray(neighbors)
    i = np.repeat(np.arange(len(nb)), 3)
    j = nb.reshape(-1)
    ok = j >= 0
    return int(np.max(np.abs(pos[i[ok]] - pos[j[ok]]), initial=0))
