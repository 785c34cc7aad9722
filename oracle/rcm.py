"""CPU oracle (TEST INFRASTRUCTURE ONLY): reverse Cuthill-McKee ordering of
the cell adjacency graph.

The pin is scipy's ``scipy.sparse.csgraph.reverse_cuthill_mckee`` (scipy
1.15.3, ``scipy/sparse/csgraph/_reordering.pyx``), whose published algorithm
this file restates step for step in plain Python:

* node degree = stored entries of the CSR row, plus one more when the row
  holds its diagonal (``_node_degrees``);
* components are seeded in the order ``numpy.argsort(degree)`` (the default,
  unstable sort: which of several minimum-degree nodes comes first is the
  sort's choice, and differs between numpy's SIMD dispatch targets); the
  caller may give that seed order explicitly;
* breadth-first search level by level; for each node the unvisited
  neighbours are appended in CSR order, then insertion-sorted by degree
  (stable: equal degrees keep CSR order);
* the final order is reversed.

The product's C++ ordering (eulersr_amd/csrc/esr_mesh.cpp) must equal this
one element for element for the same seed order. The reference has no
renumbering (the new layout is ours).
"""
from __future__ import annotations

import numpy as np


def cell_graph(neighbors):
    """CSR (indptr, indices) of the cell adjacency graph of a (N,3)
    neighbour table (-1 = none): sorted, duplicate-free, no self loops -- the
    canonical CSR scipy builds from these pairs."""
    nb = np.asarray(neighbors)
    n = len(nb)
    rows = [sorted({int(j) for j in nb[i] if j >= 0 and j != i}) for i in range(n)]
    indptr = np.zeros(n + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    indices = np.array([j for r in rows for j in r], dtype=np.int64)
    return indptr, indices


def node_degrees(indptr, indices):
    n = len(indptr) - 1
    deg = np.zeros(n, dtype=np.int32)
    for i in range(n):
        deg[i] = indptr[i + 1] - indptr[i]
        for jj in range(indptr[i], indptr[i + 1]):
            if indices[jj] == i:
                deg[i] += 1
                break
    return deg


def seed_order(neighbors):
    """The component seed order scipy uses for this graph on this machine."""
    indptr, indices = cell_graph(neighbors)
    return np.argsort(node_degrees(indptr, indices))


def rcm_csr(indptr, indices, seeds=None):
    """Reverse Cuthill-McKee of a symmetric CSR graph; ``seeds`` = component
    seed order (default ``numpy.argsort`` of the degrees). Returns perm with
    perm[new] = old."""
    n = len(indptr) - 1
    deg = node_degrees(indptr, indices)
    seeds = np.argsort(deg) if seeds is None else np.asarray(seeds)
    assert sorted(int(s) for s in seeds) == list(range(n)), "seed order must be a permutation"
    visited = np.zeros(n, dtype=bool)
    order = []
    for s in seeds:
        s = int(s)
        if visited[s]:
            continue
        visited[s] = True
        order.append(s)
        lo, hi = len(order) - 1, len(order)
        while lo < hi:
            for ii in range(lo, hi):
                i = order[ii]
                start = len(order)
                for jj in range(indptr[i], indptr[i + 1]):
                    j = int(indices[jj])
                    if not visited[j]:
                        visited[j] = True
                        order.append(j)
                # insertion sort of the new nodes by degree (stable)
                for kk in range(start + 1, len(order)):
                    v = order[kk]
                    ll = kk
                    while ll > start and deg[v] < deg[order[ll - 1]]:
                        order[ll] = order[ll - 1]
                        ll -= 1
                    order[ll] = v
            lo, hi = hi, len(order)
        if len(order) == n:
            break
    return np.array(order[::-1], dtype=np.int32)


def rcm(neighbors, seeds=None):
    """RCM of the cell adjacency graph of a (N,3) neighbour table."""
    indptr, indices = cell_graph(neighbors)
    return rcm_csr(indptr, indices, seeds)


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
