"""Writes tests/golden/rcm_scipy.npz: reverse Cuthill-McKee permutations of
cell graphs computed by scipy.sparse.csgraph.reverse_cuthill_mckee (scipy
1.15.3, symmetric_mode=True), the pin of the library's renumbering.

The cell graph of each mesh is the CSR matrix of its (N,3) neighbour table
(-1 = boundary; no self loops). scipy seeds each component at the first
unvisited node of numpy.argsort(degree); among nodes of equal degree that
order is numpy's (unstable sort, SIMD-dispatched), so each entry also keeps
the seed order scipy used here. Given that order, the permutation is fixed
by the algorithm alone, on any machine.

Contents, per mesh tag T (geometry and h):
  T_neighbors   (N,3) int32 neighbour table (this project's mesher)
  T_seeds       (N,) int32 numpy.argsort of the node degrees, as scipy made it
  T_perm        (N,) int32 scipy's permutation, perm[new] = old

Usage (from the repository root): python tests/golden/make_rcm.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

MESHES = (("diamond", 0.1), ("diamond", 0.2), ("naca0012", 0.2), ("naca2412", 0.1), ("naca4412", 0.1))


def scipy_rcm(neighbors):
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    nb = np.asarray(neighbors)
    n = len(nb)
    i = np.repeat(np.arange(n), 3)
    j = nb.reshape(-1)
    ok = j >= 0
    A = sp.csr_matrix((np.ones(int(ok.sum())), (i[ok], j[ok])), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    deg = np.diff(A.indptr).astype(np.int32)      # no diagonal in a cell graph
    return np.asarray(reverse_cuthill_mckee(A, symmetric_mode=True), dtype=np.int32), \
        np.argsort(deg).astype(np.int32)


def main():
    from eulersr_amd.meshing import build_mesh
    out = {}
    for g, h in MESHES:
        m = build_mesh(g, h)
        perm, seeds = scipy_rcm(m.neighbors)
        tag = f"{g}_h{h}"
        out[f"{tag}_neighbors"] = np.asarray(m.neighbors, dtype=np.int32)
        out[f"{tag}_seeds"] = seeds
        out[f"{tag}_perm"] = perm
        print(f"{tag}: {m.n_cells} cells, first seed {seeds[0]}")
    np.savez_compressed(HERE / "rcm_scipy.npz", **out)


if __name__ == "__main__":
    main()
