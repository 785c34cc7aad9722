"""Mesh layer: connectivity and geometry bit-exact against the loop-based
restatement of the reference conventions (oracle/mesh_ref.py), hand-checked
small meshes, the C++ reverse Cuthill-McKee order equal to scipy's
reverse_cuthill_mckee (live and against committed permutations), and mesher
quality / cell-count pins."""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

from conftest import mesh_of
from oracle import mesh_ref, rcm as RCM


@pytest.mark.parametrize("geom,h", [("diamond", 0.2), ("naca0012", 0.2), ("naca4412", 0.1)])
def test_connectivity_and_geometry_bit_exact(geom, h):
    m = mesh_of(geom, h)
    faces, nb, bnd = mesh_ref.faces_neighbors_from_tris(m.tris)
    assert np.array_equal(faces, m.faces)
    assert np.array_equal(nb, m.neighbors)
    assert np.array_equal(bnd, m.face_markers > 0)
    assert np.array_equal(mesh_ref.face_connectivity(m.faces, m.tris), m.face_connectivity)
    bary, normals, area, surface = mesh_ref.geometry(m.points, m.tris, m.faces)
    for a, b in ((bary, m.barycenter), (normals, m.normals), (area, m.area), (surface, m.surface)):
        assert np.array_equal(a.view(np.uint64), np.asarray(b).view(np.uint64))


def test_two_triangle_mesh_by_hand():
    from eulersr_amd.mesh import Mesh, faces_neighbors_from_tris
    pts = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    tris = np.array([[0, 1, 2], [0, 2, 3]])
    faces, nb, bnd, _ = faces_neighbors_from_tris(tris, 4)
    m = Mesh.from_points(pts, tris, nb, faces, np.where(bnd, 4, 0))
    # edge 1 of triangle 0 is (1,2)? no: edges are (0,1),(1,2),(2,0); (2,0) is shared
    assert nb.tolist() == [[-1, -1, 1], [0, -1, -1]]
    assert np.allclose(m.area, [0.5, 0.5])
    assert np.allclose(m.barycenter, [[2 / 3, 1 / 3], [1 / 3, 2 / 3]])
    # outward normals of the CCW triangle 0: bottom (0,-1), right (1,0), diagonal (-1,1)/sqrt2
    s = 1 / math.sqrt(2)
    assert np.allclose(m.normals[0], [[0, -1], [1, 0], [-s, s]])
    assert np.allclose(m.normals[1][0], [s, -s])           # shared edge, opposite direction
    assert np.array_equal(m.normals[1][0], -m.normals[0][2])
    assert int(np.sum(m.face_markers == 0)) == 1 and len(m.faces) == 5


@pytest.mark.parametrize("geom,h", [("diamond", 0.1), ("naca2412", 0.1)])
def test_mesh_invariants(geom, h):
    m = mesh_of(geom, h)
    n = m.n_cells
    # neighbour symmetry: nbr[nbr[i,k], k'] == i
    for i in range(0, n, 5):
        for k in range(3):
            j = m.neighbors[i, k]
            if j >= 0:
                assert i in m.neighbors[j]
    # Euler relation for triangles: 2 Nf = 3 N + Nb
    nb = int(np.sum(m.face_markers > 0))
    assert 2 * len(m.faces) == 3 * n + nb
    # counter-clockwise, positive areas, closed cells (sum of n*len = 0)
    p = m.points[m.tris]
    cross = (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - (p[:, 1, 1] - p[:, 0, 1]) * (p[:, 2, 0] - p[:, 0, 0])
    assert np.all(cross > 0)
    closure = np.sum(m.normals * m.surface[m.face_connectivity][..., None], axis=1)
    assert np.max(np.abs(closure)) < 1e-12
    # markers: wall faces lie on the body, inlet on x = x0, outlets elsewhere on the box
    fm = m.face_markers
    assert set(np.unique(fm)) == {0, 2, 3, 4}
    mid = m.points[m.faces].mean(axis=1)
    assert np.allclose(mid[fm == 3, 0], 0.0)


def test_mesher_quality_and_published_cell_counts():
    """Cell counts near the reference's published ones (README: 1.4k cells at
    h=0.1, 20.2k at h=0.025 for the diamond); meshes are not bit-identical to
    meshpy/Triangle's, so the pin is a +-20 % band. Minimum angle >= 28 deg."""
    m1 = mesh_of("diamond", 0.1)
    assert 0.8 * 1400 <= m1.n_cells <= 1.2 * 1400
    m2 = mesh_of("diamond", 0.025)
    assert 0.8 * 20200 <= m2.n_cells <= 1.2 * 20200
    p = m1.points[m1.tris]

    def ang(a, b, c):
        u, v = b - a, c - a
        return np.degrees(np.arccos(np.clip((u * v).sum(1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1),
                                            -1, 1)))
    mn = np.minimum(np.minimum(ang(p[:, 0], p[:, 1], p[:, 2]), ang(p[:, 1], p[:, 2], p[:, 0])),
                    ang(p[:, 2], p[:, 0], p[:, 1]))
    assert mn.min() >= 28.0


def _esr_rcm(nb, seeds=None):
    from eulersr_amd import _lib as L
    nb = np.ascontiguousarray(nb, dtype=np.int32)
    perm = np.empty(len(nb), dtype=np.int32)
    sp = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.int32)
    st = L.lib().esr_rcm(len(nb), nb.ctypes.data_as(C.POINTER(C.c_int32)),
                         None if sp is None else sp.ctypes.data_as(C.POINTER(C.c_int32)),
                         perm.ctypes.data_as(C.POINTER(C.c_int32)))
    assert st == 0
    return perm


def _scipy_rcm(nb, diag=None):
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    nb = np.asarray(nb)
    n = len(nb)
    i = np.repeat(np.arange(n), 3)
    j = nb.reshape(-1)
    ok = j >= 0
    r, c = i[ok], j[ok]
    if diag is not None:            # self loops on these nodes (degree + 2 in scipy's count)
        r, c = np.concatenate([r, diag]), np.concatenate([c, diag])
    A = csr_matrix((np.ones(len(r)), (r, c)), shape=(n, n))
    return np.asarray(reverse_cuthill_mckee(A, symmetric_mode=True))


RCM_MESHES = (("diamond", 0.1), ("diamond", 0.2), ("naca0012", 0.2), ("naca4412", 0.1), ("naca2412", 0.1))


@pytest.mark.parametrize("geom,h", RCM_MESHES)
def test_rcm_equals_scipy(geom, h):
    """The library's ordering, given the seed order scipy takes from
    numpy.argsort of the degrees (eulersr_amd.mesh.rcm_seed_order, what
    DeviceMesh passes), equals scipy.sparse.csgraph.reverse_cuthill_mckee
    element for element; so does the oracle's restatement."""
    from eulersr_amd.mesh import rcm_seed_order
    m = mesh_of(geom, h)
    sp = _scipy_rcm(m.neighbors)
    seeds = rcm_seed_order(m.neighbors)
    assert np.array_equal(seeds, RCM.seed_order(m.neighbors))
    assert np.array_equal(_esr_rcm(m.neighbors, seeds), sp)
    assert np.array_equal(RCM.rcm(m.neighbors), sp)
    assert RCM.bandwidth(m.neighbors, sp) < RCM.bandwidth(m.neighbors, np.arange(m.n_cells))


@pytest.mark.parametrize("geom,h", RCM_MESHES)
def test_rcm_default_seed_equals_scipy_from_that_seed(geom, h):
    """Without a seed order the library seeds at the lowest-index cell of
    least degree. scipy run on the same graph with self loops on every other
    cell (which raises their degree by 2 in scipy's count, leaves the
    neighbour order of every other cell unchanged, and so makes that cell
    scipy's unique seed) gives the same permutation: the BFS, the degree
    tie-break and the reversal are scipy's, whatever numpy's sort does."""
    m = mesh_of(geom, h)
    deg = (np.asarray(m.neighbors) >= 0).sum(axis=1)
    s = int(np.argsort(deg, kind="stable")[0])
    others = np.setdiff1d(np.arange(m.n_cells), [s])
    sp = _scipy_rcm(m.neighbors, diag=others)
    assert sp[-1] == s
    assert np.array_equal(_esr_rcm(m.neighbors), sp)
    assert np.array_equal(RCM.rcm(m.neighbors, np.argsort(deg, kind="stable")), sp)


def test_rcm_golden_scipy_permutations():
    """Permutations scipy computed (tests/golden/make_rcm.py), each with the
    seed order it used: the library and the oracle reproduce them bit for
    bit from the stored neighbour tables."""
    gold = np.load(Path(__file__).parent / "golden" / "rcm_scipy.npz")
    tags = sorted({k.rsplit("_", 1)[0] for k in gold.files})
    assert len(tags) >= 5
    for t in tags:
        nb, seeds, perm = gold[f"{t}_neighbors"], gold[f"{t}_seeds"], gold[f"{t}_perm"]
        assert np.array_equal(_esr_rcm(nb, seeds), perm), t
        assert np.array_equal(RCM.rcm(nb, seeds), perm), t


def test_rcm_oracle_equals_scipy_on_random_graphs():
    """The oracle restates scipy's algorithm on general symmetric CSR graphs:
    disconnected components, isolated nodes and self loops included."""
    import scipy.sparse as sparse
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    rng = np.random.default_rng(7)
    for _ in range(200):
        n = int(rng.integers(1, 50))
        k = int(rng.integers(0, 3 * n + 1))
        i, j = rng.integers(0, n, k), rng.integers(0, n, k)
        A = sparse.csr_matrix((np.ones(k), (i, j)), shape=(n, n))
        A = ((A + A.T) > 0).astype(np.float64).tocsr()
        A.sort_indices()
        assert np.array_equal(RCM.rcm_csr(A.indptr, A.indices), reverse_cuthill_mckee(A, symmetric_mode=True))


def test_rcm_rejects_bad_seed_order():
    from eulersr_amd import _lib as L
    nb = np.array([[1, -1, -1], [0, -1, -1]], dtype=np.int32)
    perm = np.empty(2, dtype=np.int32)
    bad = np.array([0, 0], dtype=np.int32)
    st = L.lib().esr_rcm(2, nb.ctypes.data_as(C.POINTER(C.c_int32)), bad.ctypes.data_as(C.POINTER(C.c_int32)),
                         perm.ctypes.data_as(C.POINTER(C.c_int32)))
    assert st != 0


def test_rcm_handles_disconnected_components():
    nb = np.array([[1, -1, -1], [0, -1, -1], [3, -1, -1], [2, -1, -1], [-1, -1, -1]], dtype=np.int32)
    perm = _esr_rcm(nb)
    assert np.array_equal(perm, RCM.rcm(nb, np.argsort((nb >= 0).sum(axis=1), kind="stable")))
    assert sorted(perm.tolist()) == [0, 1, 2, 3, 4]
    assert np.array_equal(_esr_rcm(nb, RCM.seed_order(nb)), _scipy_rcm(nb))


def test_periodic_pairing():
    """A periodic box (all boundary faces marker 1): every boundary face finds
    its translate, neighbours cross the domain and the opposite-face map is an
    involution (mesh.py:266-321)."""
    from eulersr_amd.mesh import Mesh, faces_neighbors_from_tris
    k = 4
    xs = np.linspace(0, 1, k + 1)
    P = np.array([[x, y] for y in xs for x in xs])
    tris = []
    for j in range(k):
        for i in range(k):
            a, b, c, d = j * (k + 1) + i, j * (k + 1) + i + 1, (j + 1) * (k + 1) + i + 1, (j + 1) * (k + 1) + i
            tris += [[a, b, c], [a, c, d]]
    tris = np.array(tris)
    faces, nb, bnd, _ = faces_neighbors_from_tris(tris, len(P))
    m = Mesh.from_points(P, tris, nb, faces, np.where(bnd, 1, 0))
    fco = m.face_connectivity_opposite
    b = np.nonzero(m.face_markers == 1)[0]
    assert np.all(fco[b] != b) and np.array_equal(fco[fco[b]], b)
    assert np.all(m.neighbors >= 0)
    mid = m.points[m.faces].mean(axis=1)
    d = mid[b] - mid[fco[b]]
    assert np.allclose(np.sort(np.abs(d), axis=1)[:, 0], 0) and np.allclose(np.abs(d).max(axis=1), 1.0)


def test_save_load_round_trip(tmp_path):
    from eulersr_amd.mesh import Mesh
    m = mesh_of("diamond", 0.2)
    p = m.save(tmp_path / "d.npz")
    m2 = Mesh.load(p)
    for k in ("points", "tris", "neighbors", "faces", "face_markers", "face_connectivity", "normals", "area"):
        assert np.array_equal(getattr(m, k), getattr(m2, k))
    assert m2.metadata["obstacle_length"] == 1.0 and m2.metadata["case"] == "diamond"
