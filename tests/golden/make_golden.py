"""Writes tests/golden/diamond_h0.3.npz, the fixture of the golden tests.

The reference ships no tests, fixtures or golden vectors for this path, and
its JAX solver cannot run in this environment, so these vectors are written
by the CPU oracle (oracle/fvm.py, oracle/rcm.py) on one small diamond mesh
made by this project's mesher. They pin the oracle, the mesh derivation and
the renumbering against regressions, and give the GPU parity tests inputs
that do not depend on the mesher.

Contents:
  mesh_*     points, tris, neighbors, faces, face_markers, metadata (JSON),
             as built by eulersr_amd.meshing.build_mesh("diamond", 0.3)
  geo_*      barycenter, area, surface, normals, face_connectivity derived
             from them (mesh.py:223-260 conventions)
  rcm_perm   reverse Cuthill-McKee order of the cell graph (oracle/rcm.py),
             seeded at the lowest-index cell of least degree
  W, inlet   a seeded smooth perturbation of the M = 1.8, alpha = 5 deg
             freestream (float32 conserved) and the inlet state
  R64_<flux>_<recon>, R32_<flux>_<recon>
             oracle residuals in float64 and float32 arithmetic for every
             flux x reconstruction; RE64_/RE32_<flux>: the entropy flag
  dt         get_dt(W, CFL=0.5) in float32
  S64_<scheme>, S32_<scheme>
             one HLLC + MUSCL step of each time scheme with that dt
  forces     CD, CL of W (rho_inf 1, U_inf = 1.8 sqrt(1.4), L = 1)

Usage (from the repository root): python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
for p in (str(ROOT), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

OUT = HERE / "diamond_h0.3.npz"
GEOM, H = "diamond", 0.3
MACH, AOA, SEED = 1.8, 5.0, 7
FLUXES = ["hllc", "roe", "rusanov", "lf", "ausm", "tadmor"]
RECONS = ["constant", "muscl"]
ENTROPY_FLUXES = ["hllc", "rusanov"]
SCHEMES = ["EE", "RK2", "SRK2", "RK4"]
MESH_FIELDS = ["points", "tris", "neighbors", "faces", "face_markers"]
GEO_FIELDS = ["barycenter", "area", "surface", "normals", "face_connectivity"]


def kwargs(inlet, flux, recon, **extra):
    kw = dict(gamma=1.4, M=1.0, reconstruction=recon, flux=flux, value=inlet, entropy=False)
    kw.update(extra)
    return kw


def main():
    from conftest import base_cfg, perturbed_state
    from eulersr_amd.meshing import build_mesh
    from oracle import fvm as F
    from oracle import rcm as RCM

    mesh = build_mesh(GEOM, H)
    cfg = base_cfg(Mach=MACH, aoa=AOA)
    W, inlet = perturbed_state(mesh, cfg, seed=SEED)
    om32, om64 = F.omesh(mesh, np.float32), F.omesh(mesh, np.float64)
    W64, inlet64 = W.astype(np.float64), inlet.astype(np.float64)
    out = {f"mesh_{k}": np.asarray(getattr(mesh, k)) for k in MESH_FIELDS}
    out["mesh_metadata"] = np.frombuffer(json.dumps(mesh.metadata).encode(), dtype=np.uint8)
    out.update({f"geo_{k}": np.asarray(getattr(mesh, k)) for k in GEO_FIELDS})
    stable = np.argsort((np.asarray(mesh.neighbors) >= 0).sum(axis=1), kind="stable")
    out["rcm_perm"] = np.asarray(RCM.rcm(mesh.neighbors, stable), dtype=np.int32)
    out["W"], out["inlet"] = W, inlet
    for flux in FLUXES:
        for recon in RECONS:
            kw = kwargs(inlet, flux, recon)
            out[f"R32_{flux}_{recon}"] = F.residual(W, om32, **kw)
            out[f"R64_{flux}_{recon}"] = F.residual(W64, om64, **dict(kw, value=inlet64))
    for flux in ENTROPY_FLUXES:
        kw = kwargs(inlet, flux, "muscl", entropy=True)
        out[f"RE32_{flux}"] = F.residual(W, om32, **kw)
        out[f"RE64_{flux}"] = F.residual(W64, om64, **dict(kw, value=inlet64))
    dt = F.get_dt(W, om32, CFL=0.5)
    out["dt"] = np.asarray(dt, dtype=np.float32)
    kw = kwargs(inlet, "hllc", "muscl")
    for sc in SCHEMES:
        out[f"S32_{sc}"] = F.SCHEMES[sc](W, om32, dt, **kw)
        out[f"S64_{sc}"] = F.SCHEMES[sc](W64, om64, np.float64(dt), **dict(kw, value=inlet64))
    U = MACH * np.sqrt(1.4)
    out["forces"] = np.asarray(F.forces(W, om32, 1.0, U, 1.0), dtype=np.float64)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT.relative_to(ROOT)}: {mesh.n_cells} cells, {len(out)} arrays, "
          f"{OUT.stat().st_size / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
