"""Run configuration and CLI, counterpart of euler/config.py.

Same option names, choices and defaults as the reference CLI
(euler/config.py:32-60); the default file is this package's config.toml.
"""
from __future__ import annotations

import re
from argparse import ArgumentParser
from copy import deepcopy
from pathlib import Path

try:
    import tomllib  # Python >= 3.11
except ModuleNotFoundError:  # pragma: no cover - Python 3.10
    import tomli as tomllib

DEFAULT_CONFIG_PATH = Path(__file__).with_name("config.toml")

# euler/config.py:15-19, plus the two NACA shapes added for the batched cases
AOA_CASES = ("diamond", "naca0012", "naca2412", "rae2822", "oneraD", "oa209", "diamond_tx",
             "diamond_txy", "diamond_alpha10", "rae2822_txy", "diamond_sym", "naca23012",
             "diamond_big", "cylinder", "naca4412", "naca0015")
KNOWN_CASES = ("bump",) + AOA_CASES
AOA_CASES_LC = tuple(c.lower() for c in AOA_CASES)
KNOWN_CASES_LC = tuple(c.lower() for c in KNOWN_CASES)


def load_default_config(config_path=None):
    path = Path(config_path) if config_path is not None else DEFAULT_CONFIG_PATH
    with path.open("rb") as fh:
        cfg = tomllib.load(fh)
    cfg["export"] = deepcopy(cfg.get("export", {}))
    return cfg


def build_parser(default_cfg):
    p = ArgumentParser(description="2D Euler finite-volume solver (MI355X)")
    p.add_argument("--case", choices=KNOWN_CASES, default=default_cfg["case"])
    p.add_argument("--mach", type=float, default=default_cfg["Mach"])
    p.add_argument("--mesh-path", type=str, default=default_cfg["mesh_path"])
    p.add_argument("--time-scheme", choices=("EE", "RK2", "RK4", "SRK2", "SSP_RK2"),
                   default=default_cfg["time_scheme"])
    p.add_argument("--flux", choices=("Rusanov", "Tadmor", "AUSM", "Roe", "HLLC"), default=default_cfg["flux"])
    p.add_argument("--reconstruction", choices=("constant", "MUSCL", "muscl"),
                   default=default_cfg["reconstruction"])
    p.add_argument("--aoa", type=float, default=default_cfg.get("aoa", 0.0))
    p.add_argument("--stationarity-threshold", type=float,
                   default=default_cfg.get("stationarity_threshold", 1e-6))
    p.add_argument("--stationarity-check-every", type=int,
                   default=default_cfg.get("stationarity_check_every", 25))
    p.add_argument("--stationarity-patience", type=int, default=default_cfg.get("stationarity_patience", 1))
    p.add_argument("--verbose", dest="verbose", action="store_true", default=default_cfg.get("verbose", True))
    p.add_argument("--quiet", dest="verbose", action="store_false")
    return p


def build_config_from_cli(default_cfg, argv=None):
    """euler/config.py:32-60"""
    args = build_parser(default_cfg).parse_args(argv)
    cfg = deepcopy(default_cfg)
    cfg["case"] = args.case
    cfg["Mach"] = args.mach
    cfg["mesh_path"] = args.mesh_path
    cfg["time_scheme"] = "SRK2" if args.time_scheme == "SSP_RK2" else args.time_scheme
    cfg["flux"] = args.flux
    cfg["reconstruction"] = "MUSCL" if str(args.reconstruction).lower() == "muscl" else args.reconstruction
    cfg["aoa"] = float(args.aoa)
    cfg["stationarity_threshold"] = float(args.stationarity_threshold)
    cfg["stationarity_check_every"] = max(1, int(args.stationarity_check_every))
    cfg["stationarity_patience"] = max(1, int(args.stationarity_patience))
    cfg["verbose"] = bool(args.verbose)
    return cfg


_GEN = re.compile(r"^gen:([a-z0-9_]+):([0-9.eE+-]+)$")


def load_mesh(cfg):
    """euler/config.py:63-79. Also accepts ``gen:<geometry>:<h>`` (built by
    eulersr_amd.meshing and cached)."""
    path = cfg.get("mesh_path")
    if path is None:
        raise ValueError("The mesh path must be given with --mesh-path or in the default config.")
    m = _GEN.match(str(path))
    if m:
        from .meshing import cached_mesh
        return cached_mesh(m.group(1), float(m.group(2)))
    from .mesh import Mesh
    return Mesh.load(path)


def format_h(mesh):
    h = mesh.metadata.get("h", None)
    return "h" if h is None else f"h{float(h):.4f}".rstrip("0").rstrip(".")


def format_condition_tag(cfg):
    mach = f"M{float(cfg.get('Mach')):.2f}"
    if str(cfg.get("case", "")).lower() in AOA_CASES_LC:
        return f"AOA{float(cfg.get('aoa', 0.0)):.2f}_{mach}"
    return mach


def format_snapshot_name(cfg, t):
    """euler/config.py:97-103"""
    return (f"{format_condition_tag(cfg)}_{str(cfg['flux']).upper()}_{str(cfg['reconstruction']).upper()}_"
            f"{str(cfg['time_scheme']).upper()}_t{t:.2f}")


def setup_dirs(cfg, mesh, root=None):
    """euler/config.py:120-139 (rooted at ``root``, default ./results)."""
    root = Path(root) if root is not None else Path("results")
    case = str(cfg.get("case", ""))
    h_dir = format_h(mesh)
    if case.lower() in AOA_CASES_LC:
        tag = f"AOA{float(cfg.get('aoa', 0.0)):.2f}"
        dirs = {"fig": root / "figures" / case / h_dir / tag, "res": root / "raw" / case / h_dir / tag}
    else:
        dirs = {"fig": root / "figures" / case / h_dir, "res": root / "results" / case / h_dir}
    for d in dirs.values():
        d.mkdir(parents=True, exist_ok=True)
    return dirs


def print_config(cfg, mesh):
    rec, ts = str(cfg["reconstruction"]).upper(), str(cfg["time_scheme"]).upper()
    h = mesh.metadata.get("h", "n/a")
    if not cfg.get("verbose", True):
        print(f"[{cfg['case']}] {format_condition_tag(cfg)} | h={h} | {cfg['flux']}/{rec}/{ts}")
        return
    print("-" * 78)
    print(f"case {cfg['case']} | mesh {cfg['mesh_path']} (h={h}, {mesh.tris.shape[0]} cells)")
    print(f"Mach {cfg['Mach']}, AoA {float(cfg.get('aoa', 0.0)):.1f} deg, gamma={cfg['gamma']}, "
          f"p_inf={cfg['p_inf']}, rho_inf={cfg['rho_inf']}")
    print(f"scheme {cfg['flux']}, {rec}, {ts} | CFL {cfg['CFL']}, tf={cfg['tf']}")
    print(f"stationarity_threshold={cfg.get('stationarity_threshold')} | "
          f"check_every={cfg.get('stationarity_check_every')} | patience={cfg.get('stationarity_patience', 1)}")
    print("-" * 78)
