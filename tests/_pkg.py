"""Imports the package directory ``exact-graphstate-sampling_amd`` (not a valid module name) as ``egs``."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "exact-graphstate-sampling_amd")


def _load():
    if "egs" in sys.modules:
        return sys.modules["egs"]
    spec = importlib.util.spec_from_file_location("egs", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["egs"] = mod
    spec.loader.exec_module(mod)
    for name in ("tableau", "graphform", "amp", "plan", "inplace", "engine", "gadgets", "instances", "planio", "gsamp",
                 "adjoint", "autotune", "sliced", "doped"):
        importlib.import_module(f"egs.{name}")
    return mod


egs = _load()
