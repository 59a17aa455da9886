"""Import helper: the package directory ``blocklinearsystems.jl_amd`` is not a Python identifier, so it is
loaded under the module name ``blsamd``."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "blocklinearsystems.jl_amd")


def load():
    if "blsamd" in sys.modules:
        return sys.modules["blsamd"]
    spec = importlib.util.spec_from_file_location("blsamd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["blsamd"] = mod
    spec.loader.exec_module(mod)
    return mod
