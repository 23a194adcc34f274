"""Import helper for the `rbf-fd-swe_amd/` package directory (its name is not
a Python identifier): `from swe_pkg import pkg` gives the package as
`rbffd_swe_amd`."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "rbf-fd-swe_amd")
NAME = "rbffd_swe_amd"


def load():
    if NAME in sys.modules:
        return sys.modules[NAME]
    spec = importlib.util.spec_from_file_location(
        NAME, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[NAME] = mod
    spec.loader.exec_module(mod)
    return mod


pkg = load()
