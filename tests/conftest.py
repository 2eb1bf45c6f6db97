"""Test configuration: the `gpu` marker, import paths, shared fixtures."""
import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path,
                                                  submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


# The package folder is named bm4d-of_amd; it is imported as bm4dof_amd.
bm4dof_amd = _load("bm4dof_amd", os.path.join(REPO, "bm4d-of_amd", "__init__.py"))
synth = _load("bm4dof_amd.synth", os.path.join(REPO, "bm4d-of_amd", "synth.py"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def ctx():
    c = bm4dof_amd.Context(0)
    yield c
    c.close()
