"""Shared test set-up: the gpu marker, the oracle and the library binding."""
import importlib.util
import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "micromagnetic-gpu-monte-carlo_amd"
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_mmc():
    """Import the library binding (the package directory name has dashes)."""
    if "mmc_amd" in sys.modules:
        return sys.modules["mmc_amd"]
    spec = importlib.util.spec_from_file_location(
        "mmc_amd", PKG_DIR / "__init__.py",
        submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mmc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def mmc():
    mod = load_mmc()
    if not os.path.exists(mod.LIB_PATH):
        mod.build()
    mod.lib()
    return mod
