"""Shared test setup: markers, package loader and fixture helpers."""
from __future__ import annotations

import importlib.util
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU test")


def load_package():
    """Import the product package from the hyphenated directory as ``mille_feuille_amd``."""
    if "mille_feuille_amd" in sys.modules:
        return sys.modules["mille_feuille_amd"]
    pkg_dir = ROOT / "mille-feuille_amd"
    spec = importlib.util.spec_from_file_location(
        "mille_feuille_amd", pkg_dir / "__init__.py", submodule_search_locations=[str(pkg_dir)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mille_feuille_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def mf():
    return load_package()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


TILE_CASES = ["add20", "poisson64", "ragged1000", "ragged1009", "densetile", "emptyrows", "unsorted"]


def load_tile_case(name):
    with np.load(GOLDEN / f"tile_{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
