"""Shared fixtures: the product package (spm-iqc_amd, loaded by path) and the
CPU oracle (oracle/pyoracle.py, the checker)."""
import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE_DIR = os.path.join(ROOT, "oracle")
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def load_pkg():
    return importlib.import_module("spm-iqc_amd")


def load_seeds():
    return importlib.import_module("spm-iqc_amd.seeds")


@pytest.fixture(scope="session")
def spm():
    return load_pkg()


@pytest.fixture(scope="session")
def seeds():
    return load_seeds()


@pytest.fixture(scope="session")
def po():
    import pyoracle

    return pyoracle
