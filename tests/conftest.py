"""Shared fixtures: the package (libbo_hip.so bindings), the CPU oracle, and
small model builders restating the reference's test models."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def pkg():
    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def po():
    from oracle import pyoracle
    return pyoracle


@pytest.fixture
def ref(po, request):
    """The reference build (oracle/_ref), skipping when it is not present.
    CPU tests only: GPU tests compare with the committed golden fixtures
    written from it, so no GPU run loads the reference build."""
    if request.node.get_closest_marker("gpu") is not None:
        pytest.fail("a GPU test asked for the reference build; use the golden fixtures")
    if po.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return po


@pytest.fixture(scope="session")
def lpdata(pkg):
    return pkg.lpdata
