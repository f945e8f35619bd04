"""Shared fixtures.  Tests marked `gpu` need an MI355X; the rest run on CPU."""
import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (run with -m gpu on the MI355X box)")


@pytest.fixture(scope="session")
def sysdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("systems"))


def _system(sysdir, name):
    from md_amd import systems
    return systems.cached_system(name, sysdir)


@pytest.fixture(scope="session")
def polymer_2k(sysdir):
    return _system(sysdir, "polymer_2k")


@pytest.fixture(scope="session")
def water_6k(sysdir):
    return _system(sysdir, "water_box_6k")


@pytest.fixture(scope="session")
def polymer_29k(sysdir):
    return _system(sysdir, "polymer_29k")


@pytest.fixture(scope="session")
def scale_100k(sysdir):
    return _system(sysdir, "scale_100k")
