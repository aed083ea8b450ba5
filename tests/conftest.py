import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_INPUT = os.path.join(GOLDEN, "ref_input")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle


def read_snapcoeff(path):
    """beta_0..beta_n of the first element of a .snapcoeff file"""
    lines = [l.split("#")[0].strip() for l in open(path)]
    lines = [l for l in lines if l]
    ne, nc = map(int, lines[0].split())
    return [float(x) for x in lines[2:2 + nc]]
