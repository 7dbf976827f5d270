import gzip
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libeh.so)")


@pytest.fixture(scope="session")
def facebook(tmp_path_factory):
    """SNAP ego-Facebook in the reference's binary format (test/data/facebook.bin)."""
    import oracle as O
    path = tmp_path_factory.mktemp("fb") / "facebook.bin"
    with gzip.open(os.path.join(GOLDEN, "facebook.bin.gz"), "rb") as f:
        path.write_bytes(f.read())
    n, off, nbr, ext = O.load_binary(str(path))
    return {"path": str(path), "n": n, "off": off, "nbr": nbr, "ext": ext}


@pytest.fixture(scope="session")
def ctx():
    import simdgraphprocessing_amd as eh
    c = eh.Context(0)
    yield c
    c.close()
