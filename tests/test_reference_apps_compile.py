"""The reference's own apps compile, unchanged, against this repository's C++ surface
(simdgraphprocessing_amd/include/emptyheaded/ over include/eh.h): the header swap of the
drop-in boundary.  ParallelBuffer, Table / ParallelTable and the global std names
(src/ParallelBuffer.hpp:4-24, src/Table.hpp:4-88, src/common.hpp:67-68) are what the apps
need beyond SparseMatrix, Set and ops.  Skipped where the reference's sources are absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("EH_REFERENCE_SRC", "/root/reference")
APPS = ["undirected_triangle_counting", "n_clique", "similar_nodes", "symbiosity", "undirected_lollipop_counting",
        "n_path"]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "apps")), reason="the reference's sources are not present")
@pytest.mark.parametrize("app", APPS)
def test_reference_app_compiles_against_the_headers(app):
    inc = os.path.join(ROOT, "simdgraphprocessing_amd", "include")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I", os.path.join(inc, "emptyheaded"),
                        "-I", inc, "-I", os.path.join(ROOT, "include"), os.path.join(REF, "apps", app + ".cpp")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
