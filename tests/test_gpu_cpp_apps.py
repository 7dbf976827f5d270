"""The C++ surface (SparseMatrix, Set, ops::set_intersect, the two apps and their
CLI) over the C-ABI, run as the reference's own tests run: facebook.bin must give
1,612,010 triangles and 30,004,668 4-cliques (test/undirected_triangle_counting_test.cpp:11,
test/n_clique_test.cpp:11) under every layout and both schedules, and get_row +
set_intersect must agree with std::set_intersection."""
import gzip
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "simdgraphprocessing_amd", "bin")


@pytest.fixture(scope="module")
def fb_path(tmp_path_factory):
    p = tmp_path_factory.mktemp("fbcpp") / "facebook.bin"
    with gzip.open(os.path.join(ROOT, "tests", "golden", "facebook.bin.gz"), "rb") as f:
        p.write_bytes(f.read())
    return str(p)


def run(args, timeout=110):
    exe = os.path.join(BIN, args[0])
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    r = subprocess.run([exe] + args[1:], capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("prog", ["test_triangles", "test_cliques", "test_set_ops"])
def test_cpp_app_tests(fb_path, prog):
    rc, out, err = run([prog, fb_path])
    assert rc == 0 and "PASSED" in out, out[-2000:] + err[-2000:]


@pytest.mark.gpu
def test_cli_triangles(fb_path):
    rc, out, err = run(["undirected_triangle_counting", f"--graph={fb_path}", "--input_type=binary", "--t=4",
                        "--layout=hybrid"])
    assert rc == 0, err
    assert "Number of triangles: 1612010" in out
    assert "ROW DATA SIZE (Bytes): 569724" in out


@pytest.mark.gpu
def test_cli_cliques(fb_path):
    rc, out, err = run(["n_clique", f"--graph={fb_path}", "--input_type=binary", "--t=4", "--layout=pshort",
                        "--schedule=plain"])
    assert rc == 0, err
    assert "Number of cliques: 30004668" in out


@pytest.mark.gpu
def test_cli_learned_layout(fb_path):
    params = os.path.join(ROOT, "simdgraphprocessing_amd", "selector_params.npy")
    env = dict(os.environ, EH_SELECTOR_PARAMS=params)
    exe = os.path.join(BIN, "undirected_triangle_counting")
    r = subprocess.run([exe, f"--graph={fb_path}", "--input_type=binary", "--t=4", "--layout=learned"],
                       capture_output=True, text=True, timeout=110, env=env)
    assert r.returncode == 0, r.stderr
    assert "Number of triangles: 1612010" in r.stdout
