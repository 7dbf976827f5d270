"""The C-ABI library loads and exports every entry point include/eh.h declares
(no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

import simdgraphprocessing_amd as eh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "eh.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(eh_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = eh.lib()
    names = declared()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(eh.EXPORTS) == names


def test_status_strings_and_version():
    L = eh.lib()
    assert L.eh_abi_version() == 5
    for s in range(7):
        assert L.eh_status_str(s)


def test_struct_sizes_match_header(tmp_path):
    """The ctypes mirrors keep the C layout: compare against gcc's sizeof."""
    import shutil, subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    prog = tmp_path / "sz.c"
    prog.write_text('#include <stdio.h>\n#include "eh.h"\nint main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n",'
                    'sizeof(eh_run_opts), sizeof(eh_row_view), sizeof(eh_gen_spec), sizeof(eh_run_stats),'
                    'sizeof(eh_matrix_info), sizeof(eh_graph_info), sizeof(eh_run_stats));return 0;}')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert got[:6] == [C.sizeof(eh.RunOpts), C.sizeof(eh.RowView), C.sizeof(eh.GenSpec), C.sizeof(eh.RunStats),
                       C.sizeof(eh.MatrixInfo), C.sizeof(eh.GraphInfo)]


def test_context_without_gpu_fails_loudly():
    import subprocess, sys
    code = ("import simdgraphprocessing_amd as eh\n"
            "try:\n eh.Context(0)\nexcept eh.EHError as e:\n print('ERR', e.status)\nelse:\n print('OK')")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT,
                         env={**os.environ, "HIP_VISIBLE_DEVICES": os.environ.get("HIP_VISIBLE_DEVICES", "")}).stdout
    assert out.strip() in ("OK", "ERR 3")
