"""Profiling driver: tuned triangle counting passes on RMAT scale 22 (ef 16, seed 2),
followed by a 1 GiB device-to-device copy whose counters calibrate FETCH_SIZE and
WRITE_SIZE.  Run it under rocprofv3, e.g.

    rocprofv3 --pmc FETCH_SIZE -T -d out -o p --output-format csv -- \
        python -u simdgraphprocessing_amd/profile_tc.py
"""
import argparse
import ctypes as C
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import simdgraphprocessing_amd as eh  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--policy", default="hybrid")
    ap.add_argument("--passes", type=int, default=5)
    a = ap.parse_args()
    ctx = eh.Context(0)
    g = ctx.generate_rmat(a.scale, 16, 2)
    M = g.build_matrix("lower", a.policy)
    c, st = M.triangle_count("tuned")
    for _ in range(a.passes):
        c, st = M.triangle_count("tuned")
    print("tuned", c, st["kernel_ms"], flush=True)
    hip = C.CDLL("libamdhip64.so")
    src, dst, n = C.c_void_p(), C.c_void_p(), 1 << 30
    assert hip.hipMalloc(C.byref(src), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(dst), C.c_size_t(n)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(n)) == 0
    for _ in range(2):
        assert hip.hipMemcpy(dst, src, C.c_size_t(n), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    print("copy done", flush=True)


if __name__ == "__main__":
    main()
