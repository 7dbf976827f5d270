"""Profiling driver for triangle counting on RMAT scale 22 (ef 16, seed 2).

run: tuned counting passes (one warm-up plus --passes), then a 1 GiB device-to-device
copy whose counters calibrate FETCH_SIZE and WRITE_SIZE.  Run it under rocprofv3, one
counter per run:

    rocprofv3 --pmc FETCH_SIZE -T -d OUT_F -o p --output-format csv -- \\
        python -u simdgraphprocessing_amd/profile_tc.py run --policy hybrid
    rocprofv3 --pmc WRITE_SIZE -T -d OUT_W -o p --output-format csv -- \\
        python -u simdgraphprocessing_amd/profile_tc.py run --policy hybrid

summarize: turn the two counter CSVs into HBM bytes per counting pass and merge them
into a JSON file keyed by the configuration tag that bench.py looks up:

    python simdgraphprocessing_amd/profile_tc.py summarize --policy hybrid \\
        --fetch OUT_F/p_counter_collection.csv --write OUT_W/p_counter_collection.csv \\
        --out profiles/tc_rmat22_traffic.json
"""
import argparse
import csv
import ctypes as C
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

COPY_BYTES = 1 << 30
COUNT_KERNELS = ("k_count", "k_small_rows", "k_bstat", "k_bshg")


def run(a):
    import simdgraphprocessing_amd as eh
    from simdgraphprocessing_amd import selector
    ctx = eh.Context(0)
    if a.policy == "learned":
        ctx.set_selector(selector.load())
    g = ctx.generate_rmat(a.scale, 16, 2)
    M = g.build_matrix("lower", a.policy)
    kw = {}
    if a.small_degree:
        kw["small_degree"] = a.small_degree
    if a.lds_words:
        kw["lds_words"] = a.lds_words
    if a.waves:
        kw["waves"] = a.waves
    if a.bstat:
        kw["bstat"] = a.bstat
    c, st = M.triangle_count("tuned", **kw)
    for _ in range(a.passes):
        c, st = M.triangle_count("tuned", **kw)
    print("tuned", a.policy, kw, c, st["kernel_ms"], flush=True)
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(COPY_BYTES)) == 0
    for _ in range(2):
        assert hip.hipMemcpy(dst, src, C.c_size_t(COPY_BYTES), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    print("copy done", flush=True)


def _rows(path, counter):
    with open(path, newline="") as f:
        return [(r["Kernel_Name"], float(r["Counter_Value"])) for r in csv.DictReader(f)
                if r["Counter_Name"] == counter]


def _is_kernel(name, base):
    """name is kernel `base` itself (template arguments or a signature may follow), not a
    longer name that starts with it (k_bstat vs k_bstat_fill)."""
    return name == base or name.startswith(base + "<") or name.startswith(base + "(")


def _per_kernel_mean(rows):
    out = {}
    for base in COUNT_KERNELS:
        v = [x for name, x in rows if _is_kernel(name, base)]
        if v:
            out[base] = sum(v) / len(v)
    return out


def _copy_kib(rows):
    v = [x for name, x in rows if "copyBuffer" in name or "copy" in name.lower()]
    assert len(v) >= 2, "calibration copy dispatches not found"
    return sum(v[-2:]) / 2.0


def summarize(a):
    fr, wr = _rows(a.fetch, "FETCH_SIZE"), _rows(a.write, "WRITE_SIZE")
    f_kib, w_kib = _per_kernel_mean(fr), _per_kernel_mean(wr)
    cf, cw = _copy_kib(fr), _copy_kib(wr)
    ff, wf = COPY_BYTES / (cf * 1024.0), COPY_BYTES / (cw * 1024.0)
    fetch = sum(f_kib.values()) * 1024.0 * ff
    write = sum(w_kib.values()) * 1024.0 * wf
    tag = f"rmat{a.scale}-ef16-s2-{a.policy}-tuned"
    entry = {"bytes_per_pass": fetch + write, "fetch_bytes": fetch, "write_bytes": write,
             "fetch_kib_counter": f_kib, "write_kib_counter": w_kib,
             "copy_calibration": {"bytes": COPY_BYTES, "fetch_kib": cf, "write_kib": cw,
                                  "fetch_factor": ff, "write_factor": wf},
             "command": "rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE) -T -d DIR -o p --output-format csv -- "
                        f"python -u simdgraphprocessing_amd/profile_tc.py run --policy {a.policy} {a.shape}".rstrip()}
    doc = {"configs": {}}
    if os.path.exists(a.out):
        with open(a.out) as f:
            old = json.load(f)
        if "configs" in old:
            doc = old
    doc["configs"][tag] = entry
    with open(a.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps({tag: entry["bytes_per_pass"]}))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--scale", type=int, default=22)
    r.add_argument("--policy", default="hybrid")
    r.add_argument("--passes", type=int, default=5)
    r.add_argument("--small-degree", type=int, default=0)
    r.add_argument("--lds-words", type=int, default=0)
    r.add_argument("--waves", type=int, default=0)
    r.add_argument("--bstat", type=int, default=0, help="B-stationary split (0: default, <0: off)")
    s = sub.add_parser("summarize")
    s.add_argument("--scale", type=int, default=22)
    s.add_argument("--policy", default="hybrid")
    s.add_argument("--shape", default="", help="launch-shape flags the run used, recorded in the entry")
    s.add_argument("--fetch", required=True)
    s.add_argument("--write", required=True)
    s.add_argument("--out", default=os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                 "profiles", "tc_rmat22_traffic.json"))
    a = ap.parse_args()
    (run if a.cmd == "run" else summarize)(a)


if __name__ == "__main__":
    main()
