"""Benchmark: triangle counting on RMAT scale 22 (edge factor 16, a/b/c/d =
.57/.19/.19/.05, seed 2), degree-oriented, uint32 ids; counts bit-exact against the CPU
restatement of the reference timed in the same run.

One step = one counting pass (eh_triangle_count) over the device-resident matrix, at the
launch shape committed for that workload in simdgraphprocessing_amd/shapes.json (the
autotuner's pick; --tune searches again).  Prints one JSON line (rank 0, N=1).

    python bench.py [--steps K] [--warmup W] [--configs c1,c3,c4,c5,bfs] [--tune]

Profiling mode (what the rocprofv3 summaries under profiles/ were taken from):

    python bench.py --profile-run CONFIG [--policy P] [--passes N]

builds CONFIG's graph and matrix exactly as the benchmark does, runs N counting passes at
the committed shape, then one 1 GiB device copy whose FETCH_SIZE / WRITE_SIZE calibrate
the counters; `python bench.py --traffic-summary ...` turns two such counter runs into
the HBM bytes per pass recorded in profiles/traffic.json.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from simdgraphprocessing_amd.workloads import C1, C3, C4, C5, RMAT, build, load_selector, run_pass, shape_tag  # noqa: E402

HBM_PEAK_GBS = 8000.0
TRAFFIC_PATH = os.path.join(ROOT, "profiles", "traffic.json")
COPY_BYTES = 1 << 30


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ host description

def cpu_info():
    """The host the CPU baseline ran on: /proc/cpuinfo's model name with family / model /
    stepping (a hypervisor may shorten the name), the vCPUs this process may use and the
    cgroup CPU quota."""
    info = {"model_name": platform.processor() or "unknown"}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    info["model_name"] = v
                elif k in ("cpu family", "model", "stepping", "vendor_id") and k not in info:
                    info[k.replace(" ", "_")] = v
                elif k == "flags":
                    info["hypervisor"] = "hypervisor" in v.split()
                    info["avx512"] = "avx512f" in v.split()
                elif k == "" and "model_name" in info and "cpu_family" in info:
                    break
    except OSError:
        pass
    info["hardware_concurrency"] = os.cpu_count()
    info["affinity_cpus"] = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, p = f.read().split()
            info["cgroup_cpus"] = None if q == "max" else int(q) / int(p)
    except (OSError, ValueError):
        pass
    return info


def cpu_threads():
    """Every CPU this process may run on (the reference's hardware_concurrency())."""
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def median_runs(fn, runs=5, warmup=1):
    """fn() -> (result, seconds); warm-up runs, then the median of `runs` timed ones."""
    for _ in range(warmup):
        fn()
    res, secs = [], []
    for _ in range(runs):
        r, s = fn()
        res.append(r)
        secs.append(s)
    assert all(r == res[0] for r in res), "CPU restatement gave different results run to run"
    return res[0], sorted(secs)[len(secs) // 2], secs


# ------------------------------------------------------------------ shapes and traffic

def committed_shape(tag):
    from simdgraphprocessing_amd import autotune
    return autotune.committed(tag)


def traffic_key(config, policy, params):
    return f"{config}|{policy}|{json.dumps(params or {}, sort_keys=True)}"


def traffic_record(config, policy, params):
    """HBM bytes per pass measured by rocprofv3 --pmc at exactly this launch shape (from
    profiles/traffic.json), or None."""
    try:
        with open(TRAFFIC_PATH) as f:
            return json.load(f).get("entries", {}).get(traffic_key(config, policy, params))
    except (OSError, ValueError):
        return None


def roofline_of(rec, kernel_ms, kernel=None):
    """Roofline on measured DRAM-side bytes: `kernel` = one kernel's bytes over its time,
    else the pass's bytes over the pass time."""
    if not rec or not kernel_ms:
        return {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None, "basis": "no counter record for this exact shape"}
    b = rec["kernels"][kernel]["bytes"] if kernel else rec["bytes_per_pass"]
    gbs = b / (kernel_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "traffic": b, "basis": "rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated, same shape"}


def byte_figures(alg_bytes, kernel_ms, rec):
    """Per-config line: algorithmic bytes (the naive canonical-CSR model) as effective GB/s,
    and the measured DRAM-side bytes as GB/s and fraction of peak."""
    out = {"alg_bytes": alg_bytes, "effective_GBs": alg_bytes / (kernel_ms * 1e-3) / 1e9 if kernel_ms else None}
    r = roofline_of(rec, kernel_ms)
    out.update({"dram_bytes": r["traffic"], "dram_GBs": r["achieved"], "roofline_frac": r["frac"]})
    return out


# ------------------------------------------------------------------ CPU baselines

def cpu_triangles(n, off, nbr, gpu_count, runs=5, budget_s=30.0):
    """The oracle (C restatement of the reference hybrid path, dynamic chunks of 100 rows) on
    every host CPU: the median of `runs` full passes after a warm-up, or of a row sample
    when a full pass would overrun the budget."""
    import oracle as O
    threads = cpu_threads()
    M = O.Matrix(n, off, nbr, "hybrid", lower=True, threads=threads)
    _, _, t_probe = M.triangles_sample(threads, 64, 0)
    stride = 1
    while t_probe * 64 / stride * (runs + 1) > budget_s and stride < 64:
        stride *= 2
    if stride == 1:
        c, secs, all_s = median_runs(lambda: M.triangles(threads), runs)
        assert c == gpu_count, f"CPU count {c} != GPU count {gpu_count}"
        edges, sample = M.cardinality(), "full graph"
    else:
        (c, edges), secs, all_s = median_runs(
            lambda: (lambda r: ((r[0], r[1]), r[2]))(M.triangles_sample(threads, stride, 0)), runs)
        sample = f"source rows i with i % {stride} == 0 ({edges} of {M.cardinality()} oriented edges)"
    return {"value": edges / secs, "unit": "edges/s", "cores": threads, "kind": "port", "sample": sample,
            "seconds": secs, "runs_s": all_s, "timing": f"median of {runs} after 1 warm-up",
            "count_checked": stride == 1, "cpu": cpu_info(),
            "oracle": "oracle/eh_oracle.c hybrid policy, dynamic chunks of 100 (common.hpp:213-276)"}


def cpu_reference_triangles(n, off, nbr, gpu_count, runs=3):
    """The reference's own triangle program (apps/undirected_triangle_counting.cpp, built from its
    sources into oracle/_ref by the oracle's Makefile) on the same graph and every host CPU: the
    median over `runs` runs of the counting phase it times itself (common::stopClock, its "Time[...]"
    line), the count checked; None when the program was not built."""
    import oracle as O
    if not O.ref_available():
        return None
    threads = cpu_threads()
    stats = [O.run_ref_stats("undirected_triangle_counting", n, off, nbr, threads=threads, layout="hybrid")
             for _ in range(runs)]
    for st in stats:
        assert st["count"] == gpu_count, f"reference count {st['count']} != GPU count {gpu_count}"
    secs = sorted(st["seconds"] for st in stats)[len(stats) // 2]
    edges = int(off[-1]) // 2
    return {"value": edges / secs, "unit": "edges/s", "cores": threads, "kind": "reference",
            "sample": "full graph", "seconds": secs, "runs_s": [st["seconds"] for st in stats],
            "timing": f"median of {runs}, the program's own counting-phase timer", "count_checked": True,
            "program": "oracle/_ref/undirected_triangle_counting --layout=hybrid"}


# ------------------------------------------------------------------ configs

def config_c1(ctx, with_cpu):
    """ER G(10000, 100000 distinct non-loop pairs), seed 1: triangles."""
    g, M = build(ctx, "c1", "hybrid")
    params = committed_shape("c1")
    c, stp = M.triangle_count("plain", repeat=20)
    c2, st = M.triangle_count("tuned", repeat=20, **params)
    assert c == c2
    m = M.info()["cardinality"]
    out = {"workload": "triangles, ER n=10000 with exactly 100000 distinct non-loop pairs, seed 1",
           "triangles": c, "oriented_edges": m, "kernel_ms": st["kernel_ms"],
           "edges_per_s": m / (st["kernel_ms"] * 1e-3), "plain_kernel_ms": stp["kernel_ms"], "params": params}
    out.update(byte_figures(st["alg_bytes"], st["kernel_ms"], traffic_record("c1", "hybrid", params)))
    if with_cpu:
        import oracle as O
        off, nbr = g.download()
        om = O.Matrix(C1[0], off, nbr, "hybrid", threads=cpu_threads())
        oc, secs, _ = median_runs(lambda: om.triangles(cpu_threads()))
        assert oc == c
        out["cpu"] = {"seconds": secs, "edges_per_s": m / secs, "cores": cpu_threads(), "kind": "port",
                      "timing": "median of 5 after 1 warm-up"}
    return out


def config_bfs(ctx, with_cpu, scale=22, ef=16, seed=2):
    """n_path's BFS (apps/n_path.cpp:27-87) on the RMAT graph of the headline, every entry kept,
    from its largest row (get_max_row_id), no depth limit."""
    g = ctx.generate_rmat(scale, ef, seed)
    M = g.build_matrix("all", "hybrid")
    i = M.info()
    start = 0  # rows are degree-relabelled: row 0 is the first row of maximum length
    runs = [M.bfs_path(start, mode="symmetric") for _ in range(5)]
    r = sorted(runs, key=lambda x: x["kernel_ms"])[len(runs) // 2]
    out = {"workload": f"BFS path length, RMAT scale {scale} ef {ef} seed {seed}, all entries, start = largest row",
           "path_length": r["path_length"], "rounds": r["rounds"], "reached": r["reached"],
           "level_sizes": r["level_sizes"], "kernel_ms": r["kernel_ms"], "bottom_up_rounds": r["bottom_up"],
           "top_down_kernel_ms": M.bfs_path(start)["kernel_ms"], "entries": i["cardinality"],
           "teps": i["cardinality"] / (r["kernel_ms"] * 1e-3)}
    if with_cpu:
        import oracle as O
        off, nbr = g.download()

        threads = cpu_threads()

        def one():
            t0 = time.perf_counter()
            ref = O.bfs_path(i["n"], off, nbr, start, cap=256, threads=threads)
            return (ref["path_length"], tuple(ref["level_sizes"])), time.perf_counter() - t0
        (pl, ls), secs, _ = median_runs(one)
        assert pl == r["path_length"] and list(ls) == r["level_sizes"]
        out["cpu"] = {"seconds": secs, "teps": i["cardinality"] / secs, "cores": threads, "kind": "port",
                      "timing": "median of 5 after 1 warm-up; each round's union over the threads, as n_path"}
    return out


def config_c3(ctx, with_cpu):
    """Chung-Lu n=1e6, gamma 2.1, 32 x n / 2 draws + 200 planted 256-vertex communities
    (p=.5), seed 3: 4-cliques."""
    g, M = build(ctx, "c3", "hybrid")
    c, stp = M.clique_count(4, "plain", repeat=2)
    params = committed_shape("c3")
    c2, st = M.clique_count(4, "tuned", repeat=5, time_parts=True, **params)
    assert c == c2
    i = M.info()
    gi = g.info()
    out = {"workload": "4-cliques, Chung-Lu n=1e6 gamma 2.1 mean weight 32 (16e6 draws) + 200x256 communities "
                       "p=.5, seed 3",
           "four_cliques": c, "oriented_edges": i["cardinality"], "mean_degree": gi["entries"] / gi["n"],
           "kernel_ms": st["kernel_ms"], "part_ms": st["part_ms"], "plain_kernel_ms": stp["kernel_ms"],
           "speedup_vs_plain": stp["kernel_ms"] / st["kernel_ms"], "params": params}
    out.update(byte_figures(st["alg_bytes"], st["kernel_ms"], traffic_record("c3", "hybrid", params)))
    if with_cpu:
        import oracle as O
        off, nbr = g.download()
        threads = cpu_threads()
        om = O.Matrix(i["n"], off, nbr, "hybrid", threads=threads)
        kc, secs, _ = median_runs(lambda: om.cliques(4, threads))
        assert kc == c, f"CPU 4-clique count {kc} != GPU {c}"
        out["cpu"] = {"seconds": secs, "sample": "full graph", "edges_per_s": i["cardinality"] / secs,
                      "cores": threads, "kind": "port", "count_checked": True, "timing": "median of 5 after 1 warm-up"}
        # the reference's own 4-clique program on the same graph and threads, beside the restatement
        if O.ref_available():
            runs = [O.run_ref_stats("n_clique", i["n"], off, nbr, threads=threads, layout="hybrid", k=4, timeout=120)
                    for _ in range(3)]
            assert all(r["count"] == c for r in runs), "reference 4-clique count differs from the GPU's"
            rs = sorted(r["seconds"] for r in runs)[1]
            ref = {"seconds": rs, "sample": "full graph", "edges_per_s": i["cardinality"] / rs, "cores": threads,
                   "kind": "reference", "count_checked": True, "program": "oracle/_ref/n_clique --layout=hybrid --n=4",
                   "timing": "median of 3, the program's own counting-phase timer"}
            # the faster of the two is the reported baseline, the other rides beside it
            if rs < secs:
                ref["port"] = out["cpu"]
                out["cpu"] = ref
            else:
                out["cpu"]["reference_program"] = ref
    return out


def config_c4(ctx, with_cpu):
    """4096 ER graphs (n in [64,8192], mean degree in [4,64)), seed 4: per-graph triangles + 4-cliques,
    the arena under the learned layouts when the trained selector is there (the density rule
    otherwise); every rule's layouts are timed beside them (``selector``)."""
    import numpy as np
    policy = "learned" if load_selector(ctx) else "hybrid"
    _, b = build(ctx, "c4", policy)
    info = b.info()
    params = committed_shape("c4")
    kw = {"tri_opts": params["tri"], "k4_opts": params["k4"]}
    tri, k4, _ = b.count("persistent", "tuned", **kw)
    ms_p = []
    for _ in range(5):
        t2, k2, ms = b.count("persistent", "tuned", **kw)
        assert np.array_equal(t2, tri) and np.array_equal(k2, k4)
        ms_p.append(ms)
    ms_pers = sorted(ms_p)[len(ms_p) // 2]
    # per-graph launches at the better of their default shape and the persistent one
    per_graph = {}
    shapes = (("default_shape", {}), ("persistent_shape", kw),
              ("short_rows_8_for_4cliques", {"k4_opts": {"small_degree": 8}}))
    for name, opts in shapes:
        tg, kg, ms_g = b.count("per_graph", "tuned", **opts)
        assert np.array_equal(tg, tri) and np.array_equal(kg, k4)
        per_graph[name] = ms_g
    ms_g = min(per_graph.values())
    tb, kb = b.alg_bytes()
    out = {"workload": "4096 ER graphs n~U[64,8192] mean degree~U[4,64) seed 4, triangles + 4-cliques per graph",
           "layout_policy": policy, "graphs": info["graphs"], "vertices": info["vertices"], "edges": info["edges"],
           "persistent_ms": ms_pers, "persistent_runs_ms": ms_p, "per_graph_ms": ms_g,
           "per_graph_by_shape_ms": per_graph, "speedup_vs_per_graph": ms_g / ms_pers, "params": params,
           "triangles_total": int(tri.sum()), "four_cliques_total": int(k4.sum()),
           "alg_bytes_tri": tb, "alg_bytes_k4": kb}
    out.update(byte_figures(tb + kb, ms_pers, traffic_record("c4", policy, params)))
    # the arena's layouts under the learned selection and every rule (rebuilt in place; the
    # persistent passes at the same shapes, each with its build time)
    from simdgraphprocessing_amd import selector
    if policy == "learned":
        out["selector"] = selector.compare_batch(ctx, b, steps=5, log=log)
    if with_cpu:
        # all 8192 counts against the oracle in one pass over the union graph
        import oracle as O
        tot_n = info["vertices"]
        n, m = b.params()
        vbase = np.concatenate([[0], np.cumsum(n.astype(np.int64))])
        off = np.zeros(tot_n + 1, np.uint64)
        chunks = []
        for gi in range(info["graphs"]):
            gn, goff, gnbr = b.graph(gi)
            off[vbase[gi] + 1: vbase[gi + 1] + 1] = goff[1:] + off[vbase[gi]]
            chunks.append(gnbr.astype(np.uint32) + np.uint32(vbase[gi]))
        nbr = np.concatenate(chunks)
        om = O.Matrix(tot_n, off, nbr, "hybrid", threads=cpu_threads())

        def one():
            t0 = time.perf_counter()
            _, rt = om.count_rows(3, cpu_threads())
            _, rk = om.count_rows(4, cpu_threads())
            return (rt.tobytes(), rk.tobytes()), time.perf_counter() - t0
        (rt, rk), secs, _ = median_runs(one, runs=3)
        rt, rk = np.frombuffer(rt, np.uint64), np.frombuffer(rk, np.uint64)
        gid = np.repeat(np.arange(info["graphs"]), n.astype(np.int64))
        ct = np.bincount(gid, weights=rt.astype(np.float64), minlength=info["graphs"]).astype(np.uint64)
        ck = np.bincount(gid, weights=rk.astype(np.float64), minlength=info["graphs"]).astype(np.uint64)
        assert np.array_equal(ct, tri) and np.array_equal(ck, k4), "batch counts differ from the oracle"
        out["cpu"] = {"seconds": secs, "cores": cpu_threads(), "kind": "port", "counts_checked": 2 * info["graphs"],
                      "timing": "median of 3 after 1 warm-up"}
    return out


def config_c5(ctx, with_cpu):
    """Learned selector (trained on engine-timed pass costs of RMAT 18-21, seeds 10-19) evaluated
    on RMAT-24, edge factor 16, seed 24: the tuned triangle pass under the learned layouts and
    under every rule the network is measured against (the reference's density rule, bitsets at
    density 1/512 and 1/4096, all uint), each with its build time, in the default pass ("tuned":
    every pair B-stationary, no layout read) and with the B-stationary kernels off ("aside":
    every pair A-side, where the layouts decide); plus the plain schedule."""
    from simdgraphprocessing_amd import selector
    g = ctx.generate_rmat(*C5)
    out = selector.evaluate_graph(ctx, g, steps=5, log=log)
    out["workload"] = f"triangles, RMAT scale {C5[0]} ef {C5[1]} seed {C5[2]}"
    # counter records exist for the profiled layout policies of the default pass
    for regime in ("tuned", "aside"):
        for name, r in out[regime]["policies"].items():
            pol = {"density_1_256": "hybrid", "learned": "learned"}.get(name) if regime == "tuned" else None
            r.update(byte_figures(out["alg_bytes"], r["kernel_ms"], traffic_record("c5", pol, {}) if pol else None))
    if with_cpu:
        off, nbr = g.download()
        del g
        out["cpu"] = cpu_triangles(1 << C5[0], off, nbr, out["tuned"]["policies"]["learned"]["count"], runs=3,
                                   budget_s=40.0)
    return out


# ------------------------------------------------------------------ profiling mode

def profile_run(args):
    """CONFIG's counting pass at its committed shape, --passes times, then the 1 GiB copy that
    calibrates FETCH_SIZE / WRITE_SIZE (run under rocprofv3 --pmc)."""
    import ctypes as C
    import simdgraphprocessing_amd as eh
    ctx = eh.Context(0)
    policy = args.policy if args.policy != "auto" else "hybrid"
    if policy == "learned":
        assert load_selector(ctx)
    params = committed_shape(shape_tag(args.profile_run, policy))
    _, M = build(ctx, args.profile_run, policy)
    r = run_pass(args.profile_run, M, params)
    for _ in range(args.passes):
        r = run_pass(args.profile_run, M, params)
    print(f"profile-run {args.profile_run} {policy} {params}: {r[0] if args.profile_run != 'c4' else r[2]}",
          flush=True)
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(COPY_BYTES)) == 0
    for _ in range(2):
        assert hip.hipMemcpy(dst, src, C.c_size_t(COPY_BYTES), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    print("copy done", flush=True)


# kernels of a counting pass (the batch adds its counter and error-word fills)
PASS_KERNELS = ("k_small_rows", "k_small_cliques4", "k_bstat", "k_bshl", "k_local4", "k_count",
                "k_zero64", "k_zero_err")


def _counter_rows(path, counter):
    import csv
    with open(path, newline="") as f:
        return [(r["Kernel_Name"], float(r["Counter_Value"])) for r in csv.DictReader(f)
                if r["Counter_Name"] == counter]


def _kernel_base(name):
    name = name.replace("(anonymous namespace)::", "")
    if name.startswith("void "):
        name = name[5:]
    for base in PASS_KERNELS:
        if name == base or name.startswith(base + "<") or name.startswith(base + "("):
            return base
    return None


def traffic_summary(args):
    """Two counter CSVs (FETCH_SIZE, WRITE_SIZE; KiB per dispatch) -> HBM bytes per pass of
    each kernel, scaled by what the same counters read for the 1 GiB copy; merged into
    profiles/traffic.json under the config's exact shape."""
    fr, wr = _counter_rows(args.fetch, "FETCH_SIZE"), _counter_rows(args.write, "WRITE_SIZE")

    def copy_kib(rows):
        v = [x for name, x in rows if "copy" in name.lower()]
        assert len(v) >= 2, "calibration copy dispatches not found"
        return sum(v[-2:]) / 2.0
    cf, cw = copy_kib(fr), copy_kib(wr)
    ff, wf = COPY_BYTES / (cf * 1024.0), COPY_BYTES / (cw * 1024.0)
    passes = args.passes + 1
    # per kernel class: the mean bytes of one dispatch times the dispatches every pass makes;
    # classes dispatched fewer times than there were passes ran outside the pass (the one-time
    # weighted pass behind the 4-clique byte model) and are left out
    acc = {}
    for rows, key, fac in ((fr, "fetch_bytes", ff), (wr, "write_bytes", wf)):
        for name, kib in rows:
            base = _kernel_base(name)
            if base:
                a = acc.setdefault(base, {"fetch_bytes": [], "write_bytes": []})
                a[key].append(kib * 1024.0 * fac)
    kernels = {}
    for base, a in acc.items():
        nd = len(a["fetch_bytes"])
        per_pass = nd // passes
        if per_pass == 0 or not a["write_bytes"]:
            continue
        k = {"dispatches": per_pass,
             "fetch_bytes": per_pass * sum(a["fetch_bytes"]) / nd,
             "write_bytes": per_pass * sum(a["write_bytes"]) / len(a["write_bytes"])}
        k["bytes"] = k["fetch_bytes"] + k["write_bytes"]
        kernels[base] = k
    policy = args.policy if args.policy != "auto" else "hybrid"
    params = committed_shape(shape_tag(args.profile_run, policy))
    entry = {"bytes_per_pass": sum(k["bytes"] for k in kernels.values()), "kernels": kernels,
             "calibration": {"copy_bytes": COPY_BYTES, "fetch_kib": cf, "write_kib": cw, "fetch_factor": ff,
                             "write_factor": wf},
             "command": f"rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE) --output-format csv -d DIR "
                        f"-o p -- python -u bench.py --profile-run {args.profile_run} --policy {policy} "
                        f"--passes {args.passes}"}
    doc = {"entries": {}}
    if os.path.exists(TRAFFIC_PATH):
        with open(TRAFFIC_PATH) as f:
            doc = json.load(f)
    doc.setdefault("entries", {})[traffic_key(args.profile_run, policy, params)] = entry
    with open(TRAFFIC_PATH, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    print(json.dumps({"config": args.profile_run, "policy": policy, "bytes_per_pass": entry["bytes_per_pass"]}))


# ------------------------------------------------------------------ headline

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--policy", default="auto", help="layouts of the timed pass: auto = learned when the "
                    "trained selector parameters exist, else hybrid")
    ap.add_argument("--configs", default="c1,c3,c4,c5,bfs", help="extra configs to report (c1,c3,c4,c5,bfs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="store_true", help="search launch shapes again instead of the committed ones")
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--profile-run", default="", help="profiling mode: c1, c2, c3, c4 or c5")
    ap.add_argument("--passes", type=int, default=5)
    ap.add_argument("--traffic-summary", action="store_true")
    ap.add_argument("--fetch", default="")
    ap.add_argument("--write", default="")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the counting path has "
              "no multi-GPU mode.", file=sys.stderr)
        sys.exit(2)
    if args.traffic_summary:
        return traffic_summary(args)
    if args.profile_run:
        return profile_run(args)

    import simdgraphprocessing_amd as eh
    from simdgraphprocessing_amd import autotune

    ctx = eh.Context(0)
    has_sel = load_selector(ctx)
    policy = ("learned" if has_sel else "hybrid") if args.policy == "auto" else args.policy
    if policy == "learned" and not has_sel:
        raise SystemExit("--policy learned needs simdgraphprocessing_amd/selector_params.npy")
    log(f"device: {ctx.device_name}")
    t0 = time.time()
    g = ctx.generate_rmat(*RMAT)
    gi = g.info()
    log(f"generated RMAT-{RMAT[0]}: n={gi['n']} entries={gi['entries']} max_deg={gi['max_degree']} "
        f"({time.time() - t0:.1f}s)")

    # plain: the reference layout rule + the reference's scheduling (one task per row, vertex order)
    Mp = g.build_matrix("lower", "hybrid")
    cp, stp = Mp.triangle_count("plain")
    plain_ms = []
    for _ in range(args.plain_steps):
        c, st = Mp.triangle_count("plain")
        assert c == cp
        plain_ms.append(st["kernel_ms"])
    plain_ms = sorted(plain_ms)[len(plain_ms) // 2] if plain_ms else stp["kernel_ms"]
    log(f"plain hybrid: count={cp} kernel {plain_ms:.3f} ms")
    del Mp

    # the density rule and (when trained) the learned layouts on the tuned schedule, each at
    # its committed launch shape; the timed pass is `policy`'s.  One-shot = layout build +
    # schedule build + first pass, from a fresh matrix, wall clock.
    legs = {}
    for pol in sorted({"hybrid", policy}):
        tw = time.perf_counter()
        Mc = g.build_matrix("lower", pol)
        params = autotune.tune(Mc, 3, "tuned", repeat=2, budget_s=45, use_cache=False)[0] if args.tune \
            else committed_shape(f"c2:{pol}")
        cc, st1 = Mc.triangle_count("tuned", **params)
        one_shot_ms = (time.perf_counter() - tw) * 1e3
        assert cc == cp, f"{pol}: tuned count {cc} != plain count {cp}"
        legs[pol] = {"matrix": Mc, "params": params, "one_shot_ms": one_shot_ms, "build_ms": Mc.info()["build_ms"],
                     "sched_ms": st1["sched_ms"], "first_pass_ms": st1["kernel_ms"]}
        log(f"{pol}: params {params}, build {legs[pol]['build_ms']:.2f} ms, schedule {st1['sched_ms']:.2f} ms, "
            f"one-shot {one_shot_ms:.1f} ms wall")
    M = legs[policy]["matrix"]
    params = legs[policy]["params"]
    info = M.info()
    for _ in range(args.warmup):
        c, st = M.triangle_count("tuned", time_parts=True, **params)
        assert c == cp, f"tuned count {c} != plain count {cp}"
    kms, parts = [], {}
    t_start = time.perf_counter()
    for _ in range(args.steps):
        c, st = M.triangle_count("tuned", time_parts=True, **params)
        kms.append(st["kernel_ms"])
        for k, v in st["part_ms"].items():
            parts[k] = parts.get(k, 0.0) + v / args.steps
    t_end = time.perf_counter()
    assert c == cp
    ms_step = (t_end - t_start) * 1e3 / args.steps
    kernel_ms = sum(kms) / len(kms)
    # the density rule's pass at its shape, for the learned-vs-rule ratio
    rule_ms = None
    if policy != "hybrid":
        Mh, ph = legs["hybrid"]["matrix"], legs["hybrid"]["params"]
        rule_ms = sorted(Mh.triangle_count("tuned", **ph)[1]["kernel_ms"] for _ in range(5))[2]
    m = info["cardinality"]
    alg_bytes = st["alg_bytes"]
    dominant = max(parts, key=parts.get)
    rec = traffic_record("c2", policy, params)
    roofline = roofline_of(rec, parts[dominant], dominant)
    roofline["kernel"] = dominant
    roofline["kernel_ms"] = parts[dominant]
    pass_roof = roofline_of(rec, kernel_ms)

    cpu = None
    if not args.no_cpu:
        off, nbr = g.download()
        cpu = cpu_triangles(gi["n"], off, nbr, cp)
        log(f"cpu baseline: {cpu['value']:.3e} edges/s on {cpu['cores']} threads ({cpu['sample']}, "
            f"{cpu['timing']})")
        # the reference's own program on the same graph and threads; the faster of the two is the
        # reported baseline, the other rides beside it
        ref = cpu_reference_triangles(gi["n"], off, nbr, cp)
        if ref is not None:
            log(f"reference program: {ref['value']:.3e} edges/s on {ref['cores']} threads ({ref['seconds']:.3f} s)")
            if ref["value"] > cpu["value"]:
                ref["cpu"] = cpu.pop("cpu")
                ref["port"] = cpu
                cpu = ref
            else:
                cpu["reference_program"] = ref
        del off, nbr
    policies = {k: {"params": v["params"], "build_ms": v["build_ms"], "sched_ms": v["sched_ms"],
                    "one_shot_ms": v["one_shot_ms"], "first_pass_ms": v["first_pass_ms"]} for k, v in legs.items()}
    if rule_ms is not None:
        policies["hybrid"]["kernel_ms"] = rule_ms
    policies[policy]["kernel_ms"] = kernel_ms
    legs.clear()
    del M
    # the learned layouts against every rule on this graph, pass and build (default launch shape)
    sel_c2 = None
    if has_sel:
        from simdgraphprocessing_amd import selector
        sel_c2 = {r: selector.compare(ctx, g, steps=5, log=log, regime=r) for r in ("tuned", "aside")}
    del g

    extra = {}
    wanted = [c for c in args.configs.split(",") if c]
    if "c1" in wanted:
        extra["c1"] = config_c1(ctx, not args.no_cpu)
    if "c3" in wanted:
        extra["c3"] = config_c3(ctx, not args.no_cpu)
    if "c4" in wanted:
        extra["c4"] = config_c4(ctx, not args.no_cpu)
    if "bfs" in wanted:
        extra["bfs"] = config_bfs(ctx, not args.no_cpu, *RMAT)
    if "c5" in wanted and has_sel:
        extra["c5"] = config_c5(ctx, not args.no_cpu)

    out = {
        "metric": "TC edges/s + HBM GB/s, RMAT scale-22; learned/batched vs plain, bit-exact",
        "value": m / (kernel_ms * 1e-3),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"triangle count, RMAT scale {RMAT[0]}, edge factor {RMAT[1]}, "
                               f"a/b/c/d .57/.19/.19/.05, seed {RMAT[2]}, degree-oriented",
                   "layout_policy": policy, "schedule": "tuned", "params": params,
                   "n": gi["n"], "oriented_edges": m, "wedges": info["wedges"]},
        "roofline": roofline,
        "pass_roofline": pass_roof,
        "cpu_baseline": cpu,
        "kernel_ms": kernel_ms,
        "part_ms": parts,
        "triangles": cp,
        "alg_bytes": alg_bytes,
        "effective_GBs": alg_bytes / (kernel_ms * 1e-3) / 1e9,
        "plain": {"policy": "hybrid", "schedule": "plain", "kernel_ms": plain_ms,
                  "edges_per_s": m / (plain_ms * 1e-3)},
        "speedup_vs_plain": plain_ms / kernel_ms,
        "learned_vs_rule": (rule_ms / kernel_ms) if rule_ms else None,
        "selector": sel_c2,
        "policies": policies,
        "rows_by_type": info["rows_by_type"],
        "configs": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
