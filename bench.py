"""Benchmark: triangle counting on RMAT scale 22 (edge factor 16, a/b/c/d =
.57/.19/.19/.05, seed 2), degree-oriented, uint32 ids, hybrid layouts; counts
bit-exact against the CPU restatement of the reference timed in the same run.

One step = one counting pass (eh_triangle_count) over the device-resident
matrix.  Prints one JSON line (rank 0, N=1).

    python bench.py [--steps K] [--warmup W] [--scale S] [--policy P]
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

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(n, 16))


def traffic_from_profiles(config_tag):
    """HBM bytes per pass measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE),
    recorded under profiles/ for this exact configuration; None otherwise."""
    path = os.path.join(ROOT, "profiles", "tc_rmat22_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        entry = d.get("configs", {}).get(config_tag)
        if entry:
            return entry.get("bytes_per_pass")
    except (OSError, ValueError):
        pass
    return None


def _stride_for(est_full_s, budget_s):
    """Power-of-two row stride that brings an estimated full CPU pass under the budget."""
    stride = 1
    while est_full_s / stride > budget_s and stride < 64:
        stride *= 2
    return stride


def cpu_baseline(n, off, nbr, gpu_count, budget_s=30.0):
    """The oracle (C restatement of the reference hybrid path) on this host."""
    import oracle as O
    threads = cpu_threads()
    M = O.Matrix(n, off, nbr, "hybrid", lower=True, threads=threads)
    # probe with a 1/64 vertex sample, then size the measured sample to the budget
    _, e_probe, t_probe = M.triangles_sample(threads, 64, 0)
    stride = _stride_for(t_probe * 64, budget_s)
    if stride == 1:
        c, secs = M.triangles(threads)
        edges = M.cardinality()
        sample = "full graph"
        assert c == gpu_count, f"CPU count {c} != GPU count {gpu_count}"
        exact = True
    else:
        c, edges, secs = M.triangles_sample(threads, stride, 0)
        sample = f"source vertices i with i % {stride} == 0 ({edges} of {M.cardinality()} oriented edges)"
        exact = False
    return {"value": edges / secs, "unit": "edges/s", "cores": threads, "kind": "port",
            "sample": sample, "seconds": secs, "cpu_model": cpu_model(), "count_checked": exact,
            "oracle": "oracle/eh_oracle.c hybrid policy, dynamic chunks of 100 (common.hpp:213-276)"}


def _tuned(M, k, budget_s=20):
    from simdgraphprocessing_amd import autotune
    params, best_ms, _ = autotune.tune(M, k, "tuned", repeat=2, budget_s=budget_s, use_cache=True)
    return params


def config_c1(ctx, with_cpu):
    """ER G(10000, 100000 draws), seed 1: triangles."""
    g = ctx.generate_er(10000, 100000, 1)
    M = g.build_matrix("lower", "hybrid")
    c, stp = M.triangle_count("plain", repeat=5)
    c2, st = M.triangle_count("tuned", repeat=5)
    assert c == c2
    m = M.info()["cardinality"]
    out = {"workload": "triangles, ER n=10000 m=100000 draws seed 1", "triangles": c, "oriented_edges": m,
           "kernel_ms": st["kernel_ms"], "edges_per_s": m / (st["kernel_ms"] * 1e-3),
           "plain_kernel_ms": stp["kernel_ms"]}
    if with_cpu:
        import oracle as O
        off, nbr = g.download()
        om = O.Matrix(10000, off, nbr, "hybrid", threads=cpu_threads())
        oc, secs = om.triangles(cpu_threads())
        assert oc == c
        out["cpu"] = {"seconds": secs, "edges_per_s": m / secs, "cores": cpu_threads(), "kind": "port"}
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
    # edges traversed: every entry of every reached row is scanned once
    out = {"workload": f"BFS path length, RMAT scale {scale} ef {ef} seed {seed}, all entries, start = largest row",
           "path_length": r["path_length"], "rounds": r["rounds"], "reached": r["reached"],
           "level_sizes": r["level_sizes"], "kernel_ms": r["kernel_ms"], "bottom_up_rounds": r["bottom_up"],
           "top_down_kernel_ms": M.bfs_path(start)["kernel_ms"], "entries": i["cardinality"],
           "teps": i["cardinality"] / (r["kernel_ms"] * 1e-3)}
    if with_cpu:
        import oracle as O
        off, nbr = g.download()
        t0 = time.perf_counter()
        ref = O.bfs_path(i["n"], off, nbr, start, cap=256)
        secs = time.perf_counter() - t0
        assert ref["path_length"] == r["path_length"] and ref["level_sizes"] == r["level_sizes"]
        out["cpu"] = {"seconds": secs, "teps": i["cardinality"] / secs, "cores": 1, "kind": "port"}
    return out


def config_c3(ctx, with_cpu):
    """Chung-Lu n=1e6, gamma 2.1, mean degree 32 + 200 planted 256-vertex communities (p=.5), seed 3: 4-cliques."""
    g = ctx.generate_chung_lu(1000000, 32, 2.1, 3, communities=200, community_size=256)
    M = g.build_matrix("lower", "hybrid")
    c, stp = M.clique_count(4, "plain", repeat=2)
    params = _tuned(M, 4)
    c2, st = M.clique_count(4, "tuned", repeat=3, **params)
    assert c == c2
    i = M.info()
    # B_4C = B_TC + 4 * sum over triangles (i>j>k) of d+(k): the weighted pass is not exposed, so the
    # triangle part is reported and the clique work is given as time and count
    out = {"workload": "4-cliques, Chung-Lu n=1e6 gamma 2.1 mean degree 32 + 200x256 communities p=.5, seed 3",
           "four_cliques": c, "oriented_edges": i["cardinality"], "kernel_ms": st["kernel_ms"],
           "plain_kernel_ms": stp["kernel_ms"], "speedup_vs_plain": stp["kernel_ms"] / st["kernel_ms"],
           "tuned_params": params}
    if with_cpu:
        import oracle as O
        off, nbr = g.download()
        om = O.Matrix(i["n"], off, nbr, "hybrid", threads=cpu_threads())
        threads = cpu_threads()
        _, _, t_probe = om.cliques_sample(4, threads, 64, 0)
        stride = _stride_for(t_probe * 64, 30.0)
        if stride == 1:
            kc, secs = om.cliques(4, threads)
            assert kc == c, f"CPU 4-clique count {kc} != GPU {c}"
            edges, sample = i["cardinality"], "full graph"
        else:
            _, edges, secs = om.cliques_sample(4, threads, stride, 0)
            sample = f"source rows i % {stride} == 0"
        out["cpu"] = {"seconds": secs, "sample": sample, "edges_per_s": edges / secs,
                      "cores": threads, "kind": "port", "count_checked": stride == 1}
    return out


def config_c4(ctx, with_cpu):
    """4096 ER graphs (n in [64,8192], mean degree in [4,64)), seed 4: per-graph triangles + 4-cliques."""
    import numpy as np
    from simdgraphprocessing_amd import autotune
    b = ctx.generate_batch(4, 4096)
    info = b.info()
    # each pass gets its own launch shape from the autotuner (timed alone, persistent mode)
    params, pass_ms = autotune.tune_batch(b, "persistent", budget_s=30, use_cache=True)
    kw = {"tri_opts": params["tri"], "k4_opts": params["k4"]}
    tri, k4, _ = b.count("persistent", "tuned", **kw)
    ms_p = []
    for _ in range(3):
        t2, k2, ms = b.count("persistent", "tuned", **kw)
        assert np.array_equal(t2, tri) and np.array_equal(k2, k4)
        ms_p.append(ms)
    # per-graph launches at the better of their default shape and the persistent one
    per_graph = {}
    shapes = (("default_shape", {}), ("persistent_shape", kw),
              ("short_rows_8_for_4cliques", {"k4_opts": {"small_degree": 8}}))
    for name, opts in shapes:
        tg, kg, ms_g = b.count("per_graph", "tuned", **opts)
        assert np.array_equal(tg, tri) and np.array_equal(kg, k4)
        per_graph[name] = ms_g
    ms_g = min(per_graph.values())
    out = {"workload": "4096 ER graphs n~U[64,8192] mean degree~U[4,64) seed 4, triangles + 4-cliques per graph",
           "graphs": info["graphs"], "vertices": info["vertices"], "edges": info["edges"],
           "persistent_ms": min(ms_p), "per_graph_ms": ms_g, "per_graph_by_shape_ms": per_graph,
           "speedup_vs_per_graph": ms_g / min(ms_p), "tuned_params": params, "pass_ms": pass_ms,
           "triangles_total": int(tri.sum()), "four_cliques_total": int(k4.sum())}
    if with_cpu:
        # all 8192 counts against the oracle in one pass over the union graph
        import oracle as O
        import time as _t
        tot_n = info["vertices"]
        n, m = b.params()
        vbase = np.concatenate([[0], np.cumsum(n.astype(np.int64))])
        off = np.zeros(tot_n + 1, np.uint64)
        chunks = []
        for g in range(info["graphs"]):
            gn, goff, gnbr = b.graph(g)
            off[vbase[g] + 1: vbase[g + 1] + 1] = goff[1:] + off[vbase[g]]
            chunks.append(gnbr.astype(np.uint32) + np.uint32(vbase[g]))
        nbr = np.concatenate(chunks)
        om = O.Matrix(tot_n, off, nbr, "hybrid", threads=cpu_threads())
        t0 = _t.perf_counter()
        _, rt = om.count_rows(3, cpu_threads())
        _, rk = om.count_rows(4, cpu_threads())
        secs = _t.perf_counter() - t0
        gid = np.repeat(np.arange(info["graphs"]), n.astype(np.int64))
        ct = np.bincount(gid, weights=rt.astype(np.float64), minlength=info["graphs"]).astype(np.uint64)
        ck = np.bincount(gid, weights=rk.astype(np.float64), minlength=info["graphs"]).astype(np.uint64)
        assert np.array_equal(ct, tri) and np.array_equal(ck, k4), "batch counts differ from the oracle"
        out["cpu"] = {"seconds": secs, "cores": cpu_threads(), "kind": "port", "counts_checked": 2 * info["graphs"]}
    return out


def config_c5(ctx, with_cpu):
    """Learned selector (trained on engine timings of RMAT 18-21, seeds 10-19) evaluated on RMAT-24,
    edge factor 16, seed 24: triangles with density-rule vs learned layouts."""
    from simdgraphprocessing_amd import selector
    out = selector.evaluate(ctx, scale=24, seed=24, steps=3)
    if with_cpu:
        import oracle as O
        g = ctx.generate_rmat(24, 16, 24)
        off, nbr = g.download()
        del g
        out["cpu"] = cpu_baseline(1 << 24, off, nbr, out["tuned_learned"]["count"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--policy", default="auto", help="layouts of the timed pass: auto = learned when the "
                    "trained selector parameters exist, else hybrid")
    ap.add_argument("--configs", default="c1,c3,c4,c5,bfs", help="extra configs to report (c1,c3,c4,c5,bfs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--params", default="", help='fixed launch shape instead of autotuning, e.g. '
                    '\'{"small_degree": 32, "lds_words": 1536, "waves": 4}\'')
    ap.add_argument("--plain-steps", type=int, default=5)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the counting path has "
              "no multi-GPU mode.", file=sys.stderr)
        sys.exit(2)

    import simdgraphprocessing_amd as eh
    from simdgraphprocessing_amd import autotune, selector

    ctx = eh.Context(0)
    if args.policy in ("auto", "learned"):
        if os.path.exists(selector.PARAMS_PATH):
            ctx.set_selector(selector.load())
            args.policy = "learned"
        elif args.policy == "auto":
            args.policy = "hybrid"
    log(f"device: {ctx.device_name}")
    t0 = time.time()
    g = ctx.generate_rmat(args.scale, args.edge_factor, args.seed)
    gi = g.info()
    log(f"generated RMAT-{args.scale}: n={gi['n']} entries={gi['entries']} max_deg={gi['max_degree']} "
        f"({time.time() - t0:.1f}s)")

    # plain: reference layout rule + reference scheduling (one task per row, vertex order)
    Mp = g.build_matrix("lower", "hybrid")
    cp, stp = Mp.triangle_count("plain")
    plain_ms = []
    for _ in range(args.plain_steps):
        c, st = Mp.triangle_count("plain")
        assert c == cp
        plain_ms.append(st["kernel_ms"])
    plain_ms = sorted(plain_ms)[len(plain_ms) // 2] if plain_ms else stp["kernel_ms"]
    log(f"plain hybrid: count={cp} kernel {plain_ms:.3f} ms")

    # candidate layouts on the tuned schedule: the density rule (the selector's
    # yardstick) and, when trained parameters exist, the learned selector; each gets
    # its launch shape from the autotuner and the faster one is the timed pass
    cands = {}
    for pol in (["hybrid", "learned"] if args.policy == "learned" else [args.policy]):
        Mc = g.build_matrix("lower", pol)
        cc, stc = Mc.triangle_count("tuned", repeat=5)
        assert cc == cp, f"{pol}: tuned count {cc} != plain count {cp}"
        prm, best = {}, stc["kernel_ms"]
        if args.params:
            prm = json.loads(args.params)
            best = Mc.triangle_count("tuned", repeat=5, **prm)[1]["kernel_ms"]
        elif not args.no_tune:
            prm, best, _ = autotune.tune(Mc, 3, "tuned", repeat=2, budget_s=45, use_cache=True)
            best = min(best, Mc.triangle_count("tuned", repeat=5, **prm)[1]["kernel_ms"])
        cands[pol] = {"matrix": Mc, "params": prm, "default_ms": stc["kernel_ms"], "tuned_ms": best}
        log(f"tuned {pol}: default shape {stc['kernel_ms']:.3f} ms, autotuned {prm} {best:.3f} ms")
    chosen = min(cands, key=lambda k: cands[k]["tuned_ms"])
    M = cands[chosen]["matrix"]
    params = cands[chosen]["params"]
    info = M.info()
    kw = dict(params) if params else {}
    for _ in range(args.warmup):
        c, st = M.triangle_count("tuned", **kw)
        assert c == cp, f"tuned count {c} != plain count {cp}"
    kms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        c, st = M.triangle_count("tuned", **kw)
        kms.append(st["kernel_ms"])
    t_end = time.perf_counter()
    assert c == cp
    ms_step = (t_end - t_start) * 1e3 / args.steps
    kernel_ms = sum(kms) / len(kms)
    m = info["cardinality"]
    alg_bytes = st["alg_bytes"]
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    config_tag = f"rmat{args.scale}-ef{args.edge_factor}-s{args.seed}-{chosen}-tuned"
    traffic = traffic_from_profiles(config_tag)

    cpu = None
    if not args.no_cpu:
        off, nbr = g.download()
        cpu = cpu_baseline(gi["n"], off, nbr, cp)
        log(f"cpu baseline: {cpu['value']:.3e} edges/s on {cpu['cores']} threads ({cpu['sample']})")
        del off, nbr
    del M, Mp, g
    policies = {k: {"default_shape_ms": v["default_ms"], "autotuned_ms": v["tuned_ms"], "params": v["params"]}
                for k, v in cands.items()}
    cands.clear()

    extra = {}
    wanted = [c for c in args.configs.split(",") if c]
    if "c1" in wanted:
        extra["c1"] = config_c1(ctx, not args.no_cpu)
    if "c3" in wanted:
        extra["c3"] = config_c3(ctx, not args.no_cpu)
    if "c4" in wanted:
        extra["c4"] = config_c4(ctx, not args.no_cpu)
    if "bfs" in wanted:
        extra["bfs"] = config_bfs(ctx, not args.no_cpu, args.scale, args.edge_factor, args.seed)
    if "c5" in wanted and args.policy == "learned":
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
        "config": {"workload": f"triangle count, RMAT scale {args.scale}, edge factor {args.edge_factor}, "
                               f"a/b/c/d .57/.19/.19/.05, seed {args.seed}, degree-oriented",
                   "layout_policy": chosen, "schedule": "tuned", "tuned_params": params,
                   "n": gi["n"], "oriented_edges": m, "wedges": info["wedges"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "kernel_ms": kernel_ms,
        "triangles": cp,
        "plain": {"policy": "hybrid", "schedule": "plain", "kernel_ms": plain_ms,
                  "edges_per_s": m / (plain_ms * 1e-3)},
        "speedup_vs_plain": plain_ms / kernel_ms,
        "policies": policies,
        "alg_bytes": alg_bytes,
        "rows_by_type": info["rows_by_type"],
        "build_ms": info["build_ms"],
        "configs": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
