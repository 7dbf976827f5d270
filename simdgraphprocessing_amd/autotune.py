"""Schedule autotuner for the counting kernels.

Searches the launch/schedule parameters of ``eh_triangle_count`` /
``eh_clique_count`` (task split size, per-wave LDS budget, waves per block) by
timing the real kernel on the real matrix, and caches the winner per
(device, matrix signature, k).  Counts never depend on these parameters; the
tuner asserts that every candidate returns the same count.
"""
import hashlib
import json
import os
import sys
import time

CACHE = os.path.join(os.path.expanduser("~"), ".cache", "simdgraphprocessing_amd", "autotune.json")
# launch shapes picked by this tuner for the benchmark workloads, committed with the package so
# that a run and its profile use one shape (tags: c1, c2:<policy>, c3, c4, c5:<policy>)
SHAPES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shapes.json")


def committed(tag):
    """The committed launch shape of a benchmark workload ({} = the engine's defaults)."""
    with open(SHAPES) as f:
        return json.load(f).get(tag, {})


def _signature(M, k, schedule):
    i = M.info()
    key = f"{i['n']}:{i['cardinality']}:{i['wedges']}:{i['rows_by_type']}:{i['payload_bytes']}:{k}:{schedule}"
    return hashlib.sha1(key.encode()).hexdigest()[:16]


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1)
        os.replace(tmp, CACHE)
    except OSError:
        pass


def candidates(M, k=3):
    """Launch/schedule grid: edge-centric row threshold x LDS budget x waves per block, and for
    triangles the B-stationary split (B rows below 64Ki / 48Ki / 32Ki ids, or off).  The
    likelier winners come first, so a tight time budget still reaches them."""
    smalls = (32, 48, 16, 24, 64) if k == 3 else (8, 16, 32, 48, 64)
    splits = (65536, 49152, 32768, -1) if k == 3 else (-1,)
    for bstat in splits:
        for small in smalls:
            for lds, waves in ((512, 4), (768, 4), (1024, 4), (1536, 4), (2048, 4), (2048, 2), (3072, 2)):
                p = {"small_degree": small, "lds_words": lds, "waves": waves}
                if k == 3:
                    p["bstat"] = bstat
                yield p


def tune(M, k=3, schedule="tuned", repeat=3, budget_s=60.0, use_cache=True, verbose=False):
    """Return (best params dict, best ms, table of trials)."""
    dev = M.ctx.device_name
    sig = f"{dev}|{_signature(M, k, schedule)}"
    db = _load() if use_cache else {}
    if use_cache and sig in db:
        return db[sig]["params"], db[sig]["ms"], []
    run = (lambda **kw: M.triangle_count(schedule, **kw)) if k == 3 else \
        (lambda **kw: M.clique_count(k, schedule, **kw))
    ref, _ = run()
    trials = []
    t0 = time.time()
    best = None
    for p in candidates(M, k):
        if time.time() - t0 > budget_s:
            break
        c, st = run(repeat=repeat, **p)
        if c != ref:
            raise AssertionError(f"autotune candidate {p} changed the count: {c} != {ref}")
        trials.append((st["kernel_ms"], p))
        if verbose:
            print(f"  autotune {p} -> {st['kernel_ms']:.3f} ms", flush=True)
        if best is None or st["kernel_ms"] < best[0]:
            best = (st["kernel_ms"], p)
    db[sig] = {"params": best[1], "ms": best[0], "time": time.time()}
    if use_cache:
        _store(db)
    return best[1], best[0], trials


def _batch_signature(b, k, schedule):
    i = b.info()
    n, m = b.params()
    key = f"batch:{i['graphs']}:{i['vertices']}:{i['edges']}:{int(n.sum())}:{int(m.sum())}:{k}:{schedule}"
    return hashlib.sha1(key.encode()).hexdigest()[:16]


def batch_candidates(k):
    """Grid for one pass of a batch: edge-centric row threshold x LDS budget x waves."""
    smalls = (32, 48, 64, 96) if k == 3 else (16, 32, 48, 64)
    for small in smalls:
        for lds, waves in ((2048, 2), (2048, 4), (1536, 4), (3072, 2)):
            yield {"small_degree": small, "lds_words": lds, "waves": waves}


def tune_batch(b, mode="persistent", schedule="tuned", repeat=2, budget_s=30.0, use_cache=True, verbose=False):
    """Tune the triangle and the 4-clique pass of a batch separately (each timed alone in
    `mode`); return ({"tri": params, "k4": params}, {"tri": ms, "k4": ms})."""
    dev = b.ctx.device_name
    db = _load() if use_cache else {}
    params, best_ms = {}, {}
    for name, k in (("tri", 3), ("k4", 4)):
        sig = f"{dev}|{mode}|{_batch_signature(b, k, schedule)}"
        if use_cache and sig in db:
            params[name], best_ms[name] = db[sig]["params"], db[sig]["ms"]
            continue
        want = dict(triangles=(k == 3), cliques=(k == 4))

        def run(p):
            opt = {"tri_opts": p} if k == 3 else {"k4_opts": p}
            return b.count(mode, schedule, **want, **opt)

        ref = run({})
        ref = ref[0] if k == 3 else ref[1]
        best = None
        t0 = time.time()
        for p in batch_candidates(k):
            if time.time() - t0 > budget_s:
                break
            ms = []
            for _ in range(repeat):
                r = run(p)
                got = r[0] if k == 3 else r[1]
                if not (got == ref).all():
                    raise AssertionError(f"batch autotune candidate {p} changed the counts")
                ms.append(r[2])
            if verbose:
                print(f"  batch autotune k={k} {p} -> {min(ms):.3f} ms", flush=True)
            if best is None or min(ms) < best[0]:
                best = (min(ms), p)
        params[name], best_ms[name] = best[1], best[0]
        db[sig] = {"params": best[1], "ms": best[0], "time": time.time()}
    if use_cache:
        _store(db)
    return params, best_ms


def refresh(out_path, repeat=4, budget_s=90.0, log=print):
    """Tune every benchmark workload afresh and write the shapes table (same tags as
    shapes.json) to out_path."""
    import simdgraphprocessing_amd as eh
    from . import workloads as W
    ctx = eh.Context(0)
    has_sel = W.load_selector(ctx)
    shapes = {}
    for cfg, k, policies in (("c1", 3, ("hybrid",)), ("c2", 3, ("hybrid", "learned")), ("c3", 4, ("hybrid",)),
                             ("c5", 3, ("hybrid", "learned"))):
        for pol in policies:
            if pol == "learned" and not has_sel:
                continue
            _, M = W.build(ctx, cfg, pol)
            p, ms, _ = tune(M, k, "tuned", repeat=repeat, budget_s=budget_s, use_cache=False)
            shapes[W.shape_tag(cfg, pol)] = p
            log(f"{W.shape_tag(cfg, pol)}: {p} {ms:.3f} ms")
            del M
    _, b = W.build(ctx, "c4", "hybrid")
    p, ms = tune_batch(b, "persistent", repeat=2, budget_s=budget_s, use_cache=False)
    shapes["c4"] = p
    log(f"c4: {p} {ms}")
    with open(out_path, "w") as f:
        json.dump(shapes, f, indent=1)


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from simdgraphprocessing_amd import autotune as _self
    _self.refresh(sys.argv[1] if len(sys.argv) > 1 else SHAPES)
