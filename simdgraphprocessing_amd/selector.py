"""Training and evaluation of the learned per-row selector.

The engine times each of its four membership arms (bitset probe, pshort walk,
uint binary search, uint galloping) on the real queries of sampled rows
(``SparseMatrix.profile_layouts``); the FP32 MLP 16 -> 64 -> 64 -> 4 is trained on
those timings with the engine's own HIP forward/backward/Adam kernels
(``Context.train_selector``), and ``build_matrix(policy="learned")`` then gives
each row its cheapest predicted arm.  The reference's density rule
(``policy="hybrid"``) stays the fallback and the yardstick.
"""
import os
import time

import numpy as np

PARAMS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_params.npy")
ARMS = ("bitset", "pshort", "uint", "uint-gallop")


def collect(ctx, scales=(18, 19, 20, 21), seeds=range(10, 20), edge_factor=16, target_rows=20000, log=None):
    """Engine-timed training rows from RMAT graphs: (features, targets, mask)."""
    xs, ts, ms = [], [], []
    for scale in scales:
        for seed in seeds:
            t0 = time.time()
            g = ctx.generate_rmat(scale, edge_factor, seed)
            M = g.build_matrix("lower", "hybrid")
            n = M.info()["n"]
            stride = max(1, n // target_rows)
            F = M.features()
            T, Mk = M.profile_layouts(stride)
            keep = Mk.max(axis=1) > 0
            xs.append(F[keep])
            ts.append(T[keep])
            ms.append(Mk[keep])
            if log:
                log(f"  profiled RMAT-{scale} seed {seed}: {int(keep.sum())} rows ({time.time() - t0:.1f}s)")
            del M, g
    return np.concatenate(xs), np.concatenate(ts), np.concatenate(ms)


def train(ctx, X, T, Mk, epochs=60, batch=512, lr=2e-3, seed=0):
    loss = ctx.train_selector(X, T, Mk, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return ctx.get_selector(), loss


def save(params, path=PARAMS_PATH):
    np.save(path, np.asarray(params, np.float32), allow_pickle=False)


def load(path=PARAMS_PATH):
    p = np.load(path, allow_pickle=False)
    assert p.shape == (5508,) and p.dtype == np.float32
    return p


def choice_cost(T, Mk, choice):
    """Measured cycles per query of a per-row arm choice over labelled rows (sum)."""
    rows = np.arange(len(choice))
    ok = Mk[rows, choice] > 0
    return float(np.sum(2.0 ** T[rows[ok], choice[ok]])), int(ok.sum())


def heuristic_choice(F):
    """The reference rule as arms: bitset when density > 1/256 and more than one element."""
    card = 2.0 ** (F[:, 0] * 16.0) - 1.0
    span = 2.0 ** (F[:, 1] * 32.0) - 1.0
    dens = np.where(span > 0, card / np.maximum(span, 1.0), 0.0)
    return np.where((card > 1.5) & (span > 0) & (dens > 1.0 / 256.0), 0, 2).astype(np.int64)


def evaluate(ctx, scale=24, seed=24, edge_factor=16, steps=5, log=None):
    """Triangle count on RMAT-scale: heuristic vs learned layouts, both on the tuned
    schedule, plus the plain schedule; counts must agree."""
    g = ctx.generate_rmat(scale, edge_factor, seed)
    out = {"workload": f"triangles, RMAT scale {scale} ef {edge_factor} seed {seed}"}
    res = {}
    for name, policy, sched in (("plain_hybrid", "hybrid", "plain"), ("tuned_hybrid", "hybrid", "tuned"),
                                ("tuned_learned", "learned", "tuned")):
        M = g.build_matrix("lower", policy)
        c, st = M.triangle_count(sched)
        ms = []
        for _ in range(steps):
            c2, st = M.triangle_count(sched)
            assert c2 == c
            ms.append(st["kernel_ms"])
        res[name] = {"count": c, "kernel_ms": float(np.median(ms)), "rows_by_type": M.info()["rows_by_type"]}
        if log:
            log(f"  {name}: {c} in {res[name]['kernel_ms']:.3f} ms")
        del M
    counts = {v["count"] for v in res.values()}
    assert len(counts) == 1, res
    out.update(res)
    out["learned_vs_heuristic"] = res["tuned_hybrid"]["kernel_ms"] / res["tuned_learned"]["kernel_ms"]
    out["learned_vs_plain"] = res["plain_hybrid"]["kernel_ms"] / res["tuned_learned"]["kernel_ms"]
    return out
