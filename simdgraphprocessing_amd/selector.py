"""Training and evaluation of the learned per-row selector.

Each row that is the B side of a pair has four arms (bitset, pshort, uint, uint searched:
``ARMS``); the tuned triangle pass runs a different algorithm for each (probe/AND, stream,
stream, per-pair search).  The labels are what each arm costs the tuned pass itself: rows are
grouped by their statistics (length, bitset words per id, how many rows hold them), and for
every group and arm the engine builds a matrix where only that group takes the arm (the rest
uint) and times its build and its tuned pass against the all-uint matrix (``label_graph``).
The difference, per row of the group, is the label of each row in it.  The FP32 MLP
16 -> 64 -> 64 -> 4 learns those costs from the rows' own features with the engine's own HIP
forward / backward / Adam kernels (``Context.train_selector``), and
``build_matrix(policy="learned")`` gives each row its cheapest predicted arm.  The reference's
density rule (``policy="hybrid"``) stays the fallback and the yardstick, with more rules as
ablations (``RULES``).

Layouts act only on A-side work (the B-stationary kernels read canonical lists whatever the
layout), and on the small training graphs nearly every pair runs B-stationary, so the labels
time the tuned pass with the B-stationary kernels off: every pair is A-side work, short source
rows through k_small_rows and the others through the task kernel.
"""
import os
import time

import numpy as np

PARAMS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_params.npy")
ARMS = ("bitset", "pshort", "uint", "uint-searched")
GUARD = 8  # the learned policy's bitset guard: at most 8 words per id (16x the uint row's bytes)
# the rules the network is measured against: (name, layout policy, bitset guard, fixed arm); the
# bitset rule is the learned policy's guard with every row a bitset
RULES = (("density_1_256", "hybrid", None, None),
         ("bitset_1_512", "learned", GUARD, 0),
         ("bitset_1_4096", "learned", 64, 0),
         ("uint", "uint", None, None))
U = 2


def slog(x):
    """Signed log2 of a cost difference in ns per row (the training targets)."""
    x = np.asarray(x, np.float64)
    return (np.sign(x) * np.log2(1.0 + np.abs(x))).astype(np.float32)


def unslog(y):
    y = np.asarray(y, np.float64)
    return np.sign(y) * (2.0 ** np.abs(y) - 1.0)


def row_stats(M):
    """(card, uses, bitset words) of every row of an oriented matrix."""
    off, nbr = M.download()
    off = off.astype(np.int64)
    card = np.diff(off)
    uses = np.bincount(nbr, minlength=len(card)) if len(nbr) else np.zeros(len(card), np.int64)
    words = np.zeros(len(card), np.int64)
    nz = card > 0
    first = nbr[off[:-1][nz]].astype(np.int64)
    last = nbr[off[1:][nz] - 1].astype(np.int64)
    words[nz] = (last >> 6) - (first >> 6) + 1
    return card, uses, words


def groups_of(card, uses, words, min_rows):
    """Row groups (lists of row ids) by length, bitset words per id and uses; only rows of at
    least two ids that some row holds, and groups of at least min_rows."""
    cand = (card >= 2) & (uses >= 1)
    cb = np.digitize(card, [4, 8, 16, 32, 64, 256])
    db = np.digitize(words / np.maximum(card, 1), [1.5, 8.5, 64.5])
    ub = np.digitize(uses, [2, 8])
    key = (cb * 4 + db) * 3 + ub
    out = []
    for k in np.unique(key[cand]):
        rows = np.nonzero(cand & (key == k))[0]
        if len(rows) >= min_rows:
            out.append(rows)
    return out


def _pass(g, types, split, bsh, reps):
    M = g.build_matrix_typed(types)
    b = M.info()["build_ms"]
    c, _ = M.triangle_count("tuned", bstat=split, bsh=bsh)
    ms = []
    for _ in range(reps):
        c2, st = M.triangle_count("tuned", bstat=split, bsh=bsh)
        assert c2 == c
        ms.append(st["kernel_ms"])
    del M
    return c, b, float(np.median(ms))


def label_graph(ctx, g, reps=3, min_frac=0.001, max_rows=40000, seed=0, log=None):
    """Labels of one graph: (features, targets n x 4, mask n x 4) of its labelled rows (a seeded
    sample of at most max_rows of them: a group's rows share their labels)."""
    ctx.set_selector_rule(GUARD, -1)
    n = g.info()["n"]
    split, bsh = -1, 0  # the B-stationary kernels off: every pair is A-side work
    M0 = g.build_matrix("lower", "uint")
    F = M0.features()
    card, uses, words = row_stats(M0)
    del M0
    cand = (card >= 2) & (uses >= 1)
    groups = groups_of(card, uses, words, max(64, int(min_frac * cand.sum())))
    base = np.full(n, U, np.uint8)
    c0, b0s, k0s = [], [], []
    for _ in range(3):
        c, b, k = _pass(g, base, split, bsh, reps)
        c0.append(c)
        b0s.append(b)
        k0s.append(k)
    b0, k0 = float(np.median(b0s)), float(np.median(k0s))
    T = np.zeros((n, 4), np.float32)
    Mk = np.zeros((n, 4), np.float32)
    bs_ok = words <= GUARD * card
    for rows in groups:
        Mk[rows, U] = 1.0
        for arm in (0, 1, 3):
            sel = rows[bs_ok[rows]] if arm == 0 else rows
            if len(sel) == 0:
                continue
            t = base.copy()
            t[sel] = arm
            c, b, k = _pass(g, t, split, bsh, reps)
            assert c == c0[0]
            T[sel, arm] = slog(1e6 * ((k - k0) + (b - b0)) / len(sel))
            Mk[sel, arm] = 1.0
    keep = Mk.max(axis=1) > 0
    idx = np.nonzero(keep)[0]
    if max_rows and len(idx) > max_rows:
        idx = np.sort(np.random.default_rng(seed).choice(idx, max_rows, replace=False))
    keep = np.zeros(n, bool)
    keep[idx] = True
    if log:
        log(f"    {len(groups)} groups, {int(keep.sum())} rows labelled; base build {b0:.2f} ms, pass {k0:.2f} ms")
    return F[keep], T[keep], Mk[keep]


def collect(ctx, scales=(18, 19, 20, 21), seeds=range(10, 20), edge_factor=16, reps=3, log=None):
    """Engine-timed training rows from RMAT graphs: (features, targets, mask, seed of each row)."""
    xs, ts, ms, gs = [], [], [], []
    for scale in scales:
        for seed in seeds:
            t0 = time.time()
            g = ctx.generate_rmat(scale, edge_factor, seed)
            X, T, Mk = label_graph(ctx, g, reps=reps, seed=1000 * scale + seed, log=log)
            xs.append(X)
            ts.append(T)
            ms.append(Mk)
            gs.append(np.full(len(X), seed, np.int64))
            if log:
                log(f"  labelled RMAT-{scale} seed {seed}: {len(X)} rows ({time.time() - t0:.1f}s)")
            del g
    return np.concatenate(xs), np.concatenate(ts), np.concatenate(ms), np.concatenate(gs)


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
    """Measured pass + build ns of a per-row arm choice over labelled rows, against all uint
    (sum); a choice without a label (a bitset past the guard) costs what uint costs, as the
    selection would fall back."""
    rows = np.arange(len(choice))
    c = np.where(Mk[rows, choice] > 0, choice, U)
    return float(np.sum(unslog(T[rows, c]))), int(len(choice))


def _card_words(F):
    card = 2.0 ** (F[:, 0] * 16.0) - 1.0
    span = card * (2.0 ** (F[:, 8] * 32.0) - 1.0)
    words = card * (2.0 ** (F[:, 6] * 16.0) - 1.0)
    return card, span, words


def heuristic_choice(F):
    """The reference rule as arms: bitset when density > 1/256 and more than one element."""
    card, span, _ = _card_words(F)
    dens = np.where(span > 0, card / np.maximum(span, 1.0), 0.0)
    return np.where((card > 1.5) & (span > 0) & (dens > 1.0 / 256.0), 0, 2).astype(np.int64)


def bitset_choice(F, guard):
    """A bitset wherever its words stay within `guard` per id, else uint."""
    card, _, words = _card_words(F)
    return np.where((card > 1.5) & (words <= guard * card + 0.5), 0, 2).astype(np.int64)


def _timed(M, steps):
    c, _ = M.triangle_count("tuned")
    ms = []
    for _ in range(steps):
        c2, st = M.triangle_count("tuned")
        assert c2 == c
        ms.append(st["kernel_ms"])
    return c, float(np.median(ms))


def compare(ctx, g, steps=5, params=None, log=None):
    """The tuned triangle pass of graph g under the learned layouts and under every rule of
    RULES, each with its build time; counts must agree.  Returns per-policy results and the
    learned policy's gain over the best rule, with and without the build."""
    res = {}
    if params is not None:
        ctx.set_selector(params)
    for name, policy, guard, fixed in RULES + (("learned", "learned", GUARD, -1),):
        if policy == "learned":
            ctx.set_selector_rule(guard, fixed)
        builds = []
        for _ in range(3):
            M = g.build_matrix("lower", policy)
            builds.append(M.info()["build_ms"])
        info = M.info()
        info["build_ms"] = float(np.median(builds))
        c, kms = _timed(M, steps)
        res[name] = {"count": c, "kernel_ms": kms, "build_ms": info["build_ms"],
                     "rows_by_type": info["rows_by_type"]}
        if log:
            log(f"  {name}: {c} in {kms:.3f} ms, build {info['build_ms']:.2f} ms, {info['rows_by_type']}")
        del M
    ctx.set_selector_rule(GUARD, -1)
    counts = {v["count"] for v in res.values()}
    assert len(counts) == 1, res
    rules = [r[0] for r in RULES]
    best_pass = min(rules, key=lambda k: res[k]["kernel_ms"])
    best_total = min(rules, key=lambda k: res[k]["kernel_ms"] + res[k]["build_ms"])
    L = res["learned"]
    return {"policies": res,
            "best_rule_pass": best_pass,
            "learned_vs_best_rule": res[best_pass]["kernel_ms"] / L["kernel_ms"],
            "best_rule_with_build": best_total,
            "learned_vs_best_rule_with_build": (res[best_total]["kernel_ms"] + res[best_total]["build_ms"]) /
                                               (L["kernel_ms"] + L["build_ms"])}


def evaluate_graph(ctx, g, steps=5, log=None):
    """compare() on graph g, plus the plain schedule under the density rule and the pass's
    algorithmic bytes; counts must agree."""
    out = compare(ctx, g, steps=steps, log=log)
    M = g.build_matrix("lower", "hybrid")
    c, st = M.triangle_count("plain")
    ms = [M.triangle_count("plain")[1]["kernel_ms"] for _ in range(steps)]
    assert c == out["policies"]["learned"]["count"]
    out["alg_bytes"] = M.triangle_count("tuned")[1]["alg_bytes"]
    out["plain_density_1_256_ms"] = float(np.median(ms))
    out["learned_vs_plain"] = out["plain_density_1_256_ms"] / out["policies"]["learned"]["kernel_ms"]
    return out


def evaluate(ctx, scale=24, seed=24, edge_factor=16, steps=5, log=None):
    """Triangle count on RMAT-scale: the learned layouts against the rules, all on the tuned
    schedule, plus the plain schedule under the density rule; counts must agree."""
    g = ctx.generate_rmat(scale, edge_factor, seed)
    out = {"workload": f"triangles, RMAT scale {scale} ef {edge_factor} seed {seed}"}
    out.update(evaluate_graph(ctx, g, steps=steps, log=log))
    return out
