"""Training and evaluation of the learned per-row selector.

Each row that is the B side of a pair has four arms (bitset, pshort, uint, uint searched:
``ARMS``); the tuned triangle pass runs a different algorithm for each (probe/AND, stream,
stream, per-pair search) wherever it reads the row's layout: in the A-side tasks and the
edge-centric short rows.  The B-stationary kernels read canonical lists whatever the layout, and
with the default split and row bound (65,536 and 512) every pair of RMAT-22 and RMAT-24 is
B-stationary, so there no layout changes the pass: the LEARNED policy is schedule-aware
(``Context.set_selector_schedule``), takes as candidates only the rows that some A-side or
edge-centric pair of the schedule reads, and leaves the rest uint (the cheapest build).

The network is trained where the arms matter: the tuned pass with the B-stationary kernels off
(``REGIMES["aside"]``), every pair A-side.  Labels are what each arm costs that pass itself: rows
are grouped by their statistics (length, bitset words per id, uses), and for every group and
arm the engine builds a matrix where only that group takes the arm (the rest uint) and times
its build and its pass against the all-uint matrix, the two passes alternated
(``label_graph``).  The group's pass difference is shared among its rows in proportion to
their stake in the pass (uses x length, the ids streamed when the row is read as uint), the
build difference evenly.  The FP32 MLP 16 -> 64 -> 64 -> 4 learns those costs from the rows'
own features with the engine's own HIP forward / backward / Adam kernels
(``Context.train_selector``), and ``build_matrix(policy="learned")`` gives each candidate row
its cheapest predicted arm.  The reference's density rule (``policy="hybrid"``) stays the
fallback and the yardstick, with more rules as ablations (``RULES``).
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
# the arms the network chooses among (the uint baseline aside): bitset and uint searched.  In the
# tuned pass a pshort B row's pairs stream its canonical list as a uint row's do and only its
# short-row queries differ; its labels sit within the timing noise of uint's, and the network's
# pshort output left its training range on larger graphs, so pshort stays a fixed arm only
LEARNED_ARMS = (0, 3)
# where the arms are timed and compared: the selector schedule the LEARNED build serves
# (Context.set_selector_schedule) and the options of the tuned pass that reads the matrix.
# "tuned": the default pass (B-stationary split 65,536, k_bshl rows up to 512 ids);
# "aside": the B-stationary kernels off, every pair A-side (where the network is trained)
REGIMES = {"tuned": (0, {}), "aside": (-1, {"bstat": -1})}


def set_regime(ctx, regime):
    split, opts = REGIMES[regime]
    ctx.set_selector_schedule(split)
    return opts


def slog(x):
    """Signed log2 of a cost difference in ns per row (the training targets)."""
    x = np.asarray(x, np.float64)
    return (np.sign(x) * np.log2(1.0 + np.abs(x))).astype(np.float32)


def unslog(y):
    y = np.asarray(y, np.float64)
    return np.sign(y) * (2.0 ** np.abs(y) - 1.0)


def row_stats(M):
    """(card, uses, bitset words) of every row of an oriented matrix (uses: the rows that hold
    it, every one of them A-side)."""
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


def tuned_uses(off, nbr, split=65536, bsh=512):
    """The uses the schedule-aware selection counts (k_uses_tuned): per row j, the rows holding
    j at or above the split that also hold some B row there of more than bsh ids."""
    off = np.asarray(off, np.int64)
    card = np.diff(off)
    n = len(card)
    row = np.repeat(np.arange(n, dtype=np.int64), card)
    nbr = np.asarray(nbr, np.int64)
    hi = nbr >= split
    bad = np.zeros(n, bool)
    bad[row[hi & (card[nbr] > bsh)]] = True
    sel = hi & bad[row] & (card[nbr] > 0)
    return np.bincount(nbr[sel], minlength=n)


def groups_of(card, uses, words, min_rows, by_uses=True):
    """Row groups (lists of row ids) by length, bitset words per id and (by_uses) uses; only
    rows of at least two ids that some row holds, and groups of at least min_rows."""
    cand = (card >= 2) & (uses >= 1)
    cb = np.digitize(card, [4, 8, 16, 32, 64, 256])
    db = np.digitize(words / np.maximum(card, 1), [1.5, 8.5, 64.5])
    ub = np.digitize(uses, [2, 8, 64, 512]) if by_uses else np.zeros(len(card), np.int64)
    key = (cb * 4 + db) * 5 + ub
    out = []
    for k in np.unique(key[cand]):
        rows = np.nonzero(cand & (key == k))[0]
        if len(rows) >= min_rows:
            out.append(rows)
    return out


def _pass_ms(M, opts):
    c, st = M.triangle_count("tuned", **opts)
    return c, st["kernel_ms"]


def row_parts(M):
    """16-bit partitions (distinct id >> 16) of every row of a matrix: a pshort row's headers."""
    off, nbr = M.download()
    card = np.diff(off.astype(np.int64))
    row = np.repeat(np.arange(len(card), dtype=np.int64), card)
    hi = np.asarray(nbr, np.int64) >> 16
    new = np.ones(len(hi), np.float64)
    if len(hi) > 1:
        new[1:] = (hi[1:] != hi[:-1]) | (row[1:] != row[:-1])
    return np.bincount(row, weights=new, minlength=len(card)).astype(np.int64)


def significant(dk, base_ms):
    """The median of alternated pass-time differences (ms), or 0 where it is within the timing
    noise: 3 standard errors (from the median absolute deviation) plus 0.1% of the pass."""
    dk = np.asarray(dk, np.float64)
    med = float(np.median(dk))
    sd = 1.4826 * float(np.median(np.abs(dk - med)))
    return med if abs(med) > 3.0 * sd / np.sqrt(len(dk)) + 1e-3 * base_ms else 0.0


def label_graph(ctx, g, reps=4, min_rows=64, per_group=500, seed=0, log=None):
    """Labels of one graph in the A-side regime: (features, targets n x 4, weights n x 4) of its
    labelled rows (a seeded sample of at most per_group rows of every group, so the few heavy
    rows -- the hubs every pass is bound by -- are not swamped by the many light ones).

    Pass: per group and arm, the pass of the matrix where the group takes the arm and the
    all-uint pass are timed alternately, reps times; the median difference, where it stands out
    of the timing noise (``significant``; else 0), is shared among the group's rows by their
    stake, uses x length (the ids streamed when the row is read as uint).
    Build: per arm, the build with every candidate in the arm and the all-uint build are timed
    alternately four times; the median difference is shared among the rows by the payload the
    arm writes for them (bitset: 8 bytes a word; uint searched writes none).  The arms labelled
    are LEARNED_ARMS, with uint the baseline.  A row's label is the signed log of the sum (ns), its weight in the training
    loss 1 + log2(1 + stake) / 4."""
    ctx.set_selector_rule(GUARD, -1)
    opts = set_regime(ctx, "aside")
    n = g.info()["n"]
    base = np.full(n, U, np.uint8)
    Mb = g.build_matrix_typed(base)
    F = Mb.features()
    card, uses, words = row_stats(Mb)
    parts = row_parts(Mb)
    cand = (card >= 2) & (uses >= 1)
    stake = (uses * card).astype(np.float64)
    groups = groups_of(card, uses, words, min_rows)
    c0, k0 = _pass_ms(Mb, opts)  # the schedule, then a warm pass
    bs_ok = words <= GUARD * card
    # build cost of each arm, shared by payload
    bcost = {3: np.zeros(n)}
    payload = {0: 8.0 * words + 8.0, 1: 2.0 * card + 4.0 * parts}
    for arm in (a for a in LEARNED_ARMS if a in payload):
        sel = cand & bs_ok if arm == 0 else cand
        t = base.copy()
        t[sel] = arm
        d = []
        for _ in range(4):
            Mv = g.build_matrix_typed(t)
            bv = Mv.info()["build_ms"]
            del Mv
            M2 = g.build_matrix_typed(base)
            d.append(bv - M2.info()["build_ms"])
            del M2
        tot = max(float(np.median(d)), 0.0)
        bc = np.zeros(n)
        if sel.any():
            bc[sel] = tot * payload[arm][sel] / payload[arm][sel].sum()
        bcost[arm] = bc
    T = np.zeros((n, 4), np.float32)
    Mk = np.zeros((n, 4), np.float32)
    wt = (1.0 + np.log2(1.0 + stake) / 4.0).astype(np.float32)
    for rows in groups:
        Mk[rows, U] = wt[rows]
        for arm in LEARNED_ARMS:
            sel = rows[bs_ok[rows]] if arm == 0 else rows
            if len(sel) == 0:
                continue
            t = base.copy()
            t[sel] = arm
            Mv = g.build_matrix_typed(t)
            c, _ = _pass_ms(Mv, opts)
            assert c == c0
            dk = []
            for _ in range(reps):
                kb = _pass_ms(Mb, opts)[1]
                kv = _pass_ms(Mv, opts)[1]
                dk.append(kv - kb)
            del Mv
            w = stake[sel] / max(stake[sel].sum(), 1.0)
            T[sel, arm] = slog(1e6 * (significant(dk, k0) * w + bcost[arm][sel]))
            Mk[sel, arm] = wt[sel]
    rng = np.random.default_rng(seed)
    keep = np.zeros(n, bool)
    for rows in groups:
        keep[rows if len(rows) <= per_group else rng.choice(rows, per_group, replace=False)] = True
    if log:
        log(f"    {len(groups)} groups, {int(keep.sum())} rows labelled; all-uint pass {k0:.2f} ms; "
            f"build +{bcost[0].sum():.2f} ms all bitset")
    del Mb
    set_regime(ctx, "tuned")
    return F[keep], T[keep], Mk[keep]


def collect(ctx, scales=(18, 19, 20, 21), seeds=range(10, 20), edge_factor=16, reps=4, log=None):
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


# Batch arenas (eh_batch): every arena row held by another is read by the batch passes'
# edge-centric wedge queries (the triangle pass's short rows, k_small_cliques4), where a bitset
# B row answers a query with one probe and a uint row with a binary search, so the arms cost
# differently there than in a single graph's A-side pairs.  The arena's rows carry feature 15 =
# 0 and its graph-relative ids, and are labelled on arenas of their own: ER batches drawn as the
# 4,096-graph workload is (seed 4), from other seeds.
BATCH_SEEDS = (10, 11, 12, 13)
BATCH_GRAPHS = 1024
# the batch comparison adds every row a bitset (no guard), a rule the arena's small graphs can
# afford and single large graphs cannot (its payload passes 32-bit offsets at RMAT-22)
BATCH_RULES = RULES + (("bitset_all", "bs", None, None),)


def batch_shapes():
    """The committed launch shapes of the 4,096-graph workload's two passes."""
    from .autotune import committed
    return committed("c4")


def _batch_pass(b, shapes):
    t, k, ms = b.count("persistent", "tuned", tri_opts=shapes.get("tri"), k4_opts=shapes.get("k4"))
    return np.concatenate([t, k]), ms


def label_batch(ctx, seed, ngraphs=BATCH_GRAPHS, reps=6, min_rows=256, per_group=2000, log=None):
    """Labels of one batch arena (seed, ngraphs ER graphs as eh_batch_generate draws them), timed
    on the arena's persistent triangle + 4-clique passes at the committed launch shapes.

    The arms interact there: the edge-centric kernels answer a wave's 64 wedge queries together,
    and a wave whose B rows mix bitsets and uint lists runs both searches, so an arena with one
    group of rows switched from uint to bitsets runs no faster than all uint (measured), while all
    bitsets save 14%.  Each arm is therefore timed around an operating point, the density rule's
    selection (the yardstick): per group (length x bitset words per id; the uses of an ER arena's
    rows are too alike to split by) and arm, an arena rebuilt with the group in that arm and the
    rest as the rule has them is timed alternately with the rule's arena, reps times, and a row's
    label for an arm is its stake's share of the difference between that arm's and uint's times.
    Build costs are left out: a batch is built once and its captured launch replayed.  Returns
    (features, targets, weights)."""
    shapes = batch_shapes()
    ctx.set_selector_rule(GUARD, -1)
    b0 = ctx.generate_batch(seed, ngraphs, policy="hybrid")
    bv = ctx.generate_batch(seed, ngraphs, policy="hybrid")
    M0 = b0.matrix()
    F = M0.features()
    base = M0.row_types()
    card, uses, words = row_stats(M0)
    del M0
    n = len(card)
    stake = (uses * card).astype(np.float64)
    groups = groups_of(card, uses, words, min_rows, by_uses=False)
    ref, _ = _batch_pass(b0, shapes)  # the schedules and the captured graph, then warm passes
    k0 = float(np.median([_batch_pass(b0, shapes)[1] for _ in range(3)]))
    bs_ok = words <= GUARD * card
    T = np.zeros((n, 4), np.float32)
    Mk = np.zeros((n, 4), np.float32)
    wt = (1.0 + np.log2(1.0 + stake) / 4.0).astype(np.float32)
    for rows in groups:
        Mk[rows, U] = wt[rows]
        d = {}
        for arm in (U,) + LEARNED_ARMS:
            sel = rows[bs_ok[rows]] if arm == 0 else rows
            if len(sel) == 0:
                continue
            t = base.copy()
            t[sel] = arm
            bv.rebuild(types=t)
            c, _ = _batch_pass(bv, shapes)
            assert np.array_equal(c, ref)
            dk = []
            for _ in range(reps):
                kb = _batch_pass(b0, shapes)[1]
                kv = _batch_pass(bv, shapes)[1]
                dk.append(kv - kb)
            d[arm] = significant(dk, k0)
        for arm in LEARNED_ARMS:
            if arm not in d:
                continue
            sel = rows[bs_ok[rows]] if arm == 0 else rows
            w = stake[sel] / max(stake[sel].sum(), 1.0)
            T[sel, arm] = slog(1e6 * (d[arm] - d[U]) * w)
            Mk[sel, arm] = wt[sel]
    rng = np.random.default_rng(seed)
    keep = np.zeros(n, bool)
    for rows in groups:
        keep[rows if len(rows) <= per_group else rng.choice(rows, per_group, replace=False)] = True
    if log:
        log(f"    {len(groups)} groups, {int(keep.sum())} rows labelled; density-rule batch passes {k0:.2f} ms")
    del b0, bv
    return F[keep], T[keep], Mk[keep]


def collect_batch(ctx, seeds=BATCH_SEEDS, ngraphs=BATCH_GRAPHS, log=None):
    """Engine-timed training rows from batch arenas: (features, targets, mask, seed of each row)."""
    xs, ts, ms, gs = [], [], [], []
    for seed in seeds:
        t0 = time.time()
        X, T, Mk = label_batch(ctx, seed, ngraphs, log=log)
        xs.append(X)
        ts.append(T)
        ms.append(Mk)
        gs.append(np.full(len(X), seed, np.int64))
        if log:
            log(f"  labelled batch seed {seed} ({ngraphs} graphs): {len(X)} rows ({time.time() - t0:.1f}s)")
    return np.concatenate(xs), np.concatenate(ts), np.concatenate(ms), np.concatenate(gs)


def compare_batch(ctx, b, steps=5, params=None, log=None):
    """The persistent triangle + 4-clique passes of batch b (committed shapes) under the learned
    layouts and under every rule of BATCH_RULES, each with its build time (the arena rebuilt in
    place; median of 3 builds); per-graph counts must agree.  The batch is left under the
    density rule."""
    shapes = batch_shapes()
    if params is not None:
        ctx.set_selector(params)
    res = {}
    ref = None
    for name, policy, guard, fixed in BATCH_RULES + (("learned", "learned", GUARD, -1),):
        if policy == "learned":
            ctx.set_selector_rule(guard, fixed)
        builds = []
        for _ in range(3):
            b.rebuild(policy)
            builds.append(b.matrix().info()["build_ms"])
        info = b.matrix().info()
        c, _ = _batch_pass(b, shapes)
        ms = []
        for _ in range(steps):
            c2, m = _batch_pass(b, shapes)
            assert np.array_equal(c2, c)
            ms.append(m)
        if ref is None:
            ref = c
        assert np.array_equal(c, ref), name
        res[name] = {"kernel_ms": float(np.median(ms)), "build_ms": float(np.median(builds)),
                     "rows_by_type": info["rows_by_type"]}
        if log:
            log(f"  batch {name}: {res[name]['kernel_ms']:.3f} ms, build {res[name]['build_ms']:.2f} ms, "
                f"{info['rows_by_type']}")
        ctx.set_selector_rule(GUARD, -1)
    b.rebuild("hybrid")
    rules = [r[0] for r in BATCH_RULES]
    best_pass = min(rules, key=lambda k: res[k]["kernel_ms"])
    best_total = min(rules, key=lambda k: res[k]["kernel_ms"] + res[k]["build_ms"])
    L = res["learned"]
    return {"regime": "batch", "shapes": shapes, "policies": res, "best_rule_pass": best_pass,
            "learned_vs_best_rule": res[best_pass]["kernel_ms"] / L["kernel_ms"],
            "best_rule_with_build": best_total,
            "learned_vs_best_rule_with_build": (res[best_total]["kernel_ms"] + res[best_total]["build_ms"]) /
                                               (L["kernel_ms"] + L["build_ms"])}


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


def _timed(M, steps, opts):
    c, _ = M.triangle_count("tuned", **opts)
    ms = []
    for _ in range(steps):
        c2, st = M.triangle_count("tuned", **opts)
        assert c2 == c
        ms.append(st["kernel_ms"])
    return c, float(np.median(ms))


def compare(ctx, g, steps=5, params=None, log=None, regime="tuned"):
    """The tuned triangle pass of graph g in `regime` under the learned layouts and under every
    rule of RULES, each with its build time; counts must agree.  Returns per-policy results and
    the learned policy's gain over the best rule, with and without the build."""
    res = {}
    if params is not None:
        ctx.set_selector(params)
    opts = set_regime(ctx, regime)
    for name, policy, guard, fixed in RULES + (("learned", "learned", GUARD, -1),):
        if policy == "learned":
            ctx.set_selector_rule(guard, fixed)
        builds = []
        for _ in range(3):
            M = g.build_matrix("lower", policy)
            builds.append(M.info()["build_ms"])
        info = M.info()
        info["build_ms"] = float(np.median(builds))
        c, kms = _timed(M, steps, opts)
        res[name] = {"count": c, "kernel_ms": kms, "build_ms": info["build_ms"],
                     "rows_by_type": info["rows_by_type"]}
        if log:
            log(f"  {regime} {name}: {c} in {kms:.3f} ms, build {info['build_ms']:.2f} ms, {info['rows_by_type']}")
        del M
    ctx.set_selector_rule(GUARD, -1)
    set_regime(ctx, "tuned")
    counts = {v["count"] for v in res.values()}
    assert len(counts) == 1, res
    rules = [r[0] for r in RULES]
    best_pass = min(rules, key=lambda k: res[k]["kernel_ms"])
    best_total = min(rules, key=lambda k: res[k]["kernel_ms"] + res[k]["build_ms"])
    L = res["learned"]
    return {"regime": regime, "pass_options": opts, "policies": res,
            "best_rule_pass": best_pass,
            "learned_vs_best_rule": res[best_pass]["kernel_ms"] / L["kernel_ms"],
            "best_rule_with_build": best_total,
            "learned_vs_best_rule_with_build": (res[best_total]["kernel_ms"] + res[best_total]["build_ms"]) /
                                               (L["kernel_ms"] + L["build_ms"])}


def evaluate_graph(ctx, g, steps=5, log=None):
    """compare() on graph g in both regimes, plus the plain schedule under the density rule and
    the pass's algorithmic bytes; counts must agree."""
    out = {r: compare(ctx, g, steps=steps, log=log, regime=r) for r in ("tuned", "aside")}
    M = g.build_matrix("lower", "hybrid")
    c, st = M.triangle_count("plain")
    ms = [M.triangle_count("plain")[1]["kernel_ms"] for _ in range(steps)]
    assert c == out["tuned"]["policies"]["learned"]["count"]
    out["alg_bytes"] = M.triangle_count("tuned")[1]["alg_bytes"]
    out["plain_density_1_256_ms"] = float(np.median(ms))
    out["learned_vs_plain"] = out["plain_density_1_256_ms"] / out["tuned"]["policies"]["learned"]["kernel_ms"]
    return out


def evaluate(ctx, scale=24, seed=24, edge_factor=16, steps=5, log=None):
    """Triangle count on RMAT-scale: the learned layouts against the rules in both regimes, plus
    the plain schedule under the density rule; counts must agree."""
    g = ctx.generate_rmat(scale, edge_factor, seed)
    out = {"workload": f"triangles, RMAT scale {scale} ef {edge_factor} seed {seed}"}
    out.update(evaluate_graph(ctx, g, steps=steps, log=log))
    return out
