"""GPU parity of batched counting: every graph of a batch against the oracle,
in both the persistent (hipGraph) and the per-graph launch modes."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def oracle_graph(seed, g):
    n, m = O.batch_params(seed, g)
    src, dst = O.gen_er(n, m, seed, stream=1 + g)
    r = O.canonical_oriented(n, src, dst)
    return n, r["relabel"][1], r["relabel"][2]


def test_batch_params_match(ctx):
    b = ctx.generate_batch(4, 32)
    n, m = b.params()
    for g in range(32):
        assert (int(n[g]), int(m[g])) == O.batch_params(4, g)


@pytest.mark.parametrize("schedule", ["plain", "tuned"])
def test_batch_counts(ctx, schedule):
    seed, ng = 4, 48
    b = ctx.generate_batch(seed, ng)
    tri_p, k4_p, ms_p = b.count("persistent", schedule)
    tri_g, k4_g, ms_g = b.count("per_graph", schedule)
    tri_p2, k4_p2, _ = b.count("persistent", schedule)   # graph replay
    assert np.array_equal(tri_p, tri_p2) and np.array_equal(k4_p, k4_p2)
    for g in range(ng):
        n, off, nbr = oracle_graph(seed, g)
        gn, goff, gnbr = b.graph(g)
        assert gn == n and np.array_equal(goff, off) and np.array_equal(gnbr, nbr), g
        om = O.Matrix(n, off, nbr, "hybrid")
        t, _ = om.triangles(4)
        k, _ = om.cliques(4, 4)
        assert tri_p[g] == t and tri_g[g] == t, (g, tri_p[g], tri_g[g], t)
        assert k4_p[g] == k and k4_g[g] == k, (g, k4_p[g], k4_g[g], k)
    print(f"batch {ng} graphs {schedule}: persistent {ms_p:.2f} ms, per-graph {ms_g:.2f} ms", flush=True)


def test_batch_pass_options_and_recapture(ctx):
    """Each pass takes its own launch shape; changing either one rebuilds its schedule,
    and the captured graph and the per-graph plans must follow (never replaying freed
    schedule buffers).  Counts stay those of the default shape in every combination."""
    b = ctx.generate_batch(4, 24)
    tri, k4, _ = b.count("persistent", "tuned")
    shapes = (({"small_degree": 64}, {"small_degree": 8}),
              ({"small_degree": 16}, {"small_degree": 64}),
              ({"small_degree": -1}, {"small_degree": -1, "lds_words": 3072, "waves": 2}),
              ({"small_degree": 64}, {"small_degree": 48}),
              ({"small_degree": 64, "lds_words": 3072, "waves": 2}, {"small_degree": 64}))
    for t_o, k_o in shapes:
        for mode in ("persistent", "per_graph", "persistent"):
            t, k, _ = b.count(mode, "tuned", tri_opts=t_o, k4_opts=k_o)
            assert np.array_equal(t, tri) and np.array_equal(k, k4), (mode, t_o, k_o)
    t, k, _ = b.count("persistent", "tuned", cliques=False, tri_opts={"small_degree": 32})
    assert k is None and np.array_equal(t, tri)
    t, k, _ = b.count("persistent", "tuned", triangles=False)
    assert t is None and np.array_equal(k, k4)
    t, k, _ = b.count("persistent", "plain")
    assert np.array_equal(t, tri) and np.array_equal(k, k4)


@pytest.mark.parametrize("ng", [1, 3, 48])
def test_batch_short_queues(ctx, ng):
    """The short rows taken from the 8 per-XCD queues (the default: the row list cut into 8
    ranges of equal wedge work) and from one queue count the same, and both equal the
    oracle's; with 1 or 3 graphs several ranges are empty or hold a few rows."""
    seed = 4
    b = ctx.generate_batch(seed, ng)
    t_q, k_q, _ = b.count("persistent", "tuned")
    t_1, k_1, _ = b.count("persistent", "tuned", short_queues=-1)
    t_q2, k_q2, _ = b.count("persistent", "tuned")  # the graph recaptured with the queues again
    assert np.array_equal(t_q, t_1) and np.array_equal(k_q, k_1)
    assert np.array_equal(t_q, t_q2) and np.array_equal(k_q, k_q2)
    for g in range(ng):
        n, off, nbr = oracle_graph(seed, g)
        om = O.Matrix(n, off, nbr, "hybrid")
        assert t_q[g] == om.triangles(4)[0], g
        assert k_q[g] == om.cliques(4, 4)[0], g


def test_arena_features_are_graph_relative(ctx):
    """An arena row's selector features are those of the same row of its graph alone (ids,
    index and n taken in the graph), feature 15 apart: 0 in the arena, 1 for a single graph."""
    b = ctx.generate_batch(5, 40, policy="uint")
    n_g, _ = b.params()
    vbase = np.concatenate([[0], np.cumsum(n_g.astype(np.int64))])
    F = b.matrix().features()
    assert F.shape == (int(vbase[-1]), 16) and np.all(F[:, 15] == 0)
    ctx.set_selector_schedule(-1)  # a single graph's every use counts, as the arena's do
    try:
        for g in (0, 7, 39):
            n, off, nbr = b.graph(g)
            Fs = ctx.matrix_from_csr(n, off, nbr, "lower", "uint").features()
            Fa = F[vbase[g]:vbase[g + 1]]
            assert np.all(Fs[:, 15] == 1)
            assert np.array_equal(Fa[:, :15], Fs[:, :15]), g
    finally:
        ctx.set_selector_schedule(0)


def test_batch_rebuild(ctx):
    """The arena's layouts rebuilt in place under other policies and with per-row arms: every
    graph's counts are unchanged in both modes (the captured graph and the per-graph plans are
    recorded again over the new matrix), and a typed rebuild gives each row its arm (a bitset
    only within the guard, rows of fewer than two ids uint)."""
    from simdgraphprocessing_amd import selector
    seed, ng = 6, 32
    ctx.set_selector(selector.load())
    b = ctx.generate_batch(seed, ng, policy="hybrid")
    tri, k4, _ = b.count("persistent", "tuned")
    tg, kg, _ = b.count("per_graph", "tuned")
    assert np.array_equal(tri, tg) and np.array_equal(k4, kg)
    for g in range(0, ng, 5):
        n, off, nbr = oracle_graph(seed, g)
        om = O.Matrix(n, off, nbr, "hybrid")
        assert tri[g] == om.triangles(4)[0] and k4[g] == om.cliques(4, 4)[0], g
    M = b.matrix()
    card, _, words = selector.row_stats(M)
    rng = np.random.default_rng(2)
    t = rng.integers(0, 4, len(card)).astype(np.uint8)
    for policy, types in (("uint", None), ("bs", None), ("pshort", None), (None, t), ("learned", None),
                          ("hybrid", None)):
        if types is not None:
            b.rebuild(types=types)
        else:
            b.rebuild(policy)
        for mode in ("persistent", "per_graph"):
            t2, k2, _ = b.count(mode, "tuned")
            assert np.array_equal(t2, tri) and np.array_equal(k2, k4), (policy, mode)
        got = b.matrix().row_types()
        if types is not None:
            want = np.where((types == 0) & (words > selector.GUARD * card), 2, types)
            want = np.where(card < 2, 2, want)
            assert np.array_equal(got, want)
        elif policy == "uint":
            assert np.all(got == 2)


@pytest.mark.parametrize("seed,ng", [(4, 48), (7, 5)])
def test_batch_fused_pass(ctx, seed, ng):
    """At the batch workload's shapes every row of both passes is a short row, and the triangle
    pass runs inside the 4-clique kernel (its local-edge hits are the triangles): the counts of
    the fused launch equal those of each pass run alone and the oracle's, in both modes."""
    from simdgraphprocessing_amd import selector
    sh = selector.batch_shapes()
    b = ctx.generate_batch(seed, ng)
    kw = {"tri_opts": sh["tri"], "k4_opts": sh["k4"]}
    for mode in ("persistent", "per_graph"):
        t, k, _ = b.count(mode, "tuned", **kw)
        t1, _, _ = b.count(mode, "tuned", cliques=False, **kw)
        _, k1, _ = b.count(mode, "tuned", triangles=False, **kw)
        assert np.array_equal(t, t1) and np.array_equal(k, k1), mode
    for g in range(ng):
        n, off, nbr = oracle_graph(seed, g)
        om = O.Matrix(n, off, nbr, "hybrid")
        assert t[g] == om.triangles(4)[0] and k[g] == om.cliques(4, 4)[0], g
