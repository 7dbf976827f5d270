"""GPU parity: triangle and 4-clique counts through libeh.so against the oracle
and the reference's own fixture values (test/undirected_triangle_counting_test.cpp:11,
test/n_clique_test.cpp:11)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

FB_TRIANGLES = 1612010
FB_4CLIQUES = 30004668
POLICIES = ["uint", "pshort", "bs", "hybrid", "hybrid3", "cost"]


@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("schedule", ["plain", "tuned"])
def test_facebook_triangles(ctx, facebook, policy, schedule):
    M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "lower", policy)
    c, st = M.triangle_count(schedule)
    print(f"facebook {policy}/{schedule}: {c} in {st['kernel_ms']:.3f} ms", flush=True)
    assert c == FB_TRIANGLES


@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("schedule", ["plain", "tuned"])
def test_facebook_4cliques(ctx, facebook, policy, schedule):
    M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "lower", policy)
    c, st = M.clique_count(4, schedule)
    print(f"facebook 4-clique {policy}/{schedule}: {c} in {st['kernel_ms']:.3f} ms", flush=True)
    assert c == FB_4CLIQUES


def test_facebook_ref_row_bytes(ctx, facebook):
    # the reference prints "ROW DATA SIZE (Bytes): 569724" for the hybrid build
    M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "lower", "hybrid")
    om = O.Matrix(facebook["n"], facebook["off"], facebook["nbr"], "hybrid")
    assert M.info()["ref_row_bytes"] == om.bytes() == 569724
    assert M.info()["rows_by_type"]["bitset"] == om.type_counts()["bitset"]


def test_er_exact_generator(ctx):
    """G(n, m) with exactly m distinct pairs: the device picks the oracle's pairs; a complete
    graph (m = n (n-1) / 2) gives C(n, 3) triangles and C(n, 4) 4-cliques."""
    for n, m, seed in ((10, 45, 2), (300, 2000, 5), (5000, 60000, 9)):
        g = ctx.generate_er_exact(n, m, seed)
        off, nbr = g.download()
        assert len(nbr) == 2 * m
        ref = O.canonical_oriented(n, *O.gen_er_exact(n, m, seed))
        assert np.array_equal(off, ref["relabel"][1]) and np.array_equal(nbr, ref["relabel"][2])
    M = ctx.generate_er_exact(10, 45, 3).build_matrix("lower", "hybrid")
    assert M.triangle_count("tuned")[0] == 120 and M.clique_count(4, "tuned")[0] == 210


def test_er_c1_parity(ctx):
    """C1: ER G(10000, 100000 distinct non-loop pairs), seed 1: device generator + build + count
    vs oracle."""
    g = ctx.generate_er_exact(10000, 100000, 1)
    off, nbr = g.download()
    assert len(nbr) == 200000
    src, dst = O.gen_er_exact(10000, 100000, 1)
    ref = O.canonical_oriented(10000, src, dst)
    roff, rnbr = ref["relabel"][1], ref["relabel"][2]
    assert np.array_equal(off, roff) and np.array_equal(nbr, rnbr)
    M = g.build_matrix("lower", "hybrid")
    om = O.Matrix(10000, roff, rnbr, "hybrid")
    exp, _ = om.triangles(8)
    for sched in ("plain", "tuned"):
        c, st = M.triangle_count(sched)
        assert c == exp, sched
        k4, _ = M.clique_count(4, sched)
        assert k4 == om.cliques(4, 8)[0]


@pytest.mark.parametrize("scale", [12, 16])
def test_rmat_parity(ctx, scale):
    g = ctx.generate_rmat(scale, 16, 2)
    off, nbr = g.download()
    src, dst = O.gen_rmat(scale, 16, 2)
    ref = O.canonical_oriented(1 << scale, src, dst)
    assert np.array_equal(off, ref["relabel"][1]) and np.array_equal(nbr, ref["relabel"][2])
    om = O.Matrix(1 << scale, off, nbr, "hybrid")
    exp, _ = om.triangles(8)
    for policy in ("hybrid", "hybrid3", "cost", "uint", "bs", "pshort"):
        M = g.build_matrix("lower", policy)
        for sched in ("plain", "tuned"):
            c, _ = M.triangle_count(sched)
            assert c == exp, (policy, sched)


@pytest.mark.parametrize("small", [-1, 2, 3, 8, 16, 32, 48, 64])
def test_small_row_kernel(ctx, facebook, small):
    """The edge-centric path for short rows must not change any count."""
    M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "lower", "hybrid")
    assert M.triangle_count("tuned", small_degree=small)[0] == FB_TRIANGLES
    assert M.clique_count(4, "tuned", small_degree=small)[0] == FB_4CLIQUES
    g = ctx.generate_rmat(14, 16, 5)
    off, nbr = g.download()
    om = O.Matrix(1 << 14, off, nbr, "hybrid")
    for pol in ("hybrid", "cost", "hybrid3"):
        M = g.build_matrix("lower", pol)
        assert M.triangle_count("tuned", small_degree=small)[0] == om.triangles(8)[0], pol
        assert M.clique_count(4, "tuned", small_degree=small)[0] == om.cliques(4, 8)[0], pol


@pytest.mark.parametrize("split", [-1, 64, 100, 1024, 4096, 524288])
def test_b_stationary_pairs(ctx, facebook, split):
    """Pairs whose B row id is below the split are counted B-stationary (row(j) as an LDS
    bitmap, the prefixes of the rows holding j streamed through it); any split, with or
    without the edge-centric short rows, and under every layout, gives the same count."""
    M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "lower", "hybrid")
    for small in (-1, 16, 48):
        assert M.triangle_count("tuned", small_degree=small, bstat=split)[0] == FB_TRIANGLES, small
    # B-stationary applies to triangles only: the 4-clique pass ignores the option
    assert M.clique_count(4, "tuned", bstat=split)[0] == FB_4CLIQUES
    g = ctx.generate_rmat(14, 16, 5)
    off, nbr = g.download()
    exp = O.Matrix(1 << 14, off, nbr, "hybrid").triangles(8)[0]
    for pol in ("hybrid", "cost", "uint", "pshort", "bs", "bitset_new"):
        M = g.build_matrix("lower", pol)
        for small in (-1, 32):
            assert M.triangle_count("tuned", small_degree=small, bstat=split)[0] == exp, (pol, small)
    # unoriented matrices keep every pair A-side (the prefix identity needs row(j) below j)
    M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "all", "hybrid")
    assert M.triangle_count("tuned", bstat=split)[0] == 6 * FB_TRIANGLES


@pytest.mark.parametrize("local_max", [-1, 64, 128, 200, 512])
def test_local_graph_cliques(ctx, facebook, local_max):
    """4-cliques of rows above small_degree counted as triangles of each row's local graph
    (LDS bit-rows built from streamed B rows, one wave per row up to 128 entries, one block
    per row up to 512): any bound and threshold gives the oracle's count."""
    M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "lower", "hybrid")
    for small in (-1, 8, 32):
        assert M.clique_count(4, "tuned", local_max=local_max, small_degree=small)[0] == FB_4CLIQUES, small
    # a dense graph: rows of 100-300 entries and many local edges per row
    g = ctx.generate_er(1500, 150000, 7)
    off, nbr = g.download()
    exp = O.Matrix(1500, off, nbr, "hybrid").cliques(4, 8)[0]
    for pol in ("hybrid", "uint", "pshort"):
        M = g.build_matrix("lower", pol)
        for small in (-1, 16):
            assert M.clique_count(4, "tuned", local_max=local_max, small_degree=small)[0] == exp, (pol, small)
    # triangle passes ignore the option
    assert M.triangle_count("tuned", local_max=local_max)[0] == O.Matrix(1500, off, nbr, "hybrid").triangles(8)[0]


def test_unoriented_matrix(ctx, facebook):
    """EH_FILTER_ALL keeps both directions: every triangle is counted 6 times
    (sum over ordered (i,j) of |N(i) & N(j)|), and the small-row path stays off."""
    M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "all", "hybrid")
    for sched in ("plain", "tuned"):
        assert M.triangle_count(sched)[0] == 6 * FB_TRIANGLES


# ------------------------------------------------------------------ deeper cliques (query_depth > 4)

@pytest.mark.parametrize("k", [5, 6])
def test_facebook_deep_cliques(ctx, facebook, k):
    """n_clique with --n=5, 6 (apps/n_clique.cpp:52-101): the general level-set kernel
    against the oracle's recursion."""
    om = O.Matrix(facebook["n"], facebook["off"], facebook["nbr"], "hybrid")
    exp, _ = om.cliques(k, 8)
    for policy in ("hybrid", "pshort"):
        M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "lower", policy)
        c, st = M.clique_count(k, "tuned")
        print(f"facebook {k}-cliques {policy}: {c} in {st['kernel_ms']:.1f} ms", flush=True)
        assert c == exp


def test_deep_cliques_closed_forms(ctx):
    """A hub joined to 5,000 vertices, ten of which form K_10: the k-cliques are those
    of K_11, C(11, k); the hub's row (5,000 entries) forces the HBM spill of the levels."""
    from math import comb
    n = 5001
    hub = n - 1
    src = [hub] * (n - 1) + [a for a in range(10) for b in range(a + 1, 10)]
    dst = list(range(n - 1)) + [b for a in range(10) for b in range(a + 1, 10)]
    off, nbr = O.edges_to_csr(n, np.array(src), np.array(dst))
    M = ctx.matrix_from_csr(n, off, nbr, "lower", "hybrid")
    for k in (5, 6, 8, 11, 12):
        assert M.clique_count(k, "tuned")[0] == comb(11, k), k


def test_deep_cliques_unoriented_and_random(ctx):
    """EH_FILTER_ALL counts ordered cliques (k! each); random RMAT graphs match the oracle."""
    from math import factorial
    src, dst = O.gen_rmat(11, 12, 5)
    r = O.canonical_oriented(2048, src, dst)
    _, off, nbr = r["relabel"]
    om = O.Matrix(2048, off, nbr, "hybrid")
    M = ctx.matrix_from_csr(2048, off, nbr, "lower", "hybrid")
    Ma = ctx.matrix_from_csr(2048, off, nbr, "all", "hybrid")
    for k in (5, 6):
        exp, _ = om.cliques(k, 8)
        assert M.clique_count(k, "tuned")[0] == exp
        assert Ma.clique_count(k, "tuned")[0] == factorial(k) * exp


def test_deep_cliques_reject_large_k(ctx, facebook):
    import simdgraphprocessing_amd as eh
    M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "lower", "hybrid")
    with pytest.raises(eh.EHError):
        M.clique_count(17, "tuned")


@pytest.mark.parametrize("split", [64, 1024])
def test_b_stationary_long_b_rows(ctx, split):
    """A dense ER graph (n 2,500, 500,000 edges): oriented row j holds about j / 6 ids, so the B
    rows above the split hold up to ~400, and the k_bshl bound (B rows up to 256 ids in four
    register rounds, up to 512 in eight) decides which rows go wholly B-stationary; every bound
    gives the oracle's count."""
    n = 2500
    g = ctx.generate_er_exact(n, 500000, 7)
    off, nbr = g.download()
    exp = O.Matrix(n, off, nbr, "hybrid").triangles(8)[0]
    ooff, _ = O.orient_lower(n, off, nbr)
    card = np.diff(ooff.astype(np.int64))
    assert card[split:].max() > 256 and card.max() <= 512
    for pol in ("hybrid", "uint"):
        M = g.build_matrix("lower", pol)
        for bsh in (0, 128, 256, 300, 512, -1):
            for small in (-1, 32):
                got = M.triangle_count("tuned", bstat=split, bsh=bsh, small_degree=small)[0]
                assert got == exp, (pol, bsh, small)


def test_complete_graph_full_remainders(ctx):
    """K_600: every B row j of the oriented matrix has the in-pairs (i, j), i > j, each with a
    prefix of j ids, so B rows 497-503 give 64 pairs of 63 remainder vectors per pass (the most
    slots a flattened pass holds: 4,032 in 63 blocks of 64) and the hub-free dense rows exercise
    the A-side and local-graph paths too; the counts are C(600, 3) and C(600, 4)."""
    from math import comb
    n = 600
    src = [a for a in range(n) for b in range(a + 1, n)]
    dst = [b for a in range(n) for b in range(a + 1, n)]
    off, nbr = O.edges_to_csr(n, np.array(src), np.array(dst))
    for pol in ("hybrid", "uint", "bs"):
        M = ctx.matrix_from_csr(n, off, nbr, "lower", pol)
        for bstat in (0, -1, 256):
            assert M.triangle_count("tuned", bstat=bstat)[0] == comb(n, 3), (pol, bstat)
        # split 256: the rows up to 512 hand their pairs above the split to k_bshl, whose B rows
        # then hold 256-511 ids (row(j) in eight register rounds); a bound of 256 or 300 leaves
        # more of them A-side
        for bsh in (256, 300, 511, 512, -1):
            assert M.triangle_count("tuned", bstat=256, bsh=bsh)[0] == comb(n, 3), (pol, bsh)
        assert M.triangle_count("plain")[0] == comb(n, 3), pol
        for local in (0, -1):
            assert M.clique_count(4, "tuned", local_max=local)[0] == comb(n, 4), (pol, local)


def _oriented_bytes(n, off, nbr):
    """B_TC and B_4C of an oriented CSR by sparse algebra: with A[i, j] = 1 for kept (i, j),
    T = (A A) o A counts for every kept (i, x) the j with i > j > x closing a triangle, so the
    4-clique term is 4 sum over (i, x) of T[i, x] card(x)."""
    import scipy.sparse as sp
    off = np.asarray(off, np.int64)
    card = np.diff(off)
    rows = np.repeat(np.arange(n), card)
    A = sp.csr_matrix((np.ones(len(nbr), np.int64), (rows, np.asarray(nbr, np.int64))), shape=(n, n))
    m, W = len(nbr), int(card[np.asarray(nbr, np.int64)].sum())
    b_tc = 4 * (m + W) + 8 * (n + 1)
    T = (A @ A).multiply(A).tocoo()
    tw = int((T.data * card[T.col]).sum())
    return b_tc, b_tc + 4 * tw, int(T.data.sum())


def test_alg_bytes_and_kernel_class_times(ctx, facebook):
    """eh_run_stats.alg_bytes is B_TC for triangles and B_4C for 4-cliques; with time_parts the
    per-class device times cover the pass and the count is unchanged."""
    import simdgraphprocessing_amd as eh
    n = facebook["n"]
    M = ctx.matrix_from_csr(n, facebook["off"], facebook["nbr"], "lower", "hybrid")
    off, nbr = M.download()
    b_tc, b_4c, tri = _oriented_bytes(n, off, nbr)
    assert tri == FB_TRIANGLES
    for sched in ("plain", "tuned"):
        t, st = M.triangle_count(sched, time_parts=True, repeat=3)
        assert t == FB_TRIANGLES and st["alg_bytes"] == b_tc
        parts = st["part_ms"]
        assert parts and set(parts) <= set(eh.PARTS), parts
        assert sum(parts.values()) <= st["kernel_ms"] * 1.02 + 0.02, (parts, st["kernel_ms"])
        k, st4 = M.clique_count(4, sched, time_parts=True)
        assert k == FB_4CLIQUES and st4["alg_bytes"] == b_4c
    assert set(M.triangle_count("tuned", time_parts=True)[1]["part_ms"]) >= {"k_bstat"}
    assert M.triangle_count("tuned")[1]["part_ms"] == {}


def test_batch_alg_bytes(ctx):
    """The batch's byte counts are the per-graph B_TC and B_4C summed over its graphs."""
    b = ctx.generate_batch(7, 6)
    tri_b, k4_b = b.alg_bytes()
    exp_t = exp_k = 0
    for g in range(b.info()["graphs"]):
        n, off, nbr = b.graph(g)  # graph g's rows, relabelled inside the graph, local ids
        ooff, onbr = O.orient_lower(n, off, nbr)
        t, k, _ = _oriented_bytes(n, ooff, onbr)
        exp_t += t
        exp_k += k
    assert (tri_b, k4_b) == (exp_t, exp_k)
