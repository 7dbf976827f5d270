"""Edge cases through the C-ABI: edgeless and ragged graphs, tiny cliques under every
layout and schedule, masked filters, duplicate/self-loop edge lists, malformed
inputs rejected by the host or the validation kernel, and the 32-bit payload limit."""
import itertools
from math import comb

import numpy as np
import pytest

import oracle as O
import simdgraphprocessing_amd as eh

pytestmark = pytest.mark.gpu

POLICIES = ["uint", "pshort", "bs", "hybrid", "hybrid3", "cost"]


def csr_of_edges(n, edges):
    if not edges:
        return np.zeros(n + 1, np.uint64), np.zeros(0, np.uint32)
    src = np.array([a for a, _ in edges]) if edges else np.zeros(0)
    dst = np.array([b for _, b in edges]) if edges else np.zeros(0)
    return O.edges_to_csr(n, src, dst)


def counts(M, schedule):
    return [M.triangle_count(schedule)[0]] + [M.clique_count(k, schedule)[0] for k in (4, 5)]


@pytest.mark.parametrize("n", [1, 5])
def test_edgeless_graphs(ctx, n):
    off, nbr = csr_of_edges(n, [])
    for pol in POLICIES:
        M = ctx.matrix_from_csr(n, off, nbr, "lower", pol)
        assert M.info()["cardinality"] == 0
        for sched in ("plain", "tuned"):
            assert counts(M, sched) == [0, 0, 0], (pol, sched)


def test_small_cliques_every_layout(ctx):
    # K_3, K_4 and K_6 on disjoint vertex sets inside 200 mostly isolated vertices (ragged rows)
    edges = []
    for part in (range(0, 3), range(10, 14), range(100, 106)):
        edges += list(itertools.combinations(part, 2))
    off, nbr = csr_of_edges(200, edges)
    exp = [1 + 4 + comb(6, 3), 1 + comb(6, 4), comb(6, 5)]
    for pol in POLICIES:
        M = ctx.matrix_from_csr(200, off, nbr, "lower", pol)
        for sched in ("plain", "tuned"):
            assert counts(M, sched) == exp, (pol, sched)


def test_mask_filter_matches_orientation(ctx, facebook):
    n, off, nbr = facebook["n"], facebook["off"], facebook["nbr"]
    rows = np.repeat(np.arange(n), np.diff(off).astype(np.int64))
    mask = (nbr < rows).astype(np.uint8)
    Mm = ctx.matrix_from_csr(n, off, nbr, "mask", "hybrid", mask=mask)
    assert Mm.triangle_count("plain")[0] == 1612010
    assert Mm.clique_count(4, "plain")[0] == 30004668


def test_edge_list_with_duplicates_and_loops(ctx):
    rng = np.random.default_rng(3)
    src = rng.integers(0, 300, 5000).astype(np.uint32)
    dst = rng.integers(0, 300, 5000).astype(np.uint32)
    src[:100] = dst[:100]  # self-loops
    src = np.concatenate([src, dst[:500]])  # reversed duplicates
    dst = np.concatenate([dst, src[:500]])
    g = ctx.graph_from_edges(300, src, dst, relabel=True)
    r = O.canonical_oriented(300, src, dst)
    off, nbr = r["sym"]
    M = g.build_matrix("lower", "hybrid")
    assert M.triangle_count("tuned")[0] == O.brute_triangles(300, off, nbr)
    assert M.clique_count(4, "tuned")[0] == O.brute_cliques(300, off, nbr, 4)


def test_malformed_inputs_rejected(ctx):
    good_off = np.array([0, 1, 2], np.uint64)
    with pytest.raises(eh.EHError):  # no vertices
        ctx.graph_from_csr(0, np.zeros(1, np.uint64), np.zeros(0, np.uint32))
    with pytest.raises(eh.EHError):  # offsets not monotone (host check)
        ctx.graph_from_csr(2, np.array([0, 2, 1], np.uint64), np.array([1, 0], np.uint32))
    with pytest.raises(eh.EHError):  # id out of range (validation kernel)
        ctx.graph_from_csr(2, good_off, np.array([1, 7], np.uint32))
    with pytest.raises(eh.EHError):  # self-loop
        ctx.graph_from_csr(2, good_off, np.array([0, 0], np.uint32))
    with pytest.raises(eh.EHError):  # unsorted row
        ctx.graph_from_csr(3, np.array([0, 2, 2, 2], np.uint64), np.array([2, 1], np.uint32))
    off, nbr = csr_of_edges(4, [(0, 1), (1, 2)])
    M = ctx.matrix_from_csr(4, off, nbr, "lower", "hybrid")
    with pytest.raises(eh.EHError):  # row pair outside the matrix (host check)
        M.intersect_rows([0, 9], [1, 2])
    with pytest.raises(eh.EHError):  # query_depth outside 3..16
        M.clique_count(2, "tuned")


def test_all_bitset_layout_limit_is_clean(ctx):
    """Every row as a bitset at RMAT-22 needs more than 2^32 payload words: the build
    reports it instead of truncating offsets; smaller graphs build and count."""
    g = ctx.generate_rmat(16, 16, 2)
    M = g.build_matrix("lower", "bs")
    off, nbr = g.download()
    assert M.triangle_count("tuned")[0] == O.Matrix(1 << 16, off, nbr, "bs").triangles(8)[0]
    del M, g
    g = ctx.generate_rmat(22, 16, 2)
    with pytest.raises(eh.EHError):
        g.build_matrix("lower", "bs")
