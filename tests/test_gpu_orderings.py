"""GPU parity of the vertex orderings (eh_graph_reorder; MutableGraph::reorder_*,
src/MutableGraph.hpp:126-285, as used by tools/undirectedEdgeListToBinary.cpp): the
device permutation and the renumbered CSR must equal the oracle's exactly."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

NAMES = list(O.ORDERINGS)


def several_components(seed):
    n = 3000
    src, dst = O.gen_er(1500, 4000, seed)
    edges = list(zip(src.tolist(), dst.tolist()))
    edges += [(1600 + i, 1601 + i) for i in range(200)]            # a long path
    edges += [(2000, 2001 + i) for i in range(300)]                 # a star
    edges += [(2400 + 3 * i, 2401 + 3 * i) for i in range(100)]    # many small components
    edges = [(a, b) for a, b in edges if a != b]                    # and isolated vertices
    off, nbr = O.edges_to_csr(n, np.array([a for a, _ in edges]), np.array([b for _, b in edges]))
    return n, off, nbr


def check(ctx, n, off, nbr, name, seed):
    g = ctx.graph_from_csr(n, off, nbr)
    g.reorder(name, seed)
    want = O.order(n, off, nbr, name, seed)
    got = g.perm()
    assert np.array_equal(got, want), (name, seed, np.nonzero(got != want)[0][:5])
    goff, gnbr = g.download()
    woff, wnbr = O.apply_order(n, off, nbr, want)
    assert np.array_equal(goff, woff) and np.array_equal(gnbr, wnbr), name


@pytest.mark.parametrize("name", NAMES)
def test_orderings_several_components(ctx, name):
    for seed in (1, 7):
        n, off, nbr = several_components(seed)
        check(ctx, n, off, nbr, name, seed)


@pytest.mark.parametrize("name", NAMES)
def test_orderings_facebook(ctx, facebook, name):
    check(ctx, facebook["n"], facebook["off"], facebook["nbr"], name, 3)


@pytest.mark.parametrize("name", ["bfs", "strong_run", "shingles"])
def test_orderings_rmat(ctx, name):
    src, dst = O.gen_rmat(14, 8, 5)
    r = O.canonical_oriented(1 << 14, src, dst)
    off, nbr = r["sym"]
    check(ctx, 1 << 14, off, nbr, name, 11)


def test_reordered_graph_counts(ctx, facebook):
    """The triangle count does not depend on the numbering (the lower orientation is valid for any)."""
    for name in NAMES:
        g = ctx.graph_from_csr(facebook["n"], facebook["off"], facebook["nbr"])
        g.reorder(name, 5)
        M = g.build_matrix("lower", "hybrid")
        assert M.triangle_count("tuned")[0] == 1612010, name
