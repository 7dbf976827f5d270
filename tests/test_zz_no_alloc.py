"""eh_intersect_rows_static -- the form of eh_intersect_rows that allocates nothing (its
pairs and counts are staged in the library's static device array) -- must give exactly
eh_intersect_rows's counts, call after call, with a later call smaller than the one before
it (the array keeps the earlier call's contents), and must reject more pairs than the array
holds.  Both are checked against numpy's set intersection of the matrix's own rows."""
import ctypes as C

import numpy as np
import pytest

import simdgraphprocessing_amd as eh

pytestmark = pytest.mark.gpu


def _rows(M):
    off, nbr = M.download()
    return off, nbr


def _numpy_counts(off, nbr, a, b):
    return np.array([np.intersect1d(nbr[off[x]:off[x + 1]], nbr[off[y]:off[y + 1]], assume_unique=True).size
                     for x, y in zip(a, b)], np.uint64)


def test_static_matches_allocating_form_on_two_calls(ctx, facebook):
    M = ctx.graph_from_csr(facebook["n"], facebook["off"], facebook["nbr"]).build_matrix("all", "hybrid")
    n = facebook["n"]
    rng = np.random.default_rng(11)
    off, nbr = _rows(M)
    # first call: 60,000 pairs (every row with itself and its neighbours, plus random pairs)
    a1 = rng.integers(0, n, 60000, dtype=np.uint32)
    b1 = rng.integers(0, n, 60000, dtype=np.uint32)
    b1[:n] = np.arange(n, dtype=np.uint32)
    a1[:n] = np.arange(n, dtype=np.uint32)
    got1 = M.intersect_rows_static(a1, b1)
    ref1 = M.intersect_rows(a1, b1)
    assert np.array_equal(got1, ref1)
    assert np.array_equal(got1[:n], np.diff(off).astype(np.uint64))  # |row & row| = |row|
    sample = rng.choice(60000, 2000, replace=False)
    assert np.array_equal(got1[sample], _numpy_counts(off, nbr, a1[sample], b1[sample]))
    # second call, smaller: the array still holds the first call's pairs and counts past it
    a2 = rng.integers(0, n, 7001, dtype=np.uint32)
    b2 = rng.integers(0, n, 7001, dtype=np.uint32)
    got2 = M.intersect_rows_static(a2, b2)
    ref2 = M.intersect_rows(a2, b2)
    assert got2.shape == (7001,)
    assert np.array_equal(got2, ref2)
    assert np.array_equal(got2[:1500], _numpy_counts(off, nbr, a2[:1500], b2[:1500]))
    # a third call of one pair, and an empty call
    assert M.intersect_rows_static([0], [0])[0] == off[1] - off[0]
    assert M.intersect_rows_static([], []).shape == (0,)


def test_static_rejects_more_pairs_than_the_array_holds(ctx):
    M = ctx.generate_er(1000, 5000, 3).build_matrix("lower", "hybrid")
    k = eh.STATIC_MAX_PAIRS + 1
    a = np.zeros(k, np.uint32)
    with pytest.raises(eh.EHError) as e:
        M.intersect_rows_static(a, a)
    assert e.value.status == 6  # EH_ERR_CAPACITY
    assert "static staging" in str(e.value)
    # a row outside the matrix is rejected before any device work
    with pytest.raises(eh.EHError) as e:
        M.intersect_rows_static(np.array([1000], np.uint32), np.array([0], np.uint32))
    assert e.value.status == 1
    # the maximum itself is accepted
    a = np.arange(eh.STATIC_MAX_PAIRS, dtype=np.uint32) % 1000
    got = M.intersect_rows_static(a, a[::-1].copy())
    assert np.array_equal(got, M.intersect_rows(a, a[::-1].copy()))
    # the library's declared maximum and the binding's agree
    assert eh.STATIC_MAX_PAIRS == 1 << 20
    assert isinstance(eh.lib().eh_intersect_rows_static, C._CFuncPtr)
