"""Per-row layout parity on the device: every row of a built matrix, as eh_matrix_row returns
it, has the type, cardinality and bytes of the oracle's image for that policy -- uinteger
(uinteger.hpp:49-64), pshort (pshort.hpp:46-100), bitset (bitset.hpp:65-118), the hybrid
density rule and its three-way band (hybrid.hpp:55-100,132-201).  For the byte-cost and learned
selectors, which the reference does not have, each row's image must be the oracle's encoding of
that row in the type the device chose."""
import numpy as np
import pytest

import oracle as O
from simdgraphprocessing_amd import selector

pytestmark = pytest.mark.gpu

TYPE_KIND = {O.BITSET: "bitset", O.PSHORT: "pshort", O.UINTEGER: "uint", O.BITSET_NEW: "bitset_new",
             O.NEW_TYPE: "new_type"}


def _rmat(scale, ef, seed):
    src, dst = O.gen_rmat(scale, ef, seed)
    _, off, nbr = O.canonical_oriented(1 << scale, src, dst)["relabel"]
    return 1 << scale, off, nbr


def _graphs(facebook):
    return (("facebook", facebook["n"], facebook["off"], facebook["nbr"]), ("rmat12",) + _rmat(12, 8, 3))


@pytest.mark.parametrize("policy", ["uint", "pshort", "bs", "hybrid", "hybrid3"])
def test_row_images_equal_oracle(ctx, facebook, policy):
    for name, n, off, nbr in _graphs(facebook):
        M = ctx.matrix_from_csr(n, off, nbr, "lower", policy)
        om = O.Matrix(n, off, nbr, policy)
        assert M.info()["ref_row_bytes"] == om.bytes()
        assert M.info()["rows_by_type"] == om.type_counts()
        for i in range(n):
            t, c, p = M.row(i)
            ot, oc, op = om.row(i)
            assert c == oc, (name, policy, i)
            if c:
                assert t == ot and np.array_equal(p, op), (name, policy, i, t, ot, len(p), len(op))


@pytest.mark.parametrize("policy", ["cost", "learned"])
def test_selected_row_images_are_exact_encodings(ctx, facebook, policy):
    import simdgraphprocessing_amd as eh
    c = eh.Context(0)
    if policy == "learned":
        c.set_selector(selector.load())
        # every pair A-side: every row held by another is a candidate (under the default schedule
        # these graphs, all below the B-stationary split, would leave every row uint)
        c.set_selector_schedule(-1)
    seen = set()
    for name, n, off, nbr in _graphs(facebook):
        M = c.matrix_from_csr(n, off, nbr, "lower", policy)
        moff, mnbr = M.download()
        for i in range(n):
            t, card, p = M.row(i)
            row = mnbr[moff[i]:moff[i + 1]]
            assert card == len(row), (name, i)
            if not card:
                continue
            seen.add(t)
            if t == O.UINTEGER:
                assert np.array_equal(p.view(np.uint32), row), (name, policy, i)
            else:
                assert np.array_equal(p, O.layout_build(TYPE_KIND[t], row)), (name, policy, i, t)
        del M
    assert len(seen) >= 2, seen
    c.close()
