"""GPU parity of the set-level intersection (eh_intersect) and of the batched
row-pair form (eh_intersect_rows) against the oracle's typed dispatcher
(ops::set_intersect, src/set/ops/intersection.hpp:2066-2118): counts, result
types and result bytes must be identical."""
import random

import numpy as np
import pytest

import oracle as O
import simdgraphprocessing_amd as eh

TYPES = {"U": O.UINTEGER, "PS": O.PSHORT, "BS": O.BITSET}


def gen_set(rng, length, span, run):
    s = set()
    length = min(length, span // 2)
    base = rng.randrange(0, max(1, span // 10))
    while len(s) < length:
        v = base + rng.randrange(0, span)
        for t in range(run):
            s.add(v + t)
    return sorted(s)[:length]


def cases(seed):
    rng = random.Random(seed)
    out = [([63], [63]), ([64], [63, 64]), ([65535, 65536], [65536, 70000]), ([0, 1, 2], [2, 3]),
           ([1 << 20], [1 << 20, (1 << 20) + 64]), ([5], [6]), (list(range(100)), list(range(50, 60))),
           ([0, 1 << 30], [1 << 30])]
    for _ in range(10):
        span = rng.choice([100, 5000, 200000, 3 << 20])
        out.append((gen_set(rng, rng.randrange(1, 300), span, rng.choice([1, 3])),
                    gen_set(rng, rng.randrange(1, 300), span, 1)))
    # several workgroup-sized chunks: > 1024 elements, > 1024 bitset words, many partitions
    out.append((gen_set(rng, 30000, 1 << 19, 2), gen_set(rng, 20000, 1 << 19, 1)))
    out.append((gen_set(rng, 5000, 1 << 22, 1), gen_set(rng, 70000, 1 << 22, 4)))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("ta", ["U", "PS", "BS"])
@pytest.mark.parametrize("tb", ["U", "PS", "BS"])
def test_intersect_matches_oracle_bytes(ctx, ta, tb):
    for a, b in cases(3 * list(TYPES).index(ta) + list(TYPES).index(tb)):
        c, rt, rb = O.set_intersect_bytes(a, TYPES[ta], b, TYPES[tb])
        gc, gt, gb = ctx.intersect(O.set_payload(a, TYPES[ta]), O.set_payload(b, TYPES[tb]))
        assert gc == c == len(set(a) & set(b)), (ta, tb, len(a), len(b))
        assert gt == rt, (ta, tb)
        assert np.array_equal(gb, rb), (ta, tb, len(a), len(b), gb[:32], rb[:32])
        k, _, _ = ctx.intersect(O.set_payload(a, TYPES[ta]), O.set_payload(b, TYPES[tb]), materialize=False)
        assert k == c


@pytest.mark.gpu
def test_intersect_empty_side(ctx):
    c, t, b = ctx.intersect(O.set_payload([], O.UINTEGER), O.set_payload([1, 2, 3], O.BITSET))
    assert c == 0 and len(b) == 0


@pytest.mark.gpu
def test_intersect_reports_small_buffer(ctx):
    a = O.set_payload(list(range(1000)), O.UINTEGER)
    with pytest.raises(eh.EHError):
        ctx.intersect(a, a, out_cap=16)


@pytest.mark.gpu
def test_intersect_rows_matches_oracle(ctx, facebook):
    n, off, nbr = facebook["n"], facebook["off"], facebook["nbr"]
    M = ctx.matrix_from_csr(n, off, nbr, "lower", "hybrid")
    om = O.Matrix(n, off, nbr, "hybrid")
    rng = np.random.default_rng(5)
    a = rng.integers(0, n, 20000, dtype=np.uint32)
    b = rng.integers(0, n, 20000, dtype=np.uint32)
    got = M.intersect_rows(a, b)
    rows = {}

    def vals(i):
        if i not in rows:
            t, card, payload = om.row(int(i))
            rows[i] = set(O.decode(t, payload, card).tolist())
        return rows[i]

    exp = np.array([len(vals(x) & vals(y)) for x, y in zip(a, b)], np.uint64)
    assert np.array_equal(got, exp)


@pytest.mark.gpu
def test_intersect_row_sets(ctx, facebook):
    """eh_intersect_row_sets: the intersection of 1 to 16 rows per request, counted and
    materialised, against numpy over the matrix's own rows; malformed requests are refused."""
    M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "all", "hybrid")
    off, nbr = M.download()
    n = facebook["n"]
    rng = np.random.default_rng(5)
    sets = [[0], [3, 3], [int(n - 1), 0]]
    # neighbourhood chains: rows that share neighbours, so deep intersections stay non-empty
    for _ in range(400):
        k = int(rng.integers(1, 17))
        v = int(rng.integers(0, n))
        row = nbr[off[v]:off[v + 1]]
        pick = [v] + (rng.choice(row, min(k - 1, len(row)), replace=False).tolist() if len(row) else [])
        sets.append([int(x) for x in pick])
    counts, elems = M.intersect_row_sets(sets, materialize=True)
    counts2, none = M.intersect_row_sets(sets)
    assert none is None and np.array_equal(counts, counts2)
    deep = 0
    for s, c, e in zip(sets, counts, elems):
        ref = nbr[off[s[0]]:off[s[0] + 1]]
        for r in s[1:]:
            ref = np.intersect1d(ref, nbr[off[r]:off[r + 1]], assume_unique=True)
        assert c == len(ref) and np.array_equal(e, ref), s
        deep += len(s) >= 3 and c > 0
    assert deep > 20
    import simdgraphprocessing_amd as eh
    for bad in ([list(range(17))], [[n]], [[]]):
        with pytest.raises(eh.EHError) as ex:
            M.intersect_row_sets(bad)
        assert ex.value.status == 1
