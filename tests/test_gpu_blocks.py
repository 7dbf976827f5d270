"""Block-bitset layouts on the device: bitset_new / new_type row images, the NEW_BITSET
hybrid, counts under both schedules, and eh_intersect over the reference's block
overloads (src/set/layouts/bitset_new.hpp:66-108, new_type.hpp:50-88,
src/set/ops/intersection.hpp:1338-1649), all against the oracle."""
import numpy as np
import pytest

import oracle as O
import simdgraphprocessing_amd as eh

pytestmark = pytest.mark.gpu

BLOCK_POLICIES = ["bitset_new", "new_type", "hybrid_block"]
# ROW DATA SIZE of the reference's triangle program on facebook.bin (one thread):
# --layout=bitset_new, --layout=new_type, --layout=hybrid built with -DNEW_BITSET
FB_BLOCK_ROW_BYTES = {"bitset_new": 513116, "new_type": 277200, "hybrid_block": 461324}
FB_TRIANGLES, FB_4CLIQUES = 1612010, 30004668


def _rmat(scale, ef, seed):
    src, dst = O.gen_rmat(scale, ef, seed)
    _, off, nbr = O.canonical_oriented(1 << scale, src, dst)["relabel"]
    return 1 << scale, off, nbr


@pytest.mark.parametrize("policy", BLOCK_POLICIES)
def test_block_row_images(ctx, facebook, policy):
    """Every row's type, cardinality and bytes equal the oracle's image."""
    for n, off, nbr in ((facebook["n"], facebook["off"], facebook["nbr"]), _rmat(12, 8, 3)):
        M = ctx.matrix_from_csr(n, off, nbr, "lower", policy)
        om = O.Matrix(n, off, nbr, policy)
        assert M.info()["ref_row_bytes"] == om.bytes()
        assert M.info()["rows_by_type"]["bitset"] == om.type_counts()["bitset"]
        for i in range(n):
            t, c, p = M.row(i)
            ot, oc, op = om.row(i)
            assert c == oc, (policy, i)
            if c:
                assert t == ot and np.array_equal(p, op), (policy, i, t, ot, len(p), len(op))
    M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "lower", policy)
    assert M.info()["ref_row_bytes"] == FB_BLOCK_ROW_BYTES[policy]


@pytest.mark.parametrize("policy", BLOCK_POLICIES + ["cost"])
@pytest.mark.parametrize("schedule", ["plain", "tuned"])
def test_block_facebook_counts(ctx, facebook, policy, schedule):
    M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "lower", policy)
    t, _ = M.triangle_count(schedule)
    c, _ = M.clique_count(4, schedule)
    print(f"facebook {policy}/{schedule}: {t} triangles, {c} 4-cliques", flush=True)
    assert t == FB_TRIANGLES and c == FB_4CLIQUES


@pytest.mark.parametrize("policy", BLOCK_POLICIES)
def test_block_rmat_counts(ctx, policy):
    """RMAT-14: triangles and 4-cliques under both schedules, several small-row
    thresholds, the unoriented matrix, and 5-cliques, against the oracle."""
    n, off, nbr = _rmat(14, 16, 9)
    om = O.Matrix(n, off, nbr, policy)
    tri, k4 = om.triangles(8)[0], om.cliques(4, 8)[0]
    M = ctx.matrix_from_csr(n, off, nbr, "lower", policy)
    for schedule in ("plain", "tuned"):
        assert M.triangle_count(schedule)[0] == tri, schedule
        assert M.clique_count(4, schedule)[0] == k4, schedule
    for sd in (0, 4, 64):
        assert M.triangle_count("tuned", small_degree=sd)[0] == tri, sd
    assert M.clique_count(5, "tuned")[0] == om.cliques(5, 8)[0]
    U = ctx.matrix_from_csr(n, off, nbr, "all", policy)
    assert U.triangle_count("tuned")[0] == 6 * tri


def _sets(seed, count=120):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        kind = int(rng.integers(0, 5))
        if kind == 0:
            v = np.unique(rng.integers(0, 1 << 24, int(rng.integers(1, 200))))
        elif kind == 1:
            s = int(rng.integers(0, 1 << 20))
            v = np.arange(s, s + int(rng.integers(1, 600)))
        elif kind == 2:
            b = 128 * int(rng.integers(0, 1000))
            v = b + np.sort(rng.choice(128, int(rng.choice([1, 2, 8, 9, 64, 128])), replace=False))
        else:
            hi = int(rng.integers(256, 30000))
            parts = [rng.integers(0, hi, int(rng.integers(1, 300)))]
            for _ in range(int(rng.integers(0, 5))):
                b = 128 * int(rng.integers(0, hi // 128 + 1))
                parts.append(b + rng.choice(128, int(rng.integers(2, 128)), replace=False))
            v = np.unique(np.concatenate(parts))
        out.append(np.asarray(v, np.uint32))
    return out


def test_block_intersect_matches_oracle(ctx):
    """eh_intersect over bitset_new x bitset_new, uint x bitset_new (both orders) and
    new_type x new_type: count, result type and bytes equal the oracle's."""
    sets = _sets(11)
    rng = np.random.default_rng(12)
    pairs = [(sets[int(rng.integers(0, len(sets)))], sets[int(rng.integers(0, len(sets)))]) for _ in range(150)]
    pairs += [(sets[0], sets[0]), (sets[1][::2].copy(), sets[1])]
    for a, b in pairs:
        for ta, tb in ((O.BITSET_NEW, O.BITSET_NEW), (O.UINTEGER, O.BITSET_NEW), (O.BITSET_NEW, O.UINTEGER),
                       (O.NEW_TYPE, O.NEW_TYPE)):
            c, rt, rbytes = O.set_intersect_bytes(a, ta, b, tb)
            gc, gt, gp = ctx.intersect(O.set_payload(a, ta), O.set_payload(b, tb))
            assert gc == c == len(np.intersect1d(a, b)), (ta, tb)
            assert gt == rt and np.array_equal(gp, rbytes), (ta, tb, len(a), len(b))
            assert ctx.intersect(O.set_payload(a, ta), O.set_payload(b, tb), materialize=False)[0] == c


def test_block_intersect_empty_and_errors(ctx):
    e = O.set_payload([], O.BITSET_NEW)
    x = O.set_payload([1, 2, 300], O.BITSET_NEW)
    c, t, p = ctx.intersect(e, x)
    assert c == 0 and t == O.UINTEGER and len(p) == 0   # the block overloads return an empty uint set
    with pytest.raises(eh.EHError):   # no reference overload for new_type x uint
        ctx.intersect(O.set_payload([1, 2], O.NEW_TYPE), O.set_payload([1, 2], O.UINTEGER))
    t, c, p = x
    bad = p.copy()
    bad[:4] = np.frombuffer(np.uint32(1 << 26).tobytes(), np.uint8)   # block id past 2^25
    with pytest.raises(eh.EHError):
        ctx.intersect((t, c, bad), x)
