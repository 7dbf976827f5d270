"""n_path's BFS (apps/n_path.cpp:27-87) and the bitset set_union / set_difference it uses
(src/set/ops/union.hpp:6-127, difference.hpp:5-84) on the device, against the oracle."""
import gzip
import os

import numpy as np
import pytest

import oracle as O
import simdgraphprocessing_amd as eh

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _dfacebook():
    raw = gzip.open(os.path.join(GOLDEN, "dfacebook.bin.gz"), "rb").read()
    n, outs, ins, ext = O.load_directed_binary(raw)
    cat = lambda rows: np.concatenate(rows).astype(np.uint32)
    off = np.zeros(2 * n + 1, np.uint64)
    off[1:] = np.cumsum([len(r) for r in outs + ins])
    return n, off, cat(outs + ins), ext


def check(M, n, off, nbr, start, depth, n_vertices=0, modes=("top_down", "symmetric")):
    """Every mode (top-down only, or bottom-up rounds where they pay) gives the oracle's
    path length, rounds, level sizes and reach."""
    want = O.bfs_path(n, off, nbr, start, depth, cap=256)
    for mode in modes:
        got = M.bfs_path(start, depth, n_vertices=n_vertices, mode=mode)
        for k in ("path_length", "rounds", "reached", "level_sizes"):
            assert got[k] == want[k], (mode, k, start, depth, got[k], want[k])
    return got


def test_bfs_fixture_dfacebook(ctx):
    """test/n_path_test.cpp:7-11: dfacebook.bin (out- and in-rows), start 0, depth 4 -> 4."""
    n, off, nbr, ext = _dfacebook()
    M = ctx.matrix_from_csr(2 * n, off, nbr, "all", "uint")
    start = int(np.nonzero(ext == 0)[0][0])
    out_off, out_nbr = off[:n + 1], nbr[:int(off[n])]
    got = check(M, n, out_off, out_nbr, start, 4, n_vertices=n, modes=("top_down", "in_rows"))
    assert got["path_length"] == 4
    for s in (5, 1000, 3000):
        check(M, n, out_off, out_nbr, s, 4, n_vertices=n, modes=("top_down", "in_rows"))
    for s in (0, 77):   # no depth limit
        check(M, n, out_off, out_nbr, s, 0xEFFFFFF, n_vertices=n, modes=("top_down", "in_rows"))


def test_bfs_in_rows_bottom_up(ctx):
    """A symmetric RMAT-12 given as out-rows + in-rows (2n rows): wide rounds run bottom-up
    over rows [n, 2n) and still give the oracle's levels."""
    n = 1 << 12
    _, off, nbr = O.canonical_oriented(n, *O.gen_rmat(12, 8, 4))["relabel"]
    off2 = np.concatenate([off, off[1:] + off[-1]]).astype(np.uint64)
    nbr2 = np.concatenate([nbr, nbr]).astype(np.uint32)
    M = ctx.matrix_from_csr(2 * n, off2, nbr2, "all", "hybrid")
    for s in (0, 5, 3000):
        check(M, n, off, nbr, s, 0xEFFFFFF, n_vertices=n, modes=("top_down", "in_rows"))
    assert M.bfs_path(0, n_vertices=n, mode="in_rows")["bottom_up"] > 0


@pytest.mark.parametrize("policy", ["uint", "hybrid", "bs", "pshort", "bitset_new"])
def test_bfs_facebook(ctx, facebook, policy):
    n, off, nbr = facebook["n"], facebook["off"], facebook["nbr"]
    M = ctx.matrix_from_csr(n, off, nbr, "all", policy)
    start = int(np.argmax(np.diff(off)))   # get_max_row_id: the first row of maximum length
    assert check(M, n, off, nbr, start, 0xEFFFFFF)["path_length"] == 5
    for s, d in ((0, 2), (17, 100), (4000, 3), (2500, 0)):
        check(M, n, off, nbr, s, d)


@pytest.mark.parametrize("scale", [12, 16])
def test_bfs_rmat(ctx, scale):
    n = 1 << scale
    _, off, nbr = O.canonical_oriented(n, *O.gen_rmat(scale, 8, scale))["relabel"]
    for policy in ("hybrid", "uint"):
        M = ctx.matrix_from_csr(n, off, nbr, "all", policy)
        rng = np.random.default_rng(scale)
        for s in rng.integers(0, n, 6).tolist() + [0, n - 1]:
            for d in (0xEFFFFFF, 3):
                g = check(M, n, off, nbr, s, d)
        assert M.bfs_path(0, mode="symmetric")["bottom_up"] > 0
        print(f"RMAT-{scale} {policy}: last BFS {g['kernel_ms']:.3f} ms, reached {g['reached']}", flush=True)


def test_bfs_edges(ctx):
    # a path, a star and isolated vertices
    src = list(range(39)) + [50] * 20
    dst = list(range(1, 40)) + list(range(51, 71))
    off, nbr = O.edges_to_csr(100, np.array(src, np.uint32), np.array(dst, np.uint32))
    M = ctx.matrix_from_csr(100, off, nbr, "all", "hybrid")
    assert check(M, 100, off, nbr, 0, 0xEFFFFFF)["path_length"] == 39
    assert check(M, 100, off, nbr, 55, 0xEFFFFFF)["path_length"] == 2
    got = check(M, 100, off, nbr, 90, 0xEFFFFFF)
    assert got["path_length"] == 0 and got["rounds"] == 1 and got["reached"] == 1
    with pytest.raises(eh.EHError):
        M.bfs_path(100)
    with pytest.raises(eh.EHError):
        M.bfs_path(0, n_vertices=101)
    with pytest.raises(eh.EHError):
        M.bfs_path(0, n_vertices=60, mode="in_rows")   # in-rows need 2 x 60 rows


def _bitset(vals):
    return O.set_payload(np.asarray(vals, np.uint32), O.BITSET)


def test_union_and_difference(ctx):
    rng = np.random.default_rng(21)
    for _ in range(60):
        a = np.unique(rng.integers(0, 4000, int(rng.integers(2, 300)))).astype(np.uint32)
        b = np.unique(rng.integers(0, 6000, int(rng.integers(1, 300)))).astype(np.uint32)
        A = _bitset(a)
        lo, hi = 64 * (int(a[0]) // 64), 64 * (int(a[-1]) // 64 + 1)   # A's word span
        inside = b[(b >= lo) & (b < hi)]
        want = np.union1d(a, inside)
        for tb in (O.BITSET, O.UINTEGER, O.PSHORT):
            c, t, p = ctx.union(A, O.set_payload(b, tb))
            assert t == O.BITSET and c == len(want)
            assert np.array_equal(p, O.layout_build("bitset", want)[:len(p)]) or \
                sorted(O.decode(O.BITSET, p, c).tolist()) == want.tolist()
            assert len(p) == len(A[2])
        c, t, p = ctx.difference(A, _bitset(b))
        ref = O.bitset_difference(A[2], _bitset(b)[2])
        assert t == O.UINTEGER and c == len(ref) == len(np.setdiff1d(a, b))
        assert np.array_equal(np.frombuffer(p.tobytes(), np.uint32), ref)
    # n_path's use: visited (span from word 0) minus the previous visited
    old = _bitset([0, 5, 64, 700])
    new = ctx.union(old, O.set_payload(np.array([5, 6, 65, 699, 700], np.uint32), O.UINTEGER))
    c, t, p = ctx.difference((O.BITSET, new[0], new[2]), old)
    assert np.frombuffer(p.tobytes(), np.uint32).tolist() == [6, 65, 699]
    with pytest.raises(eh.EHError):
        ctx.union(O.set_payload([1, 2], O.UINTEGER), _bitset([1]))


def test_bottom_up_modes_need_every_entry(ctx, facebook):
    """A LOWER (oriented) matrix holds half of each neighbourhood: the bottom-up modes, which
    search a vertex's own row for frontier members, reject it; top-down runs as the reference's
    n_path does over whatever rows the matrix keeps."""
    import simdgraphprocessing_amd as eh
    M = ctx.matrix_from_csr(facebook["n"], facebook["off"], facebook["nbr"], "lower", "hybrid")
    for mode in ("symmetric", "in_rows"):
        with pytest.raises(eh.EHError) as e:
            M.bfs_path(0, mode=mode)
        assert e.value.status == 1 and "EH_FILTER_ALL" in str(e.value)
    M.bfs_path(facebook["n"] - 1)  # top-down is accepted
