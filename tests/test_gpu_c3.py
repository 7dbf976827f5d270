"""The 4-clique workload on a Chung-Lu power-law graph with planted dense communities
(n_clique at query_depth 4, apps/n_clique.cpp:30-101), reduced from 1,000,000 vertices and 200
communities to 100,000 vertices and 20 communities of 256 at p = .5 (gamma 2.1, 32 draws' worth
of mean degree, seed 3): the device generator and canonicaliser equal the oracle's bit for bit,
and the 4-clique count equals the oracle's under three layout policies, both schedules, every
class of the tuned schedule (edge-centric short rows, local-graph rows, materialised tasks)."""
import numpy as np
import pytest

import oracle as O
import simdgraphprocessing_amd as eh
from simdgraphprocessing_amd import selector

pytestmark = pytest.mark.gpu

N, DEG, GAMMA, SEED, NCOMM, CSIZE = 100000, 32, 2.1, 3, 20, 256


@pytest.fixture(scope="module")
def c3(ctx):
    g = ctx.generate_chung_lu(N, DEG, GAMMA, SEED, communities=NCOMM, community_size=CSIZE)
    off, nbr = g.download()
    src, dst = O.gen_chung_lu(N, DEG, GAMMA, SEED, ncomm=NCOMM, csize=CSIZE)
    ref = O.canonical_oriented(N, src, dst)
    new2old, roff, rnbr = ref["relabel"]
    om = O.Matrix(N, roff, rnbr, "hybrid")
    exp = {"k4": om.cliques(4, 8)[0], "tri": om.triangles(8)[0]}
    print(f"reduced C3: {len(rnbr) // 2} edges, mean degree {len(rnbr) / N:.2f}, "
          f"{exp['tri']} triangles, {exp['k4']} 4-cliques", flush=True)
    return {"g": g, "off": off, "nbr": nbr, "perm": g.perm(), "ref": (new2old, roff, rnbr), "exp": exp}


def test_device_graph_equals_oracle(c3):
    new2old, roff, rnbr = c3["ref"]
    assert np.array_equal(c3["off"], roff) and np.array_equal(c3["nbr"], rnbr)
    assert np.array_equal(c3["perm"], new2old)


@pytest.mark.parametrize("policy", ["hybrid", "pshort", "learned"])
@pytest.mark.parametrize("schedule", ["plain", "tuned"])
def test_c3_4cliques(c3, policy, schedule):
    c = eh.Context(0)
    if policy == "learned":
        c.set_selector(selector.load())
        c.set_selector_schedule(-1)  # the 4-clique passes read every row's layout
    _, roff, rnbr = c3["ref"]
    M = c.matrix_from_csr(N, roff, rnbr, "lower", policy)
    k4, st = M.clique_count(4, schedule)
    print(f"C3/10 {policy}/{schedule}: {k4} 4-cliques in {st['kernel_ms']:.3f} ms", flush=True)
    assert k4 == c3["exp"]["k4"]
    t, _ = M.triangle_count(schedule)
    assert t == c3["exp"]["tri"]
    del M
    c.close()


@pytest.mark.parametrize("small,local", [(-1, -1), (8, -1), (8, 128), (16, 512), (64, 512), (32, 64)])
def test_c3_tuned_classes(ctx, c3, small, local):
    """Every split of rows between the short-row, local-graph and materialised kernels."""
    M = c3["g"].build_matrix("lower", "hybrid")
    k4, _ = M.clique_count(4, "tuned", small_degree=small, local_max=local)
    assert k4 == c3["exp"]["k4"], (small, local)
