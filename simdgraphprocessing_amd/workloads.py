"""The benchmark workloads, built on the device exactly as the benchmark, the autotuner's
shape refresh and the profiling runs build them.

    c1  triangles, ER G(10000, 100000 distinct non-loop pairs), seed 1
    c2  triangles, RMAT scale 22, edge factor 16, a/b/c/d .57/.19/.19/.05, seed 2
    c3  4-cliques, Chung-Lu n = 1e6, gamma 2.1, 32 n / 2 endpoint draws, plus 200 planted
        communities of 256 vertices at p = .5, seed 3
    c4  triangles + 4-cliques of 4096 ER graphs (n ~ U[64, 8192], mean degree ~ U[4, 64)), seed 4
    c5  triangles, RMAT scale 24, edge factor 16, seed 24 (the learned selector's evaluation graph;
        it trains on RMAT 18-21, seeds 10-19)
All are degree-relabelled and oriented nbr < node.
"""
import os

RMAT = (22, 16, 2)          # c2: scale, edge factor, seed
C1 = (10000, 100000, 1)     # n, distinct pairs, seed
C3 = dict(n=1000000, mean_degree=32, gamma=2.1, seed=3, communities=200, community_size=256)
C4 = (4, 4096)              # seed, graphs
C5 = (24, 16, 24)           # RMAT scale, edge factor, evaluation seed


def load_selector(ctx):
    """Give ctx the trained selector parameters when they exist; True if they did."""
    from . import selector
    if os.path.exists(selector.PARAMS_PATH):
        ctx.set_selector(selector.load())
        return True
    return False


def shape_tag(config, policy):
    return config if config in ("c1", "c3", "c4") else f"{config}:{policy}"


def build(ctx, config, policy):
    """(graph, matrix) of a config, or (None, batch) for c4."""
    if config == "c1":
        g = ctx.generate_er_exact(*C1)
        return g, g.build_matrix("lower", policy)
    if config == "c2":
        g = ctx.generate_rmat(*RMAT)
        return g, g.build_matrix("lower", policy)
    if config == "c3":
        g = ctx.generate_chung_lu(C3["n"], C3["mean_degree"], C3["gamma"], C3["seed"],
                                  communities=C3["communities"], community_size=C3["community_size"])
        return g, g.build_matrix("lower", policy)
    if config == "c4":
        return None, ctx.generate_batch(*C4, policy=policy)
    if config == "c5":
        g = ctx.generate_rmat(*C5)
        return g, g.build_matrix("lower", policy)
    raise ValueError(f"unknown workload {config}")


def run_pass(config, M, params, **kw):
    """One counting pass of the config at launch shape `params`."""
    if config == "c4":
        return M.count("persistent", "tuned", tri_opts=params["tri"], k4_opts=params["k4"])
    if config == "c3":
        return M.clique_count(4, "tuned", **params, **kw)
    return M.triangle_count("tuned", **params, **kw)
