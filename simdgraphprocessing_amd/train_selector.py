"""Train the learned selector on engine timings and evaluate it.

    python -m simdgraphprocessing_amd.train_selector [--out PATH] [--report PATH]

Labels RMAT scales 18-21 x seeds 10-19 (edge factor 16) with the tuned pass's own costs, and
batch arenas of 1,024 ER graphs (seeds 10-13) with the batch passes' own costs, trains the FP32
MLP 16 -> 64 -> 64 -> 4 on the device, saves the parameters (default:
simdgraphprocessing_amd/selector_params.npy), compares its choices with the rules on held-out
rows (the last RMAT seed and the last batch seed), and evaluates triangle counting on RMAT scale
24, edge factor 16, seed 24 (learned layouts vs every rule on the tuned schedule, pass and build
times, and the plain schedule) and the 4,096-graph batch of seed 4 (triangles + 4-cliques per
graph, learned arena layouts vs every rule), with equal counts asserted.
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import numpy as np  # noqa: E402

import simdgraphprocessing_amd as eh  # noqa: E402
from simdgraphprocessing_amd import selector  # noqa: E402


BATCH_SEED_BASE = 1 << 20


def log(*a):
    print(*a, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=selector.PARAMS_PATH)
    ap.add_argument("--report", default=None)
    ap.add_argument("--scales", default="18,19,20,21")
    ap.add_argument("--seeds", default="10-19")
    ap.add_argument("--eval-scale", type=int, default=24)
    ap.add_argument("--eval-seed", type=int, default=24)
    ap.add_argument("--epochs", type=int, default=80)
    ap.add_argument("--check-scales", default="21,22", help="held-out RMAT scales timed in the A-side regime")
    ap.add_argument("--save-data", default=None, help="write the collected labels (.npz) here")
    ap.add_argument("--data", default=None, help="train on labels written by --save-data instead of collecting")
    ap.add_argument("--batch-seeds", default="10-13", help="batch arenas labelled (empty: none)")
    ap.add_argument("--batch-graphs", type=int, default=selector.BATCH_GRAPHS)
    ap.add_argument("--eval-batch", default="4,4096", help="seed,graphs of the batch evaluated (empty: none)")
    a = ap.parse_args()
    s0, s1 = (int(x) for x in a.seeds.split("-"))
    scales = tuple(int(x) for x in a.scales.split(","))
    ctx = eh.Context(0)
    log(f"device {ctx.device_name}")
    t0 = time.time()
    if a.data:
        with np.load(a.data, allow_pickle=False) as z:
            X, T, Mk, seed_of = z["X"], z["T"], z["Mk"], z["seed_of"]
    else:
        X, T, Mk, seed_of = selector.collect(ctx, scales=scales, seeds=range(s0, s1 + 1), log=log)
        if a.batch_seeds:
            b0, b1 = (int(x) for x in a.batch_seeds.split("-"))
            Xb, Tb, Mb, sb = selector.collect_batch(ctx, seeds=range(b0, b1 + 1), ngraphs=a.batch_graphs, log=log)
            # batch seeds are told apart from RMAT seeds by an offset
            X, T, Mk = np.concatenate([X, Xb]), np.concatenate([T, Tb]), np.concatenate([Mk, Mb])
            seed_of = np.concatenate([seed_of, BATCH_SEED_BASE + sb])
        if a.save_data:
            np.savez_compressed(a.save_data, X=X, T=T, Mk=Mk, seed_of=seed_of)
    t_collect = time.time() - t0
    # hold out the graphs of the last RMAT seed and the last batch seed (their rows share no
    # group with the training rows)
    bseeds = seed_of[seed_of >= BATCH_SEED_BASE]
    test = (seed_of == s1) | ((seed_of >= BATCH_SEED_BASE) & (seed_of == (bseeds.max() if len(bseeds) else -1)))
    log(f"collected {len(X)} rows in {t_collect:.1f}s; training {a.epochs} epochs on {int((~test).sum())} rows")
    t1 = time.time()
    params, loss = selector.train(ctx, X[~test], T[~test], Mk[~test], epochs=a.epochs, batch=512, lr=2e-3, seed=1)
    t_train = time.time() - t1
    y = ctx.mlp_forward(params, X[test])
    mse = float(np.sum(Mk[test] * (y - T[test]) ** 2) / np.sum(Mk[test]))
    # the selection's choice: the smallest prediction among the arms it may take (a bitset
    # only within the guard, where the labels have one)
    learned = np.argmin(np.where(Mk[test] > 0, y, 1e9), axis=1)
    oracle_best = np.argmin(np.where(Mk[test] > 0, T[test], 1e9), axis=1)
    choices = {"learned": learned, "density_1_256": selector.heuristic_choice(X[test]),
               "bitset_1_512": selector.bitset_choice(X[test], 8),
               "bitset_1_4096": selector.bitset_choice(X[test], 64),
               "uint": np.full(int(test.sum()), 2), "best_arm": oracle_best}
    cost = {k: selector.choice_cost(T[test], Mk[test], v)[0] for k, v in choices.items()}
    best_rule = min(("density_1_256", "bitset_1_512", "bitset_1_4096", "uint"), key=lambda k: cost[k])
    log(f"collected {len(X)} rows in {t_collect:.1f}s; trained in {t_train:.1f}s; train loss {loss:.4f}; "
        f"held-out mse {mse:.4f}")
    # costs are ns of pass + build against all uint (negative: saved); the gain is the learned
    # choice's saving over the best rule's
    gain = -cost["learned"] / max(-cost[best_rule], 1e-9)
    log("held-out measured pass + build ns vs all uint: " + ", ".join(f"{k} {v:.4e}" for k, v in cost.items()) +
        f" (learned saving vs best rule {best_rule}: {gain:.3f}x)")
    selector.save(params, a.out)
    ctx.set_selector(params)
    # the pass each regime runs on a held-out graph at a training scale, then at the evaluation scale
    for sc in a.check_scales.split(","):
        if sc:
            gc = ctx.generate_rmat(int(sc), 16, s1 + 30)
            selector.compare(ctx, gc, steps=3, log=log, regime="aside")
            del gc
    ev = selector.evaluate(ctx, scale=a.eval_scale, seed=a.eval_seed, log=log)
    log(json.dumps(ev))
    ev_batch = None
    if a.eval_batch:
        bs_, bg_ = (int(x) for x in a.eval_batch.split(","))
        b = ctx.generate_batch(bs_, bg_, policy="hybrid")
        ev_batch = selector.compare_batch(ctx, b, steps=5, log=log)
        ev_batch["workload"] = f"{bg_} ER graphs, seed {bs_}, triangles + 4-cliques per graph"
        del b
        log(json.dumps(ev_batch))
    # the held-out rows of each regime apart (feature 15: 1 one graph, 0 batch arena)
    by_regime = {}
    for name, part in (("graph", X[test][:, 15] > 0.5), ("batch", X[test][:, 15] < 0.5)):
        if part.any():
            Tp, Mp, Xp = T[test][part], Mk[test][part], X[test][part]
            cp = {"learned": selector.choice_cost(Tp, Mp, learned[part])[0],
                  "density_1_256": selector.choice_cost(Tp, Mp, selector.heuristic_choice(Xp))[0],
                  "bitset_1_512": selector.choice_cost(Tp, Mp, selector.bitset_choice(Xp, 8))[0],
                  "uint": 0.0, "best_arm": selector.choice_cost(Tp, Mp, oracle_best[part])[0]}
            br = min(("density_1_256", "bitset_1_512", "uint"), key=lambda k: cp[k])
            by_regime[name] = {"rows": int(part.sum()), "ns_vs_uint": cp, "best_rule": br,
                               "saving_vs_best_rule": -cp["learned"] / max(-cp[br], 1e-9)}
    log(json.dumps(by_regime))
    report = {"train_rows": int(len(X)), "scales": scales, "seeds": [s0, s1], "epochs": a.epochs,
              "train_loss": loss, "heldout_mse": mse, "collect_s": t_collect, "train_s": t_train,
              "heldout_ns_vs_uint": cost, "heldout_best_rule": best_rule,
              "heldout_saving_vs_best_rule": gain,
              "arm_histogram_learned": np.bincount(learned, minlength=4).tolist(),
              "arm_histogram_best": np.bincount(oracle_best, minlength=4).tolist(),
              "heldout_by_regime": by_regime,
              "evaluation": ev, "evaluation_batch": ev_batch}
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    log(json.dumps({k: v for k, v in report.items() if k not in ("evaluation", "evaluation_batch")}))


if __name__ == "__main__":
    main()
