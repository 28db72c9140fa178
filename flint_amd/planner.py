"""Learned planner for integer products.

The exhaustive autotuner (fa_autotune_mpn) times all 64 candidate plans
(np, bits, trunc) of a shape on the GPU, as batched products; the planner network (fa_mlp_*) is trained
on the GPU to predict the cost of every candidate from 16 shape features
(fa_plan_features) -- the cost as log(ms / ms of the shape's fastest
candidate), which keeps the targets on one scale across sizes -- and picks
the valid candidate with the smallest prediction.  Regret of a choice =
time(choice) / time(best candidate) - 1, measured against the exhaustive
timings.

    python -m flint_amd.planner --shapes 2000 --out planner.npz

collects timings for 2000 shapes (log-uniform sizes), trains, reports the
regret of the learned and of the fixed plan rule on held-out shapes and
writes the weights with the dataset summary; `--collect-only FILE` stores
the timings of a run, `--from-dataset A,B,...` trains on stored runs.  The
package ships a dataset measured on MI355X (data/planner_timings_mi355x.npz:
19,008 training and 792 held-out shapes, 892,256 timings) and the weights
trained on it (data/planner_weights_mi355x.npz).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

import flint_amd as fa

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DATASET = os.path.join(DATA, "planner_timings_mi355x.npz")
WEIGHTS = os.path.join(DATA, "planner_weights_mi355x.npz")


def load_dataset(path=DATASET):
    """(train_shapes, train_ms, test_shapes, test_ms) of a stored autotuner run"""
    d = np.load(path, allow_pickle=False)
    return d["train_shapes"], d["train_ms"], d["test_shapes"], d["test_ms"]


def features(shapes):
    return np.array([fa.plan_features(int(a), int(b)) for a, b in shapes], dtype=np.float32)


def load_planner(path=WEIGHTS):
    """the shipped trained network"""
    net = fa.MLP(1)
    net.set_params(np.load(path, allow_pickle=False)["params"])
    return net


def plan_requests(net, reqs):
    """sets the (np, bits, trunc) of every integer request of a ctypes
    Request array to the learned plan"""
    idx = [k for k in range(len(reqs)) if reqs[k].kind == fa.REQ_MPN and reqs[k].an > 0
           and reqs[k].bn > 0]
    if not idx:
        return
    shapes = np.array([[reqs[k].an, reqs[k].bn] for k in idx], dtype=np.int64)
    plans = net.plan_mpn(shapes)
    for k, (np_, bits, tr) in zip(idx, plans):
        reqs[k].np = int(np_)
        reqs[k].bits = int(bits)
        reqs[k].trunc = fa.TRUNC_ON if tr else fa.TRUNC_OFF


def sample_shapes(n, seed, lo=4.0, hi=22.0):
    """(an, bn) pairs, each log-uniform in [2^lo, 2^hi] limbs"""
    rng = np.random.default_rng(seed)
    an = np.floor(2.0 ** rng.uniform(lo, hi, n)).astype(np.int64)
    bn = np.floor(2.0 ** rng.uniform(lo, hi, n)).astype(np.int64)
    return np.stack([an, bn], axis=1)


def collect(shapes, reps=3, progress=None):
    """features (n x 16) and candidate times (n x 64, -1 = invalid)"""
    X = np.zeros((len(shapes), 16), dtype=np.float32)
    T = np.zeros((len(shapes), 64), dtype=np.float32)
    t0 = time.time()
    for k, (an, bn) in enumerate(shapes):
        X[k] = fa.plan_features(int(an), int(bn))
        T[k] = fa.autotune_mpn(int(an), int(bn), reps)
        if progress and (k + 1) % progress == 0:
            print("# autotuned %d/%d shapes (%.1fs)" % (k + 1, len(shapes), time.time() - t0),
                  file=sys.stderr, flush=True)
    return X, T


def targets(ms):
    """cost targets log(ms / min ms of the shape) and the validity mask"""
    M = (ms > 0).astype(np.float32)
    best = np.where(ms > 0, ms, np.inf).min(axis=1, keepdims=True)
    T = np.where(ms > 0, np.log(np.maximum(ms, 1e-9) / best), 0.0).astype(np.float32)
    return T, M


def train(X, ms, epochs=1000, batch=128, lr=1e-3, seed=1):
    T, M = targets(ms)
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(X))
    net = fa.MLP(seed)
    losses = net.train(X[order], T[order], M[order], batch=batch, epochs=epochs, lr=lr)
    return net, losses


def auto_truncates(p):
    """the automatic truncation rule (fa_trunc_rule) of a plan dict"""
    d1 = p["depth"] - 12
    t = -(-p["zlen"] // 4096)
    return d1 == 11 and 20 * t <= 11 << d1


def heuristic_index(an, bn):
    """candidate index of the automatic plan: fa_mpn_choose_plan's (np, bits)
    and the automatic truncation rule"""
    p = fa.mpn_plan(int(an), int(bn))
    c = fa.mpn_candidates(int(an), int(bn))
    tr = 1 if auto_truncates(p) else 0
    for j, (np_, bits, t) in enumerate(c):
        if np_ == p["np"] and bits == p["bits"] and t == tr:
            return j
    raise ValueError("automatic plan is not a candidate")


def learned_index(net, shapes):
    plans = net.plan_mpn(shapes)
    out = []
    for (an, bn), pl in zip(shapes, plans):
        c = fa.mpn_candidates(int(an), int(bn))
        out.append(next(j for j, x in enumerate(c) if tuple(x) == tuple(pl)))
    return np.array(out)


def regret(ms, idx):
    """per-shape time(chosen) / time(best) - 1"""
    best = np.where(ms > 0, ms, np.inf).min(axis=1)
    chosen = ms[np.arange(len(ms)), idx]
    return chosen / best - 1.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--shapes", type=int, default=600)
    ap.add_argument("--test-frac", type=float, default=0.25)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=0xF1A70005)
    ap.add_argument("--epochs", type=int, default=1000)
    ap.add_argument("--out", default="")
    ap.add_argument("--from-dataset", default="", metavar="NPZ[,NPZ...]",
                    help="train on stored datasets (comma-separated; 'shipped' = the "
                         "package's) instead of collecting timings")
    ap.add_argument("--collect-only", default="", metavar="NPZ",
                    help="collect timings of --shapes shapes (seed --seed; the first "
                         "--test-frac of them held out) into NPZ and stop")
    args = ap.parse_args()
    if args.collect_only:
        shapes = sample_shapes(args.shapes, args.seed)
        _, ms = collect(shapes, args.reps, progress=25)
        ntest = int(len(shapes) * args.test_frac)
        np.savez(args.collect_only, train_shapes=shapes[ntest:], train_ms=ms[ntest:],
                 test_shapes=shapes[:ntest], test_ms=ms[:ntest])
        print(json.dumps({"shapes": int(len(shapes)), "timings": int((ms > 0).sum())}))
        return
    if args.from_dataset:
        parts = [load_dataset(DATASET if f == "shipped" else f)
                 for f in args.from_dataset.split(",")]
        trs = np.concatenate([p[0] for p in parts])
        trm = np.concatenate([p[1] for p in parts])
        tes = np.concatenate([p[2] for p in parts])
        tem = np.concatenate([p[3] for p in parts])
        Xtr = features(trs)
    else:
        shapes = sample_shapes(args.shapes, args.seed)
        X, ms = collect(shapes, args.reps, progress=50)
        ntest = int(len(shapes) * args.test_frac)
        trs, trm, tes, tem = shapes[ntest:], ms[ntest:], shapes[:ntest], ms[:ntest]
        Xtr = X[ntest:]
    ntest = len(tes)
    t0 = time.time()
    net, losses = train(Xtr, trm, epochs=args.epochs)
    ttrain = time.time() - t0
    li = learned_index(net, tes)
    hi = np.array([heuristic_index(a, b) for a, b in tes])
    rl, rh = regret(tem, li), regret(tem, hi)
    tl = tem[np.arange(ntest), li].sum()
    th = tem[np.arange(ntest), hi].sum()
    tb = np.where(tem > 0, tem, np.inf).min(axis=1).sum()
    rec = {
        "train_shapes": int(len(trs)), "timings": int((trm > 0).sum() + (tem > 0).sum()),
        "train_seconds": ttrain, "final_loss": float(losses[-1]), "test_shapes": ntest,
        "regret_learned_mean": float(rl.mean()), "regret_heuristic_mean": float(rh.mean()),
        "regret_learned_p90": float(np.percentile(rl, 90)),
        "time_learned_ms": float(tl), "time_heuristic_ms": float(th), "time_best_ms": float(tb),
    }
    print(json.dumps(rec))
    if args.out:
        np.savez(args.out, params=net.params(), summary=np.array(json.dumps(rec)))


if __name__ == "__main__":
    main()
