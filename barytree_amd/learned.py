"""Learned degree selection: training data, training, shipped models.

The treecode labels its own far-field pairs: for every pair of a plan, the
label kernel finds the least degree in 2..12 whose pair error is <= tol
(BaryTreePlan_LabelPairs), and the pair features are the scale-free
quantities of BaryTreePlan_PairFeatures.  The degree model (an 8-64-64-64-10
MLP) is trained on those pairs with the library's HIP kernels
(BaryTreeModel_Train) and then picks a degree per pair
(BaryTreePlan_UseDegreeModel).  Models are trained on particle sets other
than the ones they are evaluated on (different seeds).

    python -m barytree_amd.learned --model regularized_coulomb_cp   # writes the model file
"""
import argparse
import os
import time

import numpy as np

from . import interface as bt
from .workloads import plummer, uniform_cube

HERE = os.path.dirname(os.path.abspath(__file__))

# name -> how its training pairs are made: kernel, kernel parameters,
# singularity, compute type, theta, degree, leaf size, distribution, N,
# training seeds, label tolerance
MODELS = {
    "regularized_coulomb_cp": dict(
        kernel=bt.Kernel.REGULARIZED_COULOMB, kp=[0.01], sing=bt.Singularity.SKIPPING,
        ct=bt.ComputeType.CLUSTER_PARTICLE, theta=0.7, degree=10, leaf=2000,
        dist="plummer", n=4_000_000, seeds=(41, 42, 43), tol=1e-7, max_pairs=2_000_000),
    "coulomb_cc": dict(
        kernel=bt.Kernel.COULOMB, kp=[0.5], sing=bt.Singularity.SUBTRACTION,
        ct=bt.ComputeType.CLUSTER_CLUSTER, theta=0.8, degree=10, leaf=8000,
        dist="uniform", n=16_000_000, seeds=(51, 52), tol=1e-7, max_pairs=2_000_000),
}


def model_path(name):
    return os.path.join(HERE, f"degree_model_{name}.bin")


def load_model(name):
    """The shipped model `name` (raises if its file is missing)."""
    path = model_path(name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: train it with "
                                f"`python -m barytree_amd.learned --model {name}`")
    return bt.DegreeModel(path=path)


def particles(dist, n, seed):
    gen = plummer if dist == "plummer" else uniform_cube
    return gen(n, seed)


def plan_for(cfg, x, y, z, q, w, precision=bt.PRECISION_FP64):
    return bt.Plan((x.copy(), y.copy(), z.copy(), q.copy()), (x, y, z, q, w), cfg["kernel"],
                   cfg["kp"], cfg["sing"], bt.Approximation.LAGRANGE, cfg["ct"], cfg["theta"],
                   cfg["degree"], cfg["leaf"], cfg["leaf"], precision=precision)


def pair_dataset(plan, tol, chunk=200_000):
    """(features, labels) of every far-field pair of a plan."""
    f = plan.pair_features()
    n = len(f)
    labels = np.zeros(n, dtype=np.int32)
    for o in range(0, n, chunk):
        labels[o:o + chunk] = plan.label_pairs(tol, first=o, count=min(chunk, n - o))
    return f, labels


def training_set(cfg, log=print):
    xs, ys = [], []
    for seed in cfg["seeds"]:
        t0 = time.perf_counter()
        pts = particles(cfg["dist"], cfg["n"], seed)
        with plan_for(cfg, *pts) as p:
            f, lab = pair_dataset(p, cfg["tol"])
        xs.append(f)
        ys.append(lab)
        log(f"[learned] seed {seed}: {len(lab)} pairs labelled in {time.perf_counter() - t0:.1f}s, "
            f"label histogram {np.bincount(lab, minlength=14)[2:].tolist()}")
    x, y = np.concatenate(xs), np.concatenate(ys)
    if len(y) > cfg["max_pairs"]:
        keep = np.random.default_rng(0).choice(len(y), cfg["max_pairs"], replace=False)
        x, y = x[keep], y[keep]
    return x, y


def feature_range(cfg, log=print):
    """(lo, hi) of every pair feature over all far-field pairs of the
    configuration's training particle sets (host-only analysis: the same
    trees and lists a plan builds)."""
    lo = hi = None
    for seed in cfg["seeds"]:
        x, y, z, q, w = particles(cfg["dist"], cfg["n"], seed)
        _, f = bt.analyze_degrees((x, y, z, q), (x, y, z, q, w), cfg["kernel"], cfg["kp"], cfg["sing"],
                                  bt.Approximation.LAGRANGE, cfg["ct"], cfg["theta"], cfg["degree"],
                                  cfg["leaf"], cfg["leaf"])
        a, b = f.min(axis=0), f.max(axis=0)
        lo = a if lo is None else np.minimum(lo, a)
        hi = b if hi is None else np.maximum(hi, b)
        log(f"[learned] seed {seed}: {len(f)} pairs")
    return lo, hi


def train(x, y, epochs=12, batch=4096, lr=2e-3, seed=0):
    m = bt.DegreeModel(seed=seed)
    losses = m.train(x, y, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    return m, losses


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--model", choices=sorted(MODELS), required=True)
    ap.add_argument("--out", default=None)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--set-range", action="store_true",
                    help="only record the training range of the features in an existing model file")
    args = ap.parse_args()
    cfg = MODELS[args.model]
    if args.set_range:
        path = args.out or model_path(args.model)
        m = bt.DegreeModel(path=model_path(args.model))
        lo, hi = feature_range(cfg)
        m.set_feature_range(lo, hi)
        m.save(path)
        print(f"[learned] {args.model}: feature range {lo.tolist()} .. {hi.tolist()}; wrote {path}", flush=True)
        return
    t0 = time.perf_counter()
    x, y = training_set(cfg)
    t1 = time.perf_counter()
    m, losses = train(x, y, epochs=args.epochs)
    t2 = time.perf_counter()
    out = args.out or model_path(args.model)
    m.save(out)
    print(f"[learned] {args.model}: {len(y)} pairs, labelling {t1 - t0:.1f}s, training "
          f"{t2 - t1:.1f}s, loss {losses[0]:.4f} -> {losses[-1]:.4f}; wrote {out}", flush=True)


if __name__ == "__main__":
    main()
