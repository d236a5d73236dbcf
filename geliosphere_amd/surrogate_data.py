"""Training and evaluation spectra for the surrogate, made by the batched
engine.

Parameters come from the batch sampler (geliosphere_amd.batch.sample_one:
K0 log-uniform 1e22-1e23 cm^2/s, V 300-800 km/s or tilt 5-75 deg, r 1-80 AU,
theta 10-170 deg, dt from {50, 100, 500} s, 2D polarity +-1), with the
model fixed per set and the particle count per model set by the caller.

    python -m geliosphere_amd.surrogate_data --out FILE.npz
        [--train-per-model 1000] [--eval-per-model 40]
        [--n-train 1D_FP,1D_BP,2D_SP,2D_GE] [--n-eval ...] [--seed S]
        [--models 0,1,2,3]
    python -m geliosphere_amd.surrogate_data --out FILE.npz --join A.npz B.npz ...

writes x, y, w (training), ex, ey, ew (evaluation), the model of every row,
the particle counts used, and `sets`: one row per model generated, [model,
training spectra, training particles, evaluation spectra, evaluation
particles, seed], from which sims_for() re-draws every simulation (training
seed S, evaluation seed S + 1).  Files made in several runs are joined with
merge().
"""
import argparse
import random
import sys
import time

import numpy as np

from . import abi, batch, surrogate
from .engine import Engine

MODELS = (abi.MODEL_1D_FP, abi.MODEL_1D_BP, abi.MODEL_2D_SOLARPROP, abi.MODEL_2D_GELIOSPHERE)
DEFAULT_N_TRAIN = {abi.MODEL_1D_FP: 10_000_000, abi.MODEL_1D_BP: 100_000,
                   abi.MODEL_2D_SOLARPROP: 20_000, abi.MODEL_2D_GELIOSPHERE: 20_000}
DEFAULT_N_EVAL = {abi.MODEL_1D_FP: 100_000_000, abi.MODEL_1D_BP: 1_000_000,
                  abi.MODEL_2D_SOLARPROP: 200_000, abi.MODEL_2D_GELIOSPHERE: 200_000}


def sims_for(model, count, n_particles, seed, first_sim_seed):
    rng = random.Random(seed * 7919 + model)
    out = []
    for i in range(count):
        p = batch.sample_one(rng, first_sim_seed + i, model=model)
        p.n_particles = n_particles
        out.append(p)
    return out


def run_set(engine, sims, label, chunk=32):
    if not sims:   # an evaluation-only or training-only run
        n_out = surrogate.N_OUT
        return (np.zeros((0, surrogate.N_IN), np.float32), np.zeros((0, n_out), np.float32),
                np.zeros((0, n_out), np.float32))
    xs, ys, ws = [], [], []
    t0 = time.time()
    for i in range(0, len(sims), chunk):
        x, y, w = surrogate.generate(engine, sims[i:i + chunk], chunk=chunk)
        xs.append(x)
        ys.append(y)
        ws.append(w)
        print("%s: %d/%d spectra, %.1f s" % (label, min(i + chunk, len(sims)), len(sims), time.time() - t0),
              flush=True)
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(ws)


def redraw(sets):
    """(training simulations, evaluation simulations) of a data set's `sets`
    rows, in the order of its x and ex rows."""
    tr, ev = [], []
    for m, n_tr, p_tr, n_ev, p_ev, seed in np.asarray(sets).tolist():
        tr += sims_for(m, n_tr, p_tr, seed, 1_000_000 * (m + 1))
        ev += sims_for(m, n_ev, p_ev, seed + 1, 50_000_000 * (m + 1))
    return tr, ev


def merge(paths, out):
    """Join data sets made by separate runs (rows in the order given)."""
    ds = [np.load(p, allow_pickle=False) for p in paths]
    res = {k: np.concatenate([d[k] for d in ds]) for k in ("x", "y", "w", "m", "ex", "ey", "ew", "em", "sets")}
    res["n_train"] = np.max(np.stack([d["n_train"] for d in ds]), axis=0)
    res["n_eval"] = np.max(np.stack([d["n_eval"] for d in ds]), axis=0)
    np.savez_compressed(out, **res)
    return res


def parse_counts(s, default):
    if not s:
        return dict(default)
    vals = [int(float(v)) for v in s.split(",")]
    return dict(zip(MODELS, vals))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--train-per-model", type=int, default=1000)
    ap.add_argument("--eval-per-model", type=int, default=40)
    ap.add_argument("--n-train", default="")
    ap.add_argument("--n-eval", default="")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--models", default="0,1,2,3", help="gs_model values to generate")
    ap.add_argument("--join", nargs="+", default=None, metavar="NPZ",
                    help="write the join of these data sets to --out (no generation)")
    a = ap.parse_args(argv)
    if a.join:
        res = merge(a.join, a.out)
        print("wrote %s: %d training, %d evaluation spectra" % (a.out, len(res["x"]), len(res["ex"])), flush=True)
        return 0
    models = [int(v) for v in a.models.split(",")]
    n_train = parse_counts(a.n_train, DEFAULT_N_TRAIN)
    n_eval = parse_counts(a.n_eval, DEFAULT_N_EVAL)
    eng = Engine(0)
    parts = {"x": [], "y": [], "w": [], "m": [], "ex": [], "ey": [], "ew": [], "em": []}
    for m in models:
        tr = sims_for(m, a.train_per_model, n_train[m], a.seed, 1_000_000 * (m + 1))
        x, y, w = run_set(eng, tr, "train %s" % abi.MODEL_NAMES[m])
        parts["x"].append(x); parts["y"].append(y); parts["w"].append(w); parts["m"].append(np.full(len(x), m))
        ev = sims_for(m, a.eval_per_model, n_eval[m], a.seed + 1, 50_000_000 * (m + 1))
        x, y, w = run_set(eng, ev, "eval %s" % abi.MODEL_NAMES[m], chunk=8)
        parts["ex"].append(x); parts["ey"].append(y); parts["ew"].append(w); parts["em"].append(np.full(len(x), m))
    out = {k: np.concatenate(v) for k, v in parts.items()}
    out["n_train"] = np.array([n_train[m] if m in models else 0 for m in MODELS])
    out["n_eval"] = np.array([n_eval[m] if m in models else 0 for m in MODELS])
    out["sets"] = np.array([[m, a.train_per_model, n_train[m], a.eval_per_model, n_eval[m], a.seed] for m in models],
                           dtype=np.int64)
    np.savez_compressed(a.out, **out)
    print("wrote %s: %d training, %d evaluation spectra" % (a.out, len(out["x"]), len(out["ex"])), flush=True)
    eng.close()


if __name__ == "__main__":
    sys.exit(main())
