"""Train the surrogate on engine-made spectra and measure it.

    python -m geliosphere_amd.surrogate_eval --data A.npz [B.npz ...]
        [--epochs 1500] [--save WEIGHTS.npz] [--json OUT.json]
    python -m geliosphere_amd.surrogate_eval --data A.npz --load WEIGHTS.npz
        (measures a saved ensemble; trains only if --epochs is given too)

Reports, per model and overall, on the held-out evaluation spectra (made
with 10x more particles than the training spectra):
  * median over evaluation points and energy bins in 0.1-10 GeV of the
    relative error |J_sur / J_MC - 1| (J_sur / J_MC = exp(y_sur - y_MC));
  * the same for the nearest training spectrum in standardised input space
    (the method without a network: a lookup table of Monte Carlo runs);
  * the Monte Carlo reference's own median relative standard error;
  * the share of queries the ensemble sends back to Monte Carlo;
and the inference rate for a 65,536-query batch (spectra/s).
"""
import argparse
import json
import sys
import time

import numpy as np

from . import abi, surrogate

MODEL_NAMES = {0: "1D Fp", 1: "1D Bp", 2: "2D SolarProp-like", 3: "2D Geliosphere"}


def load(paths):
    keys = ("x", "y", "w", "m", "ex", "ey", "ew", "em")
    parts = {k: [] for k in keys}
    for p in paths:
        d = np.load(p, allow_pickle=False)
        for k in keys:
            parts[k].append(d[k])
    return {k: np.concatenate(v) for k, v in parts.items()}


def window_mask(lo=0.1, hi=10.0):
    p = abi.make_params(abi.MODEL_1D_BP)
    e = np.array([p.bin_edges[i] for i in range(p.n_bins + 1)])
    return (e[:-1] >= lo * (1 - 1e-12)) & (e[1:] <= hi * (1 + 1e-12))


def rel_errors(y_hat, y_ref, w_ref, mask):
    sel = (w_ref > 0) & mask[None, :]
    return np.abs(np.exp(y_hat - y_ref) - 1.0)[sel]


def nearest_neighbour(d, mean, sd):
    xs = (d["x"] - mean) / sd
    ex = (d["ex"] - mean) / sd
    out = np.zeros_like(d["ey"])
    for i in range(len(ex)):
        same = d["m"] == d["em"][i]
        cand = np.where(same)[0]
        j = cand[np.argmin(((xs[cand] - ex[i]) ** 2).sum(axis=1))]
        out[i] = d["y"][j]
    return out


def evaluate(sg, d):
    mask = window_mask()
    mean, sd, fb = sg.predict(d["ex"])
    nmu, nsd, _, _ = sg.get_normalization()
    nn = nearest_neighbour(d, nmu, nsd)
    res = {}
    for m in sorted(set(d["em"].tolist())) + ["all"]:
        sel = np.ones(len(d["em"]), bool) if m == "all" else d["em"] == m
        e_sur = rel_errors(mean[sel], d["ey"][sel], d["ew"][sel], mask)
        e_nn = rel_errors(nn[sel], d["ey"][sel], d["ew"][sel], mask)
        w = d["ew"][sel][(d["ew"][sel] > 0) & mask[None, :]]
        mc_se = np.sqrt(np.maximum(1.0 / w - 1e-4, 0.0))
        trusted = sel & (fb == 0)
        e_tr = rel_errors(mean[trusted], d["ey"][trusted], d["ew"][trusted], mask) if trusted.any() else np.array([])
        res[MODEL_NAMES.get(m, m)] = {
            "median_rel_err_surrogate": float(np.median(e_sur)) if e_sur.size else None,
            "median_rel_err_nearest_neighbour": float(np.median(e_nn)) if e_nn.size else None,
            "median_rel_se_mc_reference": float(np.median(mc_se)) if mc_se.size else None,
            "fallback_share": float((fb[sel] != 0).mean()),
            "median_rel_err_trusted": float(np.median(e_tr)) if e_tr.size else None,
            "eval_points": int(sel.sum()),
        }
    return res


def inference_rate(sg, n=65536, repeats=5):
    """(queries/s, seconds per call) of gs_surrogate_predict on host arrays."""
    rng = np.random.default_rng(0)
    lo, hi = sg.get_normalization()[2:]
    x = (lo + (hi - lo) * rng.uniform(size=(n, len(lo)))).astype(np.float32)
    sg.predict(x)   # warm
    t0 = time.perf_counter()
    for _ in range(repeats):
        sg.predict(x)
    dt = (time.perf_counter() - t0) / repeats
    return n / dt, dt


def inference_rate_device(sg, n=65536, repeats=5):
    """(queries/s, seconds per call) of gs_surrogate_predict_device on
    device-resident inputs and outputs (mean and flags, as a serving loop
    reads them)."""
    from .devmem import DeviceArray
    rng = np.random.default_rng(0)
    lo, hi = sg.get_normalization()[2:]
    x = DeviceArray.from_host((lo + (hi - lo) * rng.uniform(size=(n, len(lo)))).astype(np.float32))
    mean = DeviceArray((n, surrogate.N_OUT), np.float32)
    fb = DeviceArray((n,), np.uint8)
    sg.predict_device(x, mean, None, fb)   # warm
    t0 = time.perf_counter()
    for _ in range(repeats):
        sg.predict_device(x, mean, None, fb)
    dt = (time.perf_counter() - t0) / repeats
    for b in (x, mean, fb):
        b.free()
    return n / dt, dt


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--data", nargs="+", required=True)
    ap.add_argument("--epochs", type=int, default=None, help="default 1500, or 0 with --load")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--save", default="")
    ap.add_argument("--load", default="", help="start from a saved ensemble")
    ap.add_argument("--json", default="")
    a = ap.parse_args(argv)
    d = load(a.data)
    print("training spectra %d, evaluation spectra %d" % (len(d["x"]), len(d["ex"])), flush=True)
    sg = surrogate.Surrogate(0, batch=a.batch, lr=a.lr, seed=a.seed)
    if a.epochs is None:
        a.epochs = 0 if a.load else 1500
    if a.load:
        sg.load(a.load)
    t0 = time.perf_counter()
    done, loss = 0, None
    while done < a.epochs:
        step = min(100, a.epochs - done)
        loss = sg.train(d["x"], d["y"], d["w"], step)
        done += step
        print("epoch %d loss %.4e (%.1f s)" % (done, loss, time.perf_counter() - t0), flush=True)
    train_s = time.perf_counter() - t0
    res = evaluate(sg, d)
    rate, dt = inference_rate(sg)
    out = {"epochs": a.epochs, "train_seconds": train_s, "final_loss": loss, "inference_spectra_per_s": rate,
           "inference_batch_seconds": dt, "per_model": res}
    print(json.dumps(out, indent=1), flush=True)
    if a.save:
        sg.save(a.save)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
