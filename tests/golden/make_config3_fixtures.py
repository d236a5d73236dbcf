"""Writes tests/golden/config3_oracle_spectra.json: the oracle's spectra for
the twelve 2D runs of the baseline's third configuration (SolarProp-like and
Geliosphere; tilt 10, 45, 75 deg; polarity +1 and -1; r = 1 AU, theta = 90
deg, K0 = 5e22 cm^2/s, V = 400 km/s, dt = 50 s, log-uniform 0.01-100 GeV,
30 log bins), N = 40,000 particles each, with the power-law, Webber-Higbie,
Usoskin and Burger weights.  Every run has its own particle seed, so that the
twelve spectra's Monte Carlo fluctuations are independent of each other.  The
GPU statistical parity tests compare independent-seed GPU runs with these
spectra.

Run from the repository root:  python3 tests/golden/make_config3_fixtures.py
(about 18 minutes on 8 cores)."""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from geliosphere_amd import abi  # noqa: E402
from oracle import oracle_ffi  # noqa: E402

N = 40_000
SEED = 3_000_003
WEIGHTS = (abi.W_POWERLAW, abi.W_WEBER, abi.W_USOSKIN, abi.W_BURGER)
OUT = os.path.join(ROOT, "tests", "golden", "config3_oracle_spectra.json")


def cases():
    for model in (abi.MODEL_2D_SOLARPROP, abi.MODEL_2D_GELIOSPHERE):
        for tilt in (10.0, 45.0, 75.0):
            for pol in (1.0, -1.0):
                yield model, tilt, pol


def main():
    threads = int(os.environ.get("ORACLE_THREADS", "8"))
    out = {"N": N, "seed": SEED, "weights": list(WEIGHTS), "cases": []}
    for i, (model, tilt, pol) in enumerate(cases()):
        p = abi.make_params(model, N=N, seed=SEED + 1000 * i, tilt_deg=tilt, polarity=pol, weights=WEIGHTS)
        t = time.time()
        rc, s, _, _ = oracle_ffi.run(p, n_threads=threads)
        assert rc == 0 and s.n_particles == N
        out["cases"].append({
            "model": model, "tilt_deg": tilt, "polarity": pol, "seed": p.seed,
            "count": [s.count[b] for b in range(30)],
            "sum_w": [[s.sum_w[k][b] for b in range(30)] for k in range(len(WEIGHTS))],
            "sum_w2": [[s.sum_w2[k][b] for b in range(30)] for k in range(len(WEIGHTS))],
            "n_events": s.n_events, "n_outside": s.n_outside, "total_steps": s.total_steps,
        })
        print("%-18s tilt %4.0f pol %+.0f  %.0f s  steps/particle %.0f" % (
            abi.MODEL_NAMES[model], tilt, pol, time.time() - t, s.total_steps / N), flush=True)
        with open(OUT, "w") as f:
            json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
