"""The engine's autotuner over every model: the launch shape (workgroup
size, blocks per CU, ids per queue grab) of each model's plain and batch
trajectory kernel, measured on this device and written to the autotuner's
cache (gs_autotune_save), which the engine applies with gs_autotune_load.

    python3 -m geliosphere_amd.autotune [--out FILE] [--repeats 2]

Plain kernels are tuned on one simulation of the model (the benchmark's
parameters: K0 = 5e22 cm^2/s, V = 400 km/s, log-uniform 0.01-100 GeV) large
enough that throughput, not the longest trajectory, sets its time: 1D B-p
2e7 particles, 2D 1e7 (the 2D sweep's size; at 2e6, about ten particles per
lane, the longest trajectories still set a third of the time and the search
then favours fewer lanes), 1D F-p 1e8 (its runs are always set by their
longest trajectory).  Batch kernels are tuned on 16 simulations of the model
drawn by the batch benchmark's sampler at 10 % of its particle counts.  The
default cache is geliosphere_amd/data/autotune.txt.
"""
import argparse
import os
import random
import sys
import time

from . import abi, batch
from .engine import Engine

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_CACHE = os.path.join(HERE, "data", "autotune.txt")
PLAIN_N = {abi.MODEL_1D_FP: 100_000_000, abi.MODEL_1D_BP: 20_000_000,
           abi.MODEL_2D_SOLARPROP: 10_000_000, abi.MODEL_2D_GELIOSPHERE: 10_000_000}
BATCH_SCALE = 0.1


def batch_pilots(model, n=16, seed=77):
    rng = random.Random(seed + model)
    return [batch.sample_one(rng, 500 + i, n_scale=BATCH_SCALE, model=model) for i in range(n)]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_CACHE)
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--models", default="0,1,2,3")
    a = ap.parse_args(argv)
    os.environ.setdefault("GS_AUTOTUNE_LOG", "1")   # a line per measured shape (stderr)
    eng = Engine(0)
    for m in [int(v) for v in a.models.split(",")]:
        t0 = time.time()
        p = abi.make_params(m, N=PLAIN_N[m], seed=11)
        sh, rate = eng.autotune_shape([p], abi.KERNEL_PLAIN, a.repeats)
        print("%-18s plain: %d threads, %d blocks/CU, grab %d: %.3e particle-steps/s (%.0f s)"
              % (abi.MODEL_NAMES[m], sh[0], sh[1], sh[2], rate, time.time() - t0), flush=True)
        t0 = time.time()
        sh, rate = eng.autotune_shape(batch_pilots(m), abi.KERNEL_BATCH, a.repeats)
        print("%-18s batch: %d threads, %d blocks/CU, grab %d: %.3e particle-steps/s (%.0f s)"
              % (abi.MODEL_NAMES[m], sh[0], sh[1], sh[2], rate, time.time() - t0), flush=True)
    eng.autotune_save(a.out)
    print("wrote", a.out, flush=True)
    eng.close()


if __name__ == "__main__":
    sys.exit(main())
