"""Batch workloads for gs_run_batch.

`sample_batch` draws heterogeneous simulations the way the batch benchmark
defines them: model uniform over the four models, K0 log-uniform 1e22-1e23
cm^2/s, V uniform 300-800 km/s (1D) or tilt uniform 5-75 deg (2D, V = 400
km/s), injection r uniform 1-80 AU, theta uniform 10-170 deg, N log-uniform
1e5-1e7, dt from {50, 100, 500} s.  2D polarity is drawn +-1 (the surrogate
takes it as an input).  The parameter sampler has its own seed; particle
streams are keyed by each simulation's seed.

The reference reads such rows from a CSV (columns year, month, K0, V, dt, N,
r, theta, pathToCustomSettingsFile, name, model; Algorithm/src/BatchRun.cpp:
102-193) and runs them one after another (BatchRun.cpp:50-78).
"""
import math
import random

from . import abi

MODEL_CHOICES = (abi.MODEL_1D_FP, abi.MODEL_1D_BP, abi.MODEL_2D_SOLARPROP, abi.MODEL_2D_GELIOSPHERE)


def sample_one(rng, sim_seed, n_scale=1.0, n_range=(1e5, 1e7), model=None):
    """One simulation; `model` fixes the model instead of drawing it."""
    drawn = rng.choice(MODEL_CHOICES)
    model = drawn if model is None else model
    two_d = model in (abi.MODEL_2D_SOLARPROP, abi.MODEL_2D_GELIOSPHERE)
    K0 = 10.0 ** rng.uniform(22.0, 23.0)
    V = 400.0 if two_d else rng.uniform(300.0, 800.0)
    tilt = rng.uniform(5.0, 75.0) if two_d else 10.0
    r = rng.uniform(1.0, 80.0)
    theta = rng.uniform(10.0, 170.0)
    lo, hi = n_range
    N = int(round(10.0 ** rng.uniform(math.log10(lo), math.log10(hi)) * n_scale))
    dt = rng.choice((50.0, 100.0, 500.0))
    pol = rng.choice((1.0, -1.0)) if two_d else 1.0
    return abi.make_params(model, K0_cm2s=K0, V_kms=V, dt=dt, N=max(N, 1), seed=sim_seed, r_inj=r,
                           theta_deg=theta, tilt_deg=tilt, polarity=pol)


def sample_batch(n=256, seed=2024, n_scale=1.0, first_sim_seed=1):
    """n simulations; simulation i uses particle seed first_sim_seed + i."""
    rng = random.Random(seed)
    return [sample_one(rng, first_sim_seed + i, n_scale) for i in range(n)]
