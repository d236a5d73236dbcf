"""The committed surrogate data set against the code that made it.

geliosphere_amd/data/surrogate_dataset.npz holds the inputs of 4 x 1000
training and 4 x 40 evaluation simulations drawn by the batch sampler from
fixed seeds (geliosphere_amd.surrogate_data, seed 2026).  Its input rows must
be exactly the features the current code gives those simulations, so a change
to the features cannot silently leave a stale data set (or weights trained on
one) behind.
"""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "geliosphere_amd", "data", "surrogate_dataset.npz")


def test_dataset_inputs_are_the_features_of_its_simulations():
    from geliosphere_amd import surrogate, surrogate_data as sd
    d = np.load(DATA, allow_pickle=False)
    x, ex = [], []
    for m in sd.MODELS:
        tr = sd.sims_for(m, 1000, int(d["n_train"][m]), 2026, 1_000_000 * (m + 1))
        ev = sd.sims_for(m, 40, int(d["n_eval"][m]), 2027, 50_000_000 * (m + 1))
        x += [surrogate.features(p) for p in tr]
        ex += [surrogate.features(p) for p in ev]
    assert d["x"].shape == (4000, surrogate.N_IN) and d["ex"].shape == (160, surrogate.N_IN)
    assert np.array_equal(np.array(x, dtype=np.float32), d["x"])
    assert np.array_equal(np.array(ex, dtype=np.float32), d["ex"])
    assert np.array_equal(d["m"], np.repeat(np.arange(4), 1000)) and np.array_equal(d["em"], np.repeat(np.arange(4), 40))


def test_weights_match_the_input_width():
    from geliosphere_amd import surrogate
    w = np.load(os.path.join(ROOT, "geliosphere_amd", "data", "surrogate_weights.npz"), allow_pickle=False)
    assert w["mean"].shape == (surrogate.N_IN,) and w["lo"].shape == (surrogate.N_IN,)
    assert np.all(w["sd"] > 0) and np.all(w["lo"] <= w["hi"])
