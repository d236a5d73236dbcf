"""The committed surrogate data set against the code that made it.

geliosphere_amd/data/surrogate_dataset.npz holds the inputs of training and
evaluation simulations drawn by the batch sampler from fixed seeds
(geliosphere_amd.surrogate_data); its `sets` rows record, per model, the
counts, particle numbers and seed of each run that made it.  Its input rows
must be exactly the features the current code gives those simulations, so a
change to the features or the sampler cannot silently leave a stale data set
(or weights trained on one) behind.
"""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "geliosphere_amd", "data", "surrogate_dataset.npz")


def test_dataset_inputs_are_the_features_of_its_simulations():
    from geliosphere_amd import surrogate, surrogate_data as sd
    d = np.load(DATA, allow_pickle=False)
    tr, ev = sd.redraw(d["sets"])
    assert d["x"].shape == (len(tr), surrogate.N_IN) and d["ex"].shape == (len(ev), surrogate.N_IN)
    assert np.array_equal(np.array([surrogate.features(p) for p in tr], dtype=np.float32), d["x"])
    assert np.array_equal(np.array([surrogate.features(p) for p in ev], dtype=np.float32), d["ex"])
    assert np.array_equal(d["m"], np.array([p.model for p in tr])) and np.array_equal(d["em"], np.array([p.model for p in ev]))
    for m, n_tr, p_tr, n_ev, p_ev, seed in d["sets"].tolist():
        assert d["n_train"][m] == p_tr and d["n_eval"][m] == p_ev
        assert all(p.n_particles == p_tr for p in tr if p.model == m)


def test_weights_match_the_input_width():
    from geliosphere_amd import surrogate
    w = np.load(os.path.join(ROOT, "geliosphere_amd", "data", "surrogate_weights.npz"), allow_pickle=False)
    assert w["mean"].shape == (surrogate.N_IN,) and w["lo"].shape == (surrogate.N_IN,)
    assert np.all(w["sd"] > 0) and np.all(w["lo"] <= w["hi"])
