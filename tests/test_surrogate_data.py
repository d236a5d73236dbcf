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
    # per model, the largest particle count of its parts
    for m in set(d["m"].tolist()) | set(d["em"].tolist()):
        rows = [r for r in d["sets"].tolist() if r[0] == m]
        assert d["n_train"][m] == max(r[2] for r in rows if r[1] > 0 or r[3] > 0)
        assert d["n_eval"][m] == max(r[4] for r in rows)
    # each part's simulations carry its particle counts
    for r in d["sets"].tolist():
        part_tr, part_ev = sd.redraw([r])
        assert all(p.n_particles == r[2] for p in part_tr) and all(p.n_particles == r[4] for p in part_ev)


def test_weights_match_the_input_width():
    from geliosphere_amd import surrogate
    w = np.load(os.path.join(ROOT, "geliosphere_amd", "data", "surrogate_weights.npz"), allow_pickle=False)
    assert w["mean"].shape == (surrogate.N_IN,) and w["lo"].shape == (surrogate.N_IN,)
    assert np.all(w["sd"] > 0) and np.all(w["lo"] <= w["hi"])


def test_merge_joins_rows_and_keeps_provenance(tmp_path):
    """merge() concatenates every row array in order and keeps each part's
    `sets` rows; per-model particle counts are the parts' (one part per model)."""
    from geliosphere_amd import surrogate_data as sd
    parts = []
    for k, m in enumerate((1, 3)):
        n, e = 3 + k, 2
        d = {"x": np.full((n, 8), m, np.float32), "y": np.zeros((n, 30), np.float32),
             "w": np.ones((n, 30), np.float32), "m": np.full(n, m), "ex": np.full((e, 8), -m, np.float32),
             "ey": np.zeros((e, 30), np.float32), "ew": np.ones((e, 30), np.float32), "em": np.full(e, m),
             "n_train": np.array([0, 1000 if m == 1 else 0, 0, 500 if m == 3 else 0]),
             "n_eval": np.array([0, 10000 if m == 1 else 0, 0, 5000 if m == 3 else 0]),
             "sets": np.array([[m, n, 1000 if m == 1 else 500, e, 10000 if m == 1 else 5000, 31]])}
        p = tmp_path / ("part%d.npz" % k)
        np.savez(p, **d)
        parts.append(str(p))
    out = tmp_path / "joined.npz"
    res = sd.merge(parts, str(out))
    j = np.load(out, allow_pickle=False)
    assert j["x"].shape == (7, 8) and list(j["m"]) == [1, 1, 1, 3, 3, 3, 3]
    assert list(j["em"]) == [1, 1, 3, 3] and np.all(j["ex"][:2] == -1)
    assert list(j["n_train"]) == [0, 1000, 0, 500] and list(j["n_eval"]) == [0, 10000, 0, 5000]
    assert j["sets"].tolist() == [[1, 3, 1000, 2, 10000, 31], [3, 4, 500, 2, 5000, 31]]
    assert np.array_equal(res["x"], j["x"])
