"""The FP64 surrogate oracle (oracle/surrogate_ref.py) and the surrogate's
host logic, on the CPU.

The oracle's backward pass is pinned by central finite differences of its
own loss; bf16 rounding by hand-checked values; features and targets of a
spectrum by their definitions on constructed inputs.
"""
import math

import numpy as np
import pytest

from geliosphere_amd import abi
from oracle import surrogate_ref as R


def small_layout(hidden=16, n_hidden=2):
    dims = [R.PAD] + [hidden] * n_hidden + [R.PAD]
    lay, off = [], 0
    for l in range(len(dims) - 1):
        o, i = dims[l + 1], dims[l]
        lay.append((off, off + o * i, o, i))
        off += o * i + o
    return lay, off


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    lay, size = small_layout()
    p = rng.normal(0, 0.4, size)
    x = rng.normal(0, 1, (9, R.PAD))
    y = rng.normal(0, 1, (9, R.PAD))
    w = rng.uniform(0, 1, (9, R.PAD))
    w[:, 30:] = 0.0
    loss, grads = R.loss_and_grads(R.unpack(p, lay), x, y, w)
    g = R.flatten(grads, lay, size)
    idx = rng.choice(size, 60, replace=False)
    for k in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_and_grads(R.unpack(pp, lay), x, y, w)[0] - R.loss_and_grads(R.unpack(pm, lay), x, y, w)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
    # a planted sign error in the output gradient would flip every dW of the last layer
    assert loss > 0


def test_silu_derivative():
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    fd = (R.silu(z + h) - R.silu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - R.dsilu(z))) < 1e-8


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 1.0 + 2 ** -9])
    r = R.bf16_round(v)
    # 8 significant bits: 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 -> 1 + 2^-6 (tie to even)
    assert list(r) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5, 1.0]


def test_features_of_parameter_blocks():
    from geliosphere_amd.surrogate import features
    p = abi.make_params(abi.MODEL_1D_BP, K0_cm2s=2e22, V_kms=550.0, r_inj=3.0, theta_deg=40.0)
    f = features(p)
    assert f[:2] == [0.0, 1.0] and f[2] == pytest.approx(math.log10(2e22), abs=1e-12)
    assert f[3] == pytest.approx(550.0, rel=1e-12) and f[4] == 1.0 and f[5] == 3.0 and f[6] == pytest.approx(40.0)
    assert len(f) == 8 and f[7] == pytest.approx(math.log10(p.dt), abs=1e-12)
    assert features(abi.make_params(abi.MODEL_1D_FP, dt=50.0))[7] == pytest.approx(math.log10(50.0), abs=1e-12)
    q = abi.make_params(abi.MODEL_2D_SOLARPROP, tilt_deg=33.0, polarity=-1.0)
    g = features(q)
    assert g[:2] == [1.0, 0.0] and g[3] == pytest.approx(33.0) and g[4] == -1.0
    assert features(abi.make_params(abi.MODEL_2D_GELIOSPHERE))[:2] == [1.0, 1.0]
    assert features(abi.make_params(abi.MODEL_1D_FP))[:2] == [0.0, 0.0]


def test_targets_of_a_spectrum():
    from geliosphere_amd.surrogate import targets, j_lis, bin_centres
    p = abi.make_params(abi.MODEL_1D_BP)
    s = abi.GsSpectrum()
    s.n_bins, s.n_weights = 30, 1
    s.count[3], s.sum_w[0][3], s.sum_w2[0][3] = 100, 50.0, 30.0     # mean 0.5, E[w^2] 0.3
    s.count[4], s.sum_w[0][4], s.sum_w2[0][4] = 2, 1.0, 0.5         # too few events
    y, w = targets(p, s)
    tc = bin_centres(p)[3]
    assert y[3] == pytest.approx(math.log(0.5 / j_lis(tc)), rel=1e-12)
    rel_var = (0.3 - 0.25) / 100 / 0.25
    assert w[3] == pytest.approx(1.0 / (rel_var + 1e-4), rel=1e-12)
    assert w[4] == 0.0 and w[0] == 0.0
