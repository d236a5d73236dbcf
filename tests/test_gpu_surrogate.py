"""The surrogate's gfx950 kernels (BF16 MFMA, FP32 accumulate) against the
FP64 oracle network, and its training, determinism and fallback logic.
"""
import numpy as np
import pytest

from geliosphere_amd.surrogate import N_IN as NI
from oracle import surrogate_ref as R

pytestmark = pytest.mark.gpu


def make(**kw):
    from geliosphere_amd.surrogate import Surrogate
    return Surrogate(0, **kw)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def reference_grads(s, x, y, w, emulate):
    pm, lay = s.layout()
    P = s.get_params().astype(np.float64)
    mean, sd, _, _ = s.get_normalization()
    xp = R.pad_inputs(x.astype(np.float64), mean.astype(np.float64), sd.astype(np.float64))
    yp = np.zeros((x.shape[0], R.PAD))
    wp = np.zeros((x.shape[0], R.PAD))
    yp[:, :y.shape[1]] = y
    wp[:, :w.shape[1]] = w
    out = []
    for m in range(s.cfg.n_members):
        layers = R.unpack(P[m * pm:(m + 1) * pm], lay)
        loss, grads = R.loss_and_grads(layers, xp, yp, wp, emulate_bf16=emulate)
        out.append((loss, R.flatten(grads, lay, pm)))
    return out


def test_gradients_match_fp64_oracle():
    rng = np.random.default_rng(11)
    s = make(seed=5)
    n = 200
    x = rng.normal(0, 1, (n, NI)).astype(np.float32)
    y = rng.normal(-1, 0.5, (n, 30)).astype(np.float32)
    w = rng.uniform(0, 2, (n, 30)).astype(np.float32)
    w[rng.uniform(size=w.shape) < 0.2] = 0.0
    s.set_normalization(np.zeros(NI), np.ones(NI), np.full(NI, -10.0), np.full(NI, 10.0))
    g, loss = s.gradients(x, y, w)
    pm, _ = s.layout()
    emu = reference_grads(s, x, y, w, emulate=True)
    exact = reference_grads(s, x, y, w, emulate=False)
    for m in range(s.cfg.n_members):
        gm = g[m * pm:(m + 1) * pm].astype(np.float64)
        e_loss, e_g = emu[m]
        x_loss, x_g = exact[m]
        print("member %d: loss gpu %.6f bf16-emulated %.6f fp64 %.6f; grad rel L2 vs emulated %.2e vs fp64 %.2e"
              % (m, loss[m], e_loss, x_loss, rel_l2(gm, e_g), rel_l2(gm, x_g)))
        assert abs(loss[m] - e_loss) <= 2e-3 * abs(e_loss)
        assert rel_l2(gm, e_g) < 1e-2
        assert rel_l2(gm, x_g) < 5e-2
        cos = float(gm @ x_g / (np.linalg.norm(gm) * np.linalg.norm(x_g)))
        assert cos > 0.999


def test_fused_inference_equals_training_forward():
    """predict runs the fused all-layer kernel; the training path runs one
    GEMM per layer.  With one member the prediction is that member's output,
    so taking it as the training target must give a loss and a gradient of
    exactly zero: the two forward passes agree bit for bit.  300 rows leave a
    partial 128-row tile."""
    rng = np.random.default_rng(13)
    s = make(seed=4, n_members=1)
    s.set_normalization(np.zeros(NI), np.ones(NI), np.full(NI, -10.0), np.full(NI, 10.0))
    x = rng.normal(0, 1, (300, NI)).astype(np.float32)
    mean, _, _ = s.predict(x)
    assert np.all(np.isfinite(mean)) and np.abs(mean).max() > 0
    g, loss = s.gradients(x, mean, np.ones_like(mean))
    assert loss[0] == 0.0
    assert np.abs(g).max() == 0.0


def test_predict_matches_oracle_forward():
    rng = np.random.default_rng(12)
    s = make(seed=9)
    s.set_normalization(np.zeros(NI), np.ones(NI), np.full(NI, -10.0), np.full(NI, 10.0))
    x = rng.normal(0, 1, (300, NI)).astype(np.float32)
    mean, sd, fb = s.predict(x)
    pm, lay = s.layout()
    P = s.get_params().astype(np.float64)
    xp = R.pad_inputs(x.astype(np.float64), np.zeros(NI), np.ones(NI))
    outs = np.array([R.forward(R.unpack(P[m * pm:(m + 1) * pm], lay), xp, emulate_bf16=True)[0][:, :30]
                     for m in range(s.cfg.n_members)])
    assert np.max(np.abs(mean - outs.mean(axis=0))) < 2e-3
    assert np.max(np.abs(sd - outs.std(axis=0))) < 2e-3


def synthetic(n, rng):
    x = rng.uniform(-1, 1, (n, NI)).astype(np.float32)
    b = np.arange(30) / 29.0
    y = (np.sin(2.0 * x[:, :1] + b) * 0.5 + x[:, 1:2] * x[:, 2:3] * b - 0.3 * x[:, 3:4] ** 2).astype(np.float32)
    return x, y, np.ones_like(y)


def test_training_reduces_loss_and_generalises():
    rng = np.random.default_rng(13)
    x, y, w = synthetic(8192, rng)
    xt, yt, _ = synthetic(1024, rng)
    s = make(seed=21, batch=512, lr=2e-3)
    first = s.train(x, y, w, 1)
    last = s.train(x, y, w, 60)
    mean, sd, fb = s.predict(xt)
    err = float(np.sqrt(np.mean((mean - yt) ** 2)))
    print("loss first epoch %.4e, after 61 epochs %.4e, held-out rms %.4f, mean ensemble sd %.4f"
          % (first, last, err, float(sd.mean())))
    assert last < first / 20
    assert err < 0.05


def test_training_is_deterministic():
    rng = np.random.default_rng(14)
    x, y, w = synthetic(2048, rng)
    a = make(seed=3, batch=256)
    b = make(seed=3, batch=256)
    a.train(x, y, w, 3)
    b.train(x, y, w, 3)
    assert np.array_equal(a.get_params(), b.get_params())


def test_fallback_flags_and_save_load(tmp_path):
    rng = np.random.default_rng(15)
    x, y, w = synthetic(2048, rng)
    s = make(seed=4, batch=256, std_limit=1e9)
    s.train(x, y, w, 2)
    inside = np.zeros((1, NI), dtype=np.float32)
    outside = np.full((1, NI), 5.0, dtype=np.float32)
    _, _, fb = s.predict(np.vstack([inside, outside]))
    assert fb[0] == 0 and fb[1] & 0x2
    path = tmp_path / "sg.npz"
    s.save(path)
    t = make(seed=99, batch=256, std_limit=0.0)
    t.load(path)
    m1, sd1, _ = s.predict(x[:64])
    m2, sd2, fb2 = t.predict(x[:64])
    assert np.array_equal(m1, m2) and np.array_equal(sd1, sd2)
    assert np.all(fb2 & 0x1)          # std_limit 0: every query falls back
    # without the spread the flags and means are the same
    m3, sd3, fb3 = t.predict(x[:64], want_sd=False)
    assert sd3 is None and np.array_equal(m3, m2) and np.array_equal(fb3, fb2)
    # flags against the rule on the host, with a limit that about half the queries pass
    sel = [c for c in range(30) if (s.cfg.std_outputs >> c) & 1]
    u = make(seed=4, batch=256, std_limit=float(np.median(sd1[:, sel].max(axis=1))))
    u.load(path)
    q = np.vstack([x[:64], outside])
    m4, sd4, fb4 = u.predict(q)
    lo, hi = u.get_normalization()[2:]
    mask = u.cfg.std_outputs
    want = np.array([(0x1 if any((mask >> c) & 1 and not sd4[r, c] <= u.cfg.std_limit for c in range(30)) else 0) |
                     (0x2 if not np.all((q[r] >= lo) & (q[r] <= hi)) else 0) for r in range(len(q))], dtype=np.uint8)
    assert np.array_equal(fb4, want) and 0 < np.count_nonzero(fb4 & 0x1) < len(q)


def test_predict_device_equals_host_predict():
    """gs_surrogate_predict_device (device buffers, no host copies) returns
    the host entry point's means, spreads and flags bit for bit, over more
    rows than one 65,536-row chunk and with optional outputs left out."""
    from geliosphere_amd.devmem import DeviceArray
    rng = np.random.default_rng(16)
    x, y, w = synthetic(2048, rng)
    s = make(seed=5, batch=256)
    s.train(x, y, w, 2)
    sd_all = s.predict(x)[1]
    u = make(seed=5, batch=256, std_limit=float(np.median(sd_all.max(axis=1))))
    u.set_params(s.get_params())
    u.set_normalization(*s.get_normalization())
    q = np.vstack([rng.uniform(-1.2, 1.2, (70_000, NI)).astype(np.float32), np.full((1, NI), 5.0, np.float32)])
    m_h, sd_h, fb_h = u.predict(q)
    xd = DeviceArray.from_host(q)
    sd_d = DeviceArray((len(q), 30), np.float32)
    m_d, sd_d, fb_d = u.predict_device(xd, None, sd_d, None)
    assert np.array_equal(m_d.to_host(), m_h) and np.array_equal(sd_d.to_host(), sd_h)
    assert np.array_equal(fb_d.to_host(), fb_h)
    assert 0 < np.count_nonzero(fb_h & 0x1) < len(q) and fb_h[-1] & 0x2
    x100 = DeviceArray.from_host(q[:100])
    m2, _, fb2 = u.predict_device(x100)
    assert np.array_equal(m2.to_host(), m_h[:100]) and np.array_equal(fb2.to_host(), fb_h[:100])
