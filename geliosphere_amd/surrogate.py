"""Learned surrogate of the Monte Carlo spectra (ctypes front end of
include/gs_surrogate.h) and the host logic around it: features and targets
of a simulation, data set generation with the batched engine, training,
prediction with Monte Carlo fallback.

Inputs (8): the model folded to two flags (2D?, and forward-vs-backward for
1D / Geliosphere-vs-SolarProp for 2D), log10 K0 [cm^2/s], V [km/s] (1D) or
tilt [deg] (2D), polarity, injection r [AU], injection theta [deg], log10 dt
[s] (the 1D F-p spectrum depends on the time step).
Outputs (30): per energy bin ln(mean event weight / J_LIS(T_c)) with the
power-law LIS J = (T + 0.94)^-2.78 at the bin's geometric centre, i.e. the
modulated-to-LIS ratio of the spectrum the engine's histogram estimates.
Loss weights: inverse variance of each target (from the exact sum of w^2),
with a 1 % floor; empty or near-empty bins carry no weight.
"""
import ctypes
import math

import numpy as np

from . import abi
from .engine import lib as _core_lib, GsError

K0_CONV = abi.K0_CM2_TO_AU2
V_CONV = abi.V_KMS_TO_AUS
N_IN, N_OUT = 8, 30
FALLBACK_STD, FALLBACK_DOMAIN = 0x1, 0x2


class SurrogateConfig(ctypes.Structure):
    _fields_ = [("n_members", ctypes.c_uint32), ("n_inputs", ctypes.c_uint32), ("n_outputs", ctypes.c_uint32),
                ("hidden", ctypes.c_uint32), ("n_hidden_layers", ctypes.c_uint32), ("batch", ctypes.c_uint32),
                ("seed", ctypes.c_uint64), ("lr", ctypes.c_float), ("beta1", ctypes.c_float),
                ("beta2", ctypes.c_float), ("eps", ctypes.c_float), ("weight_decay", ctypes.c_float),
                ("std_limit", ctypes.c_float), ("std_outputs", ctypes.c_uint32)]


_bound = False


def lib():
    global _bound
    L = _core_lib()
    if not _bound:
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        fp = P(ctypes.c_float)
        L.gs_surrogate_default_config.argtypes = [P(SurrogateConfig)]
        L.gs_surrogate_create.argtypes = [ctypes.c_int, P(SurrogateConfig), P(vp)]
        L.gs_surrogate_destroy.argtypes = [vp]
        L.gs_surrogate_destroy.restype = None
        L.gs_surrogate_last_error.argtypes = [vp]
        L.gs_surrogate_last_error.restype = ctypes.c_char_p
        L.gs_surrogate_train.argtypes = [vp, fp, fp, fp, ctypes.c_uint32, ctypes.c_uint32, fp]
        L.gs_surrogate_predict.argtypes = [vp, fp, ctypes.c_uint32, fp, fp, P(ctypes.c_uint8)]
        L.gs_surrogate_predict_device.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, vp]
        L.gs_surrogate_gradients.argtypes = [vp, fp, fp, fp, ctypes.c_uint32, fp, fp]
        L.gs_surrogate_layout.argtypes = [vp, P(ctypes.c_uint64), P(ctypes.c_uint64), ctypes.c_uint32]
        L.gs_surrogate_get_params.argtypes = [vp, fp, ctypes.c_uint64]
        L.gs_surrogate_set_params.argtypes = [vp, fp, ctypes.c_uint64]
        L.gs_surrogate_get_normalization.argtypes = [vp, fp, fp, fp, fp]
        L.gs_surrogate_set_normalization.argtypes = [vp, fp, fp, fp, fp]
        for f in ("default_config", "create", "train", "predict", "predict_device", "gradients", "layout", "get_params",
                  "set_params", "get_normalization", "set_normalization"):
            getattr(L, "gs_surrogate_" + f).restype = ctypes.c_int
        _bound = True
    return L


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def window_outputs(lo=0.1, hi=10.0):
    """Bit mask of the energy bins (30 log bins, 0.01-100 GeV) inside [lo, hi] GeV:
    the outputs whose ensemble spread decides the Monte Carlo fallback."""
    mask = 0
    for b in range(N_OUT):
        e0, e1 = 0.01 * 10 ** (4 * b / 30), 0.01 * 10 ** (4 * (b + 1) / 30)
        if e0 >= lo * (1 - 1e-12) and e1 <= hi * (1 + 1e-12):
            mask |= 1 << b
    return mask


def default_config(**over):
    """The C defaults, with the fallback std test limited to 0.1-10 GeV."""
    c = SurrogateConfig()
    lib().gs_surrogate_default_config(ctypes.byref(c))
    c.std_outputs = window_outputs()
    for k, v in over.items():
        setattr(c, k, v)
    return c


class Surrogate:
    """An MLP ensemble on one device."""

    def __init__(self, device=0, **config):
        self.cfg = default_config(**config)
        self._h = ctypes.c_void_p()
        rc = lib().gs_surrogate_create(device, ctypes.byref(self.cfg), ctypes.byref(self._h))
        if rc:
            raise GsError(rc, "gs_surrogate_create")

    def close(self):
        if self._h:
            lib().gs_surrogate_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            raise GsError(rc, "%s: %s" % (what, lib().gs_surrogate_last_error(self._h).decode()))

    def layout(self):
        pm = ctypes.c_uint64()
        n_layers = self.cfg.n_hidden_layers + 1
        buf = (ctypes.c_uint64 * (4 * n_layers))()
        self._check(lib().gs_surrogate_layout(self._h, ctypes.byref(pm), buf, 4 * n_layers), "layout")
        return pm.value, [tuple(buf[4 * l:4 * l + 4]) for l in range(n_layers)]

    def get_params(self):
        pm, _ = self.layout()
        a = np.zeros(pm * self.cfg.n_members, dtype=np.float32)
        self._check(lib().gs_surrogate_get_params(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), a.size),
                    "get_params")
        return a

    def set_params(self, a):
        a, p = _f32(a)
        self._check(lib().gs_surrogate_set_params(self._h, p, a.size), "set_params")

    def get_normalization(self):
        out = [np.zeros(self.cfg.n_inputs, dtype=np.float32) for _ in range(4)]
        self._check(lib().gs_surrogate_get_normalization(self._h, *[_f32(o)[1] for o in out]), "get_normalization")
        return out

    def set_normalization(self, mean, sd, lo, hi):
        arrs = [_f32(v) for v in (mean, sd, lo, hi)]
        self._check(lib().gs_surrogate_set_normalization(self._h, *[a[1] for a in arrs]), "set_normalization")

    def train(self, x, y, w, epochs):
        x, px = _f32(x)
        y, py = _f32(y)
        w, pw = _f32(w)
        loss = ctypes.c_float()
        self._check(lib().gs_surrogate_train(self._h, px, py, pw, x.shape[0], epochs, ctypes.byref(loss)), "train")
        return loss.value

    def gradients(self, x, y, w):
        x, px = _f32(x)
        y, py = _f32(y)
        w, pw = _f32(w)
        g = np.zeros(self.layout()[0] * self.cfg.n_members, dtype=np.float32)
        loss = np.zeros(self.cfg.n_members, dtype=np.float32)
        self._check(lib().gs_surrogate_gradients(self._h, px, py, pw, x.shape[0],
                                                 g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                 loss.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "gradients")
        return g, loss

    def predict(self, x, want_sd=True):
        """(mean, sd, fallback flags) of n queries; sd is None when not wanted,
        which halves the bytes copied back."""
        x, px = _f32(x)
        n = x.shape[0]
        mean = np.zeros((n, self.cfg.n_outputs), dtype=np.float32)
        sd = np.zeros_like(mean) if want_sd else None
        fb = np.zeros(n, dtype=np.uint8)
        psd = sd.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if want_sd else None
        self._check(lib().gs_surrogate_predict(self._h, px, n, mean.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                               psd, fb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))), "predict")
        return mean, sd, fb

    def predict_device(self, x, mean=None, sd=None, fallback=None):
        """gs_surrogate_predict_device on device buffers (devmem.DeviceArray)
        of the surrogate's device: x float32 [n, n_inputs]; mean, sd float32
        [n, n_outputs] and fallback uint8 [n], allocated here when not given
        (sd only on request: pass a buffer).  Returns (mean, sd, fallback)."""
        from .devmem import DeviceArray
        n, no = x.shape[0], self.cfg.n_outputs
        if x.dtype != np.float32 or x.shape != (n, self.cfg.n_inputs):
            raise ValueError("x must be a float32 [n, %d] device array" % self.cfg.n_inputs)
        if mean is None:
            mean = DeviceArray((n, no), np.float32)
        if fallback is None:
            fallback = DeviceArray((n,), np.uint8)
        for t, shape, dt in ((mean, (n, no), np.float32), (sd, (n, no), np.float32), (fallback, (n,), np.uint8)):
            if t is not None and (t.shape != shape or t.dtype != dt):
                raise ValueError("output buffer of shape %s, dtype %s expected" % (shape, np.dtype(dt)))
        self._check(lib().gs_surrogate_predict_device(self._h, x.ptr, n, mean.ptr, sd.ptr if sd is not None else None,
                                                      fallback.ptr), "predict_device")
        return mean, sd, fallback

    def save(self, path):
        mean, sd, lo, hi = self.get_normalization()
        np.savez(path, params=self.get_params(), mean=mean, sd=sd, lo=lo, hi=hi,
                 config=np.array([self.cfg.n_members, self.cfg.hidden, self.cfg.n_hidden_layers]))

    def load(self, path):
        d = np.load(path, allow_pickle=False)
        if d["mean"].shape != (self.cfg.n_inputs,):
            raise ValueError("%s holds a network of %d inputs, this ensemble has %d"
                             % (path, d["mean"].shape[0], self.cfg.n_inputs))
        self.set_params(d["params"])
        self.set_normalization(d["mean"], d["sd"], d["lo"], d["hi"])


# ------------------------------------------------------------------ features and targets

def features(p):
    """The 8 inputs of a parameter block."""
    two_d = p.model in (abi.MODEL_2D_SOLARPROP, abi.MODEL_2D_GELIOSPHERE)
    sub = 1.0 if p.model in (abi.MODEL_1D_BP, abi.MODEL_2D_GELIOSPHERE) else 0.0
    v_or_tilt = math.degrees(p.tilt) if two_d else p.V / V_CONV
    return [1.0 if two_d else 0.0, sub, math.log10(p.K0 / K0_CONV), v_or_tilt,
            p.polarity if two_d else 1.0, p.r_inj, math.degrees(p.theta_inj), math.log10(p.dt)]


def bin_centres(p):
    e = [p.bin_edges[i] for i in range(p.n_bins + 1)]
    return np.sqrt(np.array(e[:-1]) * np.array(e[1:]))


def j_lis(T):
    return (T + 0.94) ** -2.78


def targets(p, s, weight_slot=0, min_count=5, floor=1e-4):
    """(y, w) of a spectrum: ln(mean weight / J_LIS(T_c)) and inverse-variance weights."""
    nb = p.n_bins
    cnt = np.array([s.count[b] for b in range(nb)], dtype=np.float64)
    sw = np.array([s.sum_w[weight_slot][b] for b in range(nb)])
    sw2 = np.array([s.sum_w2[weight_slot][b] for b in range(nb)])
    y = np.zeros(nb)
    w = np.zeros(nb)
    ok = (cnt >= min_count) & (sw > 0)
    mean = np.where(ok, sw / np.maximum(cnt, 1), 1.0)
    y[ok] = np.log(mean[ok] / j_lis(bin_centres(p)[ok]))
    rel_var = np.where(ok, np.maximum(sw2 / np.maximum(cnt, 1) - mean * mean, 0.0) / np.maximum(cnt, 1) /
                       (mean * mean), 1.0)
    w[ok] = 1.0 / (rel_var[ok] + floor)
    return y, w


def generate(engine, sims, weight_slot=0, chunk=64):
    """Run sims with the batched engine; returns (x, y, w) arrays."""
    X, Y, W = [], [], []
    for i in range(0, len(sims), chunk):
        part = sims[i:i + chunk]
        outs = engine.run_batch(part)
        for p, s in zip(part, outs):
            y, w = targets(p, s, weight_slot)
            X.append(features(p))
            Y.append(y)
            W.append(w)
    return np.array(X, dtype=np.float32), np.array(Y, dtype=np.float32), np.array(W, dtype=np.float32)


# Models the ensemble serves.  1D F-p is Monte Carlo only: a 1 AU crossing
# is rare (tens per 1e6 particles), so a training spectrum carries
# percent-level noise in every bin and the members never agree to 1 %.
SURROGATE_MODELS = (abi.MODEL_1D_BP, abi.MODEL_2D_SOLARPROP, abi.MODEL_2D_GELIOSPHERE)


def query(surrogate, engine, sims):
    """Spectra targets for parameter blocks: surrogate mean where it is
    trusted, Monte Carlo (batched engine) where the ensemble disagrees, the
    query is outside the training box or its model is not served by the
    ensemble (1D F-p).  Returns (y [n][30], from_mc [n]); a bin that a Monte
    Carlo run leaves with too few events to estimate is NaN."""
    x = np.array([features(p) for p in sims], dtype=np.float32)
    mean, _, fb = surrogate.predict(x, want_sd=False)
    fb = fb.copy()
    for i, p in enumerate(sims):
        if p.model not in SURROGATE_MODELS:
            fb[i] |= FALLBACK_DOMAIN
    y = mean.astype(np.float64)
    redo = [i for i in range(len(sims)) if fb[i]]
    if redo:
        outs = engine.run_batch([sims[i] for i in redo])
        for i, s in zip(redo, outs):
            yi, wi = targets(sims[i], s)
            yi[wi == 0] = np.nan
            y[i] = yi
    return y, fb != 0
