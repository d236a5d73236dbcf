"""ctypes mirror of include/gs_api.h and parameter helpers.

Unit conversions follow the reference's input layer:
K0 [cm^2/s] * 4.4683705e-27 -> AU^2/s (Input/src/InputValidation.cpp:76),
V [km/s] * 6.68458712e-9 -> AU/s (Input/src/InputValidation.cpp:93).
"""
import ctypes
import math

GS_MAX_BINS = 64
GS_MAX_WEIGHTS = 4
GS_ACC_LIMBS = 16
GS_ACC_LSB_EXP = -320

GS_OK = 0
GS_ERR_INVALID_ARG = 1
GS_ERR_HIP = 2
GS_ERR_OUT_OF_MEMORY = 3
GS_ERR_OVERFLOW = 4
GS_ERR_NO_DEVICE = 5
GS_ERR_IO = 6
GS_ERR_UNSUPPORTED = 7

MODEL_1D_FP = 0
MODEL_1D_BP = 1
MODEL_2D_SOLARPROP = 2
MODEL_2D_GELIOSPHERE = 3
MODEL_NAMES = {MODEL_1D_FP: "1D Fp", MODEL_1D_BP: "1D Bp",
               MODEL_2D_SOLARPROP: "2D SolarProp-like",
               MODEL_2D_GELIOSPHERE: "2D Geliosphere"}

INJ_LOG_UNIFORM = 0
INJ_GRID = 1
INJ_SOLARPROP_BINS = 2
INJ_FIXED = 3

W_POWERLAW = 0
W_YAMADA = 1
W_WEBER = 2
W_USOSKIN = 3
W_BURGER = 4

FLAG_SOLARPROP_GPU_DRIFT = 0x1

K0_CM2_TO_AU2 = 4.4683705e-27
V_KMS_TO_AUS = 6.68458712e-9


class GsParams(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int32),
        ("injection", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("n_bins", ctypes.c_uint32),
        ("n_weights", ctypes.c_uint32),
        ("weights", ctypes.c_int32 * GS_MAX_WEIGHTS),
        ("grid_n", ctypes.c_uint32),
        ("K0", ctypes.c_double),
        ("V", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("r_inj", ctypes.c_double),
        ("theta_inj", ctypes.c_double),
        ("tilt", ctypes.c_double),
        ("polarity", ctypes.c_double),
        ("ratio", ctypes.c_double),
        ("delta0", ctypes.c_double),
        ("T_min", ctypes.c_double),
        ("T_max", ctypes.c_double),
        ("n_particles", ctypes.c_uint64),
        ("first_particle", ctypes.c_uint64),
        ("seed", ctypes.c_uint64),
        ("max_steps", ctypes.c_uint64),
        ("bin_edges", ctypes.c_double * (GS_MAX_BINS + 1)),
    ]


class GsSpectrum(ctypes.Structure):
    _fields_ = [
        ("n_bins", ctypes.c_uint32),
        ("n_weights", ctypes.c_uint32),
        ("count", ctypes.c_uint64 * GS_MAX_BINS),
        ("sum_w", (ctypes.c_double * GS_MAX_BINS) * GS_MAX_WEIGHTS),
        ("sum_w2", (ctypes.c_double * GS_MAX_BINS) * GS_MAX_WEIGHTS),
        ("n_particles", ctypes.c_uint64),
        ("n_events", ctypes.c_uint64),
        ("n_outside", ctypes.c_uint64),
        ("total_steps", ctypes.c_uint64),
        ("n_nan", ctypes.c_uint64),
        ("n_step_cap", ctypes.c_uint64),
        ("n_acc_inexact", ctypes.c_uint64),
        ("n_acc_overflow", ctypes.c_uint64),
        ("acc", (((ctypes.c_uint32 * GS_ACC_LIMBS) * GS_MAX_BINS) * 2) * GS_MAX_WEIGHTS),
    ]


class GsExitRecord(ctypes.Structure):
    _fields_ = [
        ("t_inj", ctypes.c_double),
        ("t_exit", ctypes.c_double),
        ("r", ctypes.c_double),
        ("theta", ctypes.c_double),
        ("sumac", ctypes.c_double),
        ("id", ctypes.c_uint64),
    ]


KERNEL_PLAIN = 0
KERNEL_BATCH = 1


class GsLaunchShape(ctypes.Structure):
    _fields_ = [
        ("block_threads", ctypes.c_uint32),
        ("blocks_per_cu", ctypes.c_uint32),
        ("chunk", ctypes.c_uint32),
    ]

    def as_tuple(self):
        return (self.block_threads, self.blocks_per_cu, self.chunk)


class GsRunStats(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("lane_cycles", ctypes.c_uint64),
        ("active_cycles", ctypes.c_uint64),
        ("grid_blocks", ctypes.c_uint32),
        ("block_threads", ctypes.c_uint32),
        ("max_particle_steps", ctypes.c_uint64),
    ]


def log_edges(n_bins=30, t_min=0.01, t_max=100.0):
    """n_bins log-spaced edges t_min * (t_max/t_min)^(k/n_bins), k = 0..n_bins."""
    return [t_min * (t_max / t_min) ** (k / n_bins) for k in range(n_bins + 1)]


def make_params(model, *, K0_cm2s=5e22, V_kms=400.0, dt=None, N=1000, seed=1,
                r_inj=1.0, theta_deg=90.0, tilt_deg=10.0, polarity=1.0, ratio=None,
                delta0=8.7e-5, injection=INJ_LOG_UNIFORM, T_min=0.01, T_max=100.0,
                grid_n=10000, edges=None, weights=(W_POWERLAW,), first=0, max_steps=0,
                flags=0):
    """Parameters of one run.  Defaults: K0 = 5e22 cm^2/s, V = 400 km/s,
    dt = 500 s (1D) or 50 s (2D), log-uniform 0.01-100 GeV, 30 log bins."""
    p = GsParams()
    p.model = model
    p.injection = injection
    p.flags = flags
    if edges is None:
        edges = log_edges(30, 0.01, 100.0)
    p.n_bins = len(edges) - 1
    for i, e in enumerate(edges):
        p.bin_edges[i] = e
    p.n_weights = len(weights)
    for i, w in enumerate(weights):
        p.weights[i] = w
    p.grid_n = grid_n
    p.K0 = K0_cm2s * K0_CM2_TO_AU2
    p.V = V_kms * V_KMS_TO_AUS
    if dt is None:
        dt = 500.0 if model in (MODEL_1D_FP, MODEL_1D_BP) else 50.0
    p.dt = dt
    p.r_inj = r_inj
    p.theta_inj = math.radians(theta_deg)
    p.tilt = math.radians(tilt_deg)
    p.polarity = polarity
    if ratio is None:
        ratio = 0.02 if model == MODEL_2D_SOLARPROP else 0.2
    p.ratio = ratio
    p.delta0 = delta0
    p.T_min = T_min
    p.T_max = T_max
    p.n_particles = N
    p.first_particle = first
    p.seed = seed
    p.max_steps = max_steps
    return p


def clone_params(p, **changes):
    q = GsParams()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(GsParams))
    for k, v in changes.items():
        setattr(q, k, v)
    return q


def spectrum_arrays(s):
    """(count[nb], sum_w[nw][nb], sum_w2[nw][nb]) as Python lists."""
    nb, nw = s.n_bins, s.n_weights
    count = [s.count[b] for b in range(nb)]
    sw = [[s.sum_w[k][b] for b in range(nb)] for k in range(nw)]
    sw2 = [[s.sum_w2[k][b] for b in range(nb)] for k in range(nw)]
    return count, sw, sw2


def acc_int(s, k, which, b):
    """Exact accumulator value as a Python integer (units of 2^GS_ACC_LSB_EXP)."""
    limbs = s.acc[k][which][b]
    v = 0
    for i in reversed(range(GS_ACC_LIMBS)):
        v = (v << 32) | limbs[i]
    return v
