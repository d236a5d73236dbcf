"""ctypes binding of oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product path never imports this module.
"""
import ctypes
import os
import subprocess

from geliosphere_amd.abi import GsParams, GsSpectrum, GsExitRecord

# gs_params.flags bits read by the oracle only (oracle.h)
FLAG_PERTURB = 0x80000000
FLAG_NO_NOISE = 0x40000000
FLAG_FAITHFUL = 0x20000000

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        L.oracle_philox4x32_10.argtypes = [P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32)]
        L.oracle_philox4x32_10.restype = None
        L.oracle_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_normal.restype = ctypes.c_double
        L.oracle_injection_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_injection_uniform.restype = ctypes.c_double
        L.oracle_injection_energy.argtypes = [P(GsParams), ctypes.c_uint64]
        L.oracle_injection_energy.restype = ctypes.c_double
        L.oracle_run.argtypes = [P(GsParams), P(GsSpectrum), P(GsExitRecord), ctypes.c_uint64,
                                 P(ctypes.c_uint64), ctypes.c_int]
        L.oracle_run.restype = ctypes.c_int
        L.oracle_particle.argtypes = [P(GsParams), ctypes.c_uint64, P(GsExitRecord), ctypes.c_uint64,
                                      P(ctypes.c_uint64), P(ctypes.c_int)]
        L.oracle_particle.restype = ctypes.c_int64
        L.oracle_particle_state.argtypes = [P(GsParams), ctypes.c_uint64, P(ctypes.c_double),
                                            P(ctypes.c_uint64), P(ctypes.c_int)]
        L.oracle_particle_state.restype = ctypes.c_int64
        L.oracle_event_weight.argtypes = [P(GsParams), ctypes.c_int32, P(GsExitRecord)]
        L.oracle_event_weight.restype = ctypes.c_double
        L.oracle_2d_coeffs.argtypes = [P(GsParams), ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, P(ctypes.c_double)]
        L.oracle_2d_coeffs.restype = ctypes.c_int
        L.oracle_count_flops.argtypes = [P(GsParams), ctypes.c_uint64, P(ctypes.c_double)]
        L.oracle_count_flops.restype = ctypes.c_int
        L.oracle_acc_to_double.argtypes = [P(ctypes.c_uint32)]
        L.oracle_acc_to_double.restype = ctypes.c_double
        L.oracle_acc_add.argtypes = [P(ctypes.c_uint32), ctypes.c_double]
        L.oracle_acc_add.restype = ctypes.c_int
        _lib = L
    return _lib


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, k, o)
    return list(o)


def normal(seed, pid, n):
    return lib().oracle_normal(seed, pid, n)


def injection_energy(p, pid):
    return lib().oracle_injection_energy(ctypes.byref(p), pid)


def default_threads():
    """Host threads for the oracle: all cores, at most 16 (one GPU's share of
    a multi-GPU host)."""
    return max(1, min(16, os.cpu_count() or 1))


def run(p, n_threads=0, records=False, rec_cap=None):
    """Whole run; returns (status, spectrum, records-or-None, n_records)."""
    if n_threads <= 0:
        n_threads = default_threads()
    s = GsSpectrum()
    n = ctypes.c_uint64(0)
    recs = None
    if records:
        cap = rec_cap if rec_cap is not None else max(16, 4 * p.n_particles)
        recs = (GsExitRecord * cap)()
        rc = lib().oracle_run(ctypes.byref(p), ctypes.byref(s), recs, cap, ctypes.byref(n), n_threads)
        return rc, s, recs[: min(cap, n.value)], n.value
    rc = lib().oracle_run(ctypes.byref(p), ctypes.byref(s), None, 0, None, n_threads)
    return rc, s, None, 0


def particle(p, pid, cap=4096):
    recs = (GsExitRecord * cap)()
    steps = ctypes.c_uint64(0)
    status = ctypes.c_int(0)
    n = lib().oracle_particle(ctypes.byref(p), pid, recs, cap, ctypes.byref(steps), ctypes.byref(status))
    return list(recs[: min(max(n, 0), cap)]), steps.value, status.value


def particle_state(p, pid):
    """(state dict, steps, status, n_events) of one particle's last state."""
    st = (ctypes.c_double * 5)()
    steps = ctypes.c_uint64(0)
    status = ctypes.c_int(0)
    n = lib().oracle_particle_state(ctypes.byref(p), pid, st, ctypes.byref(steps), ctypes.byref(status))
    return ({"r": st[0], "theta": st[1], "T": st[2], "p": st[3], "sumac": st[4]},
            steps.value, status.value, n)


def event_weight(p, kind, ev):
    return lib().oracle_event_weight(ctypes.byref(p), kind, ctypes.byref(ev))


def coeffs_2d(p, r, theta, T):
    o = (ctypes.c_double * 16)()
    rc = lib().oracle_2d_coeffs(ctypes.byref(p), r, theta, T, o)
    assert rc == 0
    return list(o)


def count_flops(p, n_particles):
    o = (ctypes.c_double * 5)()
    rc = lib().oracle_count_flops(ctypes.byref(p), n_particles, o)
    assert rc == 0
    return {"flops_per_step": o[0], "trans_per_step": o[1], "divs_per_step": o[2],
            "sqrts_per_step": o[3], "steps": o[4]}


def acc_add(limbs, v):
    return lib().oracle_acc_add(limbs, v)


def acc_to_double(limbs):
    return lib().oracle_acc_to_double(limbs)
