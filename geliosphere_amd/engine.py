"""ctypes front end of libgeliosphere_amd.so (the gfx950 trajectory engine).

The library is the product: this module only loads it and moves plain C
structs across the C ABI declared in include/gs_api.h.  There is no CPU
fallback: if the library or a HIP device is missing, the calls raise.
"""
import ctypes
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgeliosphere_amd.so")
_lib = None

EXPORTS = ("gs_abi_version", "gs_create", "gs_destroy", "gs_last_error", "gs_default_params",
           "gs_validate_params", "gs_run_plain", "gs_run_plain_noalloc", "gs_run_batch", "gs_run_batch_records", "gs_set_batch_budget",
           "gs_set_plain_limits",
           "gs_set_launch", "gs_set_launch_shape", "gs_get_launch_shape", "gs_set_order", "gs_autotune",
           "gs_autotune_shape", "gs_autotune_save", "gs_autotune_load", "gs_last_run_stats", "gs_debug_math",
           "gs_debug_normals", "gs_host_math", "gs_spectrum_merge")

ORDER_ID, ORDER_ENERGY = 0, 1

DBG_RCP, DBG_RSQRT, DBG_SQRT, DBG_LOG, DBG_EXP, DBG_SIN_TURN, DBG_COS_TURN, DBG_LOG_U32, \
    DBG_RAW_RCP, DBG_RAW_RSQ, DBG_RAW_SQRT, DBG_SIN, DBG_COS, DBG_ATAN, DBG_ATAN_FRAC, \
    DBG_SHEET_F, DBG_SHEET_G, DBG_RSQRT4 = range(18)
TWO_ARG = (DBG_ATAN_FRAC, DBG_SHEET_F, DBG_SHEET_G)


class GsError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("gs status %d: %s" % (status, msg))
        self.status = status


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgeliosphere_amd.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        L.gs_abi_version.restype = ctypes.c_int
        L.gs_create.argtypes = [P(vp), ctypes.c_int]
        L.gs_create.restype = ctypes.c_int
        L.gs_destroy.argtypes = [vp]
        L.gs_destroy.restype = None
        L.gs_last_error.argtypes = [vp]
        L.gs_last_error.restype = ctypes.c_char_p
        L.gs_default_params.argtypes = [P(abi.GsParams), ctypes.c_int32]
        L.gs_default_params.restype = ctypes.c_int
        L.gs_validate_params.argtypes = [P(abi.GsParams), ctypes.c_char_p, ctypes.c_uint32]
        L.gs_validate_params.restype = ctypes.c_int
        L.gs_run_plain.argtypes = [vp, P(abi.GsParams), P(abi.GsSpectrum), P(abi.GsExitRecord),
                                   ctypes.c_uint64, P(ctypes.c_uint64)]
        L.gs_run_plain.restype = ctypes.c_int
        L.gs_run_plain_noalloc.argtypes = [vp, P(abi.GsParams), P(abi.GsSpectrum)]
        L.gs_run_plain_noalloc.restype = ctypes.c_int
        L.gs_run_batch.argtypes = [vp, P(abi.GsParams), ctypes.c_uint32, P(abi.GsSpectrum)]
        L.gs_run_batch.restype = ctypes.c_int
        L.gs_run_batch_records.argtypes = [vp, P(abi.GsParams), ctypes.c_uint32, P(abi.GsSpectrum),
                                           P(P(abi.GsExitRecord)), P(ctypes.c_uint64), P(ctypes.c_uint64)]
        L.gs_run_batch_records.restype = ctypes.c_int
        L.gs_set_batch_budget.argtypes = [vp, ctypes.c_uint64]
        L.gs_set_batch_budget.restype = ctypes.c_int
        L.gs_set_plain_limits.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64]
        L.gs_set_plain_limits.restype = ctypes.c_int
        L.gs_set_launch.argtypes = [vp, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32]
        L.gs_set_launch.restype = ctypes.c_int
        L.gs_set_order.argtypes = [vp, ctypes.c_int32]
        L.gs_set_order.restype = ctypes.c_int
        L.gs_autotune.argtypes = [vp, P(abi.GsParams), ctypes.c_uint32, P(ctypes.c_uint32),
                                  P(ctypes.c_double)]
        L.gs_autotune.restype = ctypes.c_int
        L.gs_set_launch_shape.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, P(abi.GsLaunchShape)]
        L.gs_set_launch_shape.restype = ctypes.c_int
        L.gs_get_launch_shape.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, P(abi.GsLaunchShape)]
        L.gs_get_launch_shape.restype = ctypes.c_int
        L.gs_autotune_shape.argtypes = [vp, P(abi.GsParams), ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32,
                                        P(abi.GsLaunchShape), P(ctypes.c_double)]
        L.gs_autotune_shape.restype = ctypes.c_int
        L.gs_autotune_save.argtypes = [vp, ctypes.c_char_p]
        L.gs_autotune_save.restype = ctypes.c_int
        L.gs_autotune_load.argtypes = [vp, ctypes.c_char_p, P(ctypes.c_uint32)]
        L.gs_autotune_load.restype = ctypes.c_int
        L.gs_last_run_stats.argtypes = [vp, P(abi.GsRunStats)]
        L.gs_last_run_stats.restype = ctypes.c_int
        dp = P(ctypes.c_double)
        L.gs_debug_math.argtypes = [vp, ctypes.c_int32, dp, dp, ctypes.c_uint64]
        L.gs_debug_math.restype = ctypes.c_int
        L.gs_debug_normals.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, dp]
        L.gs_debug_normals.restype = ctypes.c_int
        L.gs_host_math.argtypes = [ctypes.c_int32, dp, dp, ctypes.c_uint64]
        L.gs_spectrum_merge.argtypes = [P(abi.GsSpectrum), ctypes.c_uint32, P(abi.GsSpectrum)]
        L.gs_spectrum_merge.restype = ctypes.c_int
        L.gs_host_math.restype = ctypes.c_int
        _lib = L
    return _lib


def host_math(fn, xs):
    """Host evaluation of the math layer's polynomial parts (no device);
    two-argument functions as in Engine.debug_math."""
    m = len(xs)
    n = m // 2 if fn in TWO_ARG else m
    x = (ctypes.c_double * m)(*xs)
    y = (ctypes.c_double * n)()
    rc = lib().gs_host_math(fn, x, y, n)
    if rc:
        raise GsError(rc, "gs_host_math")
    return list(y)


def merge_spectra(parts):
    """gs_spectrum_merge: the spectrum of a run made of these id-range shards
    (host only).  Returns (status, spectrum)."""
    n = len(parts)
    arr = (abi.GsSpectrum * n)(*parts)
    out = abi.GsSpectrum()
    rc = lib().gs_spectrum_merge(arr, n, ctypes.byref(out))
    return rc, out


def validate(p):
    buf = ctypes.create_string_buffer(256)
    rc = lib().gs_validate_params(ctypes.byref(p), buf, 256)
    return rc, buf.value.decode()


def default_params(model):
    p = abi.GsParams()
    rc = lib().gs_default_params(ctypes.byref(p), model)
    if rc:
        raise GsError(rc, "gs_default_params")
    return p


class Engine:
    """One gs_ctx on one HIP device."""

    def __init__(self, device=0):
        self._ctx = ctypes.c_void_p()
        rc = lib().gs_create(ctypes.byref(self._ctx), device)
        if rc:
            raise GsError(rc, "gs_create(device=%d) failed" % device)

    def close(self):
        if self._ctx:
            lib().gs_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            raise GsError(rc, "%s: %s" % (what, lib().gs_last_error(self._ctx).decode()))

    def run_plain(self, p, records=False, rec_cap=None):
        """Returns (spectrum, records or None, n_events)."""
        s = abi.GsSpectrum()
        n = ctypes.c_uint64(0)
        if records:
            cap = rec_cap if rec_cap is not None else max(16, int(2 * p.n_particles))
            recs = (abi.GsExitRecord * cap)()
            rc = lib().gs_run_plain(self._ctx, ctypes.byref(p), ctypes.byref(s), recs, cap, ctypes.byref(n))
            self._check(rc, "gs_run_plain")
            return s, recs[: min(cap, n.value)], n.value
        rc = lib().gs_run_plain(self._ctx, ctypes.byref(p), ctypes.byref(s), None, 0, ctypes.byref(n))
        self._check(rc, "gs_run_plain")
        return s, None, n.value

    def run_plain_noalloc(self, p):
        """gs_run_plain_noalloc: the histogram of one run, computed in the
        device's fixed scratch without any allocation (no exit records)."""
        s = abi.GsSpectrum()
        rc = lib().gs_run_plain_noalloc(self._ctx, ctypes.byref(p), ctypes.byref(s))
        self._check(rc, "gs_run_plain_noalloc")
        return s

    def run_plain_status(self, p):
        """Status code of a run (no raise), for error-path tests."""
        s = abi.GsSpectrum()
        return lib().gs_run_plain(self._ctx, ctypes.byref(p), ctypes.byref(s), None, 0, None)

    def run_batch(self, sims):
        n = len(sims)
        arr = (abi.GsParams * n)(*sims)
        outs = (abi.GsSpectrum * n)()
        rc = lib().gs_run_batch(self._ctx, arr, n, outs)
        self._check(rc, "gs_run_batch")
        return list(outs)

    def run_batch_records(self, sims, rec_caps):
        """(spectra, records per simulation, event counts); rec_caps[i] = 0
        asks no records of simulation i."""
        n = len(sims)
        arr = (abi.GsParams * n)(*sims)
        outs = (abi.GsSpectrum * n)()
        bufs = [(abi.GsExitRecord * c)() if c else None for c in rec_caps]
        ptrs = (ctypes.POINTER(abi.GsExitRecord) * n)(
            *[ctypes.cast(b, ctypes.POINTER(abi.GsExitRecord)) if b is not None else None for b in bufs])
        caps = (ctypes.c_uint64 * n)(*rec_caps)
        nrec = (ctypes.c_uint64 * n)()
        rc = lib().gs_run_batch_records(self._ctx, arr, n, outs, ptrs, caps, nrec)
        self._check(rc, "gs_run_batch_records")
        recs = [list(b[:min(c, k)]) if b is not None else None for b, c, k in zip(bufs, rec_caps, nrec)]
        return list(outs), recs, list(nrec)

    def run_batch_status(self, sims):
        """Status code of a batch (no raise), for error-path tests."""
        n = len(sims)
        arr = (abi.GsParams * n)(*sims)
        outs = (abi.GsSpectrum * max(n, 1))()
        return lib().gs_run_batch(self._ctx, arr, n, outs)

    def set_batch_budget(self, max_events):
        self._check(lib().gs_set_batch_budget(self._ctx, max_events), "gs_set_batch_budget")

    def set_plain_limits(self, max_chunk=0, max_events=0):
        """Particles per plain launch and event records held (0: defaults)."""
        self._check(lib().gs_set_plain_limits(self._ctx, max_chunk, max_events), "gs_set_plain_limits")

    def set_launch(self, model, block_threads=0, blocks_per_cu=0):
        self._check(lib().gs_set_launch(self._ctx, model, block_threads, blocks_per_cu), "gs_set_launch")

    def set_launch_shape(self, model, kernel, block_threads=0, blocks_per_cu=0, chunk=0):
        sh = abi.GsLaunchShape(block_threads, blocks_per_cu, chunk)
        self._check(lib().gs_set_launch_shape(self._ctx, model, kernel, ctypes.byref(sh)), "gs_set_launch_shape")

    def launch_shape(self, model, kernel):
        """(block_threads, blocks_per_cu, chunk) a launch of the kernel would use."""
        sh = abi.GsLaunchShape()
        self._check(lib().gs_get_launch_shape(self._ctx, model, kernel, ctypes.byref(sh)), "gs_get_launch_shape")
        return sh.as_tuple()

    def autotune_shape(self, pilots, kernel, repeats=1):
        """Tune one kernel of the pilots' model; returns ((block, blocks/CU,
        chunk), particle-steps/s).  Plain: one pilot; batch: pilots of one model."""
        n = len(pilots)
        arr = (abi.GsParams * n)(*pilots)
        sh = abi.GsLaunchShape()
        r = ctypes.c_double(0.0)
        self._check(lib().gs_autotune_shape(self._ctx, arr, n, kernel, repeats, ctypes.byref(sh), ctypes.byref(r)),
                    "gs_autotune_shape")
        return sh.as_tuple(), r.value

    def autotune_save(self, path):
        self._check(lib().gs_autotune_save(self._ctx, path.encode()), "gs_autotune_save")

    def autotune_load(self, path):
        """Apply the cache's shapes for this device; returns how many."""
        n = ctypes.c_uint32(0)
        self._check(lib().gs_autotune_load(self._ctx, path.encode(), ctypes.byref(n)), "gs_autotune_load")
        return n.value

    def set_order(self, mode):
        """ORDER_ENERGY (default) or ORDER_ID: the order in which plain runs
        hand particles to the lanes; results do not depend on it."""
        self._check(lib().gs_set_order(self._ctx, mode), "gs_set_order")

    def autotune(self, pilot, repeats=1):
        """Time the pilot at every blocks-per-CU setting; apply and return
        (best blocks per CU, its particle-steps/s)."""
        b = ctypes.c_uint32(0)
        r = ctypes.c_double(0.0)
        self._check(lib().gs_autotune(self._ctx, ctypes.byref(pilot), repeats, ctypes.byref(b),
                                      ctypes.byref(r)), "gs_autotune")
        return b.value, r.value

    def debug_math(self, fn, xs):
        """fn over xs on the device; the two-argument functions (TWO_ARG) take
        xs = first arguments + second arguments (2n values), n results."""
        m = len(xs)
        n = m // 2 if fn in TWO_ARG else m
        x = (ctypes.c_double * m)(*xs)
        y = (ctypes.c_double * n)()
        self._check(lib().gs_debug_math(self._ctx, fn, x, y, n), "gs_debug_math")
        return list(y)

    def debug_normals(self, seed, pid, n):
        y = (ctypes.c_double * n)()
        self._check(lib().gs_debug_normals(self._ctx, seed, pid, n, y), "gs_debug_normals")
        return list(y)

    def last_stats(self):
        st = abi.GsRunStats()
        self._check(lib().gs_last_run_stats(self._ctx, ctypes.byref(st)), "gs_last_run_stats")
        return st
