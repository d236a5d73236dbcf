"""The C-ABI library: it loads, exports every entry point include/gs_api.h
declares, and its host-side logic (validation, defaults, errors without a
device) behaves.  CPU only: no kernel is launched."""
import ctypes
import os
import re

import pytest

from geliosphere_amd import abi, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="gs_api.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gs_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = engine.lib()
    names = declared_functions()
    assert set(names) == set(engine.EXPORTS)
    for n in names:
        assert hasattr(L, n), n


def test_library_exports_every_surrogate_symbol():
    L = engine.lib()
    names = declared_functions("gs_surrogate.h")
    assert "gs_surrogate_predict_device" in names
    for n in names:
        assert hasattr(L, n), n


def test_abi_version_and_struct_sizes():
    assert engine.lib().gs_abi_version() == 1
    assert ctypes.sizeof(abi.GsExitRecord) == 48
    # gs_params: 5 x 4 B + 4 x 4 B + 4 B + 11 x 8 B + 4 x 8 B + 65 x 8 B, no padding
    assert ctypes.sizeof(abi.GsParams) == 4 * 5 + 4 * 4 + 4 + 11 * 8 + 4 * 8 + 65 * 8


def test_default_params_are_the_baseline_configuration():
    p = engine.default_params(abi.MODEL_1D_BP)
    assert p.K0 == pytest.approx(5e22 * 4.4683705e-27)
    assert p.V == pytest.approx(400 * 6.68458712e-9)
    assert p.dt == 500.0 and p.r_inj == 1.0
    assert p.n_bins == 30 and p.injection == abi.INJ_LOG_UNIFORM
    for k in range(31):
        assert p.bin_edges[k] == pytest.approx(0.01 * 10 ** (4 * k / 30), rel=1e-14)
    q = engine.default_params(abi.MODEL_2D_GELIOSPHERE)
    assert q.dt == 50.0 and q.ratio == 0.2
    assert engine.default_params(abi.MODEL_2D_SOLARPROP).ratio == 0.02


@pytest.mark.parametrize("field,value", [
    ("dt", 0.0), ("dt", float("nan")), ("V", -1.0), ("K0", float("inf")), ("n_bins", 0),
    ("n_bins", 65), ("n_weights", 0), ("model", 7), ("injection", 9), ("max_steps", 1 << 33),
    ("T_min", 0.0), ("r_inj", float("nan")),
])
def test_validation_rejects(field, value):
    p = abi.make_params(abi.MODEL_1D_BP)
    setattr(p, field, value)
    rc, msg = engine.validate(p)
    assert rc == abi.GS_ERR_INVALID_ARG and msg != "ok"


def test_validation_rejects_unsorted_edges():
    p = abi.make_params(abi.MODEL_1D_BP)
    p.bin_edges[5] = p.bin_edges[4]
    assert engine.validate(p)[0] == abi.GS_ERR_INVALID_ARG
    assert engine.validate(abi.make_params(abi.MODEL_1D_BP))[0] == abi.GS_OK


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    ctx = ctypes.c_void_p()
    assert engine.lib().gs_create(ctypes.byref(ctx), 0) == abi.GS_ERR_NO_DEVICE
    assert not ctx.value


def test_spectrum_merge_of_shards_equals_whole_run():
    """gs_spectrum_merge (host only) adds id-range shards of one run into the
    whole run's spectrum bit for bit, in any order.  The shards come from
    the oracle, whose spectra have the same exact-sum layout."""
    from oracle import oracle_ffi
    p = abi.make_params(abi.MODEL_1D_BP, N=3000, seed=17, r_inj=40.0,
                        weights=(abi.W_POWERLAW, abi.W_YAMADA, abi.W_WEBER, abi.W_USOSKIN))
    rc, whole, _, _ = oracle_ffi.run(p, n_threads=4)
    assert rc == 0
    cuts = [0, 700, 1900, 3000]
    parts = []
    for a, b in zip(cuts, cuts[1:]):
        rc, s, _, _ = oracle_ffi.run(abi.clone_params(p, first_particle=a, n_particles=b - a), n_threads=4)
        assert rc == 0
        parts.append(s)
    for order in ([0, 1, 2], [2, 0, 1]):
        rc, m = engine.merge_spectra([parts[i] for i in order])
        assert rc == abi.GS_OK
        assert bytes(m) == bytes(whole)
    # a single part is returned as it is; mismatched layouts are refused
    rc, m = engine.merge_spectra([whole])
    assert rc == abi.GS_OK and bytes(m) == bytes(whole)
    other = abi.GsSpectrum()
    ctypes.memmove(ctypes.byref(other), ctypes.byref(parts[0]), ctypes.sizeof(other))
    other.n_bins = 12
    rc, _ = engine.merge_spectra([parts[0], other])
    assert rc == abi.GS_ERR_INVALID_ARG


def test_spectrum_merge_reports_overflow():
    """A carry out of the top limb is counted and reported, not dropped."""
    a = abi.GsSpectrum()
    a.n_bins, a.n_weights = 1, 1
    for i in range(abi.GS_ACC_LIMBS):
        a.acc[0][0][0][i] = 0xFFFFFFFF
    b = abi.GsSpectrum()
    b.n_bins, b.n_weights = 1, 1
    b.acc[0][0][0][0] = 1
    rc, m = engine.merge_spectra([a, b])
    assert rc == abi.GS_ERR_OVERFLOW and m.n_acc_overflow == 1
