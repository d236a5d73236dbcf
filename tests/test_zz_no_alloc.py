"""gs_run_plain_noalloc (the entry point that allocates nothing) against
gs_run_plain: identical spectra, bit for bit, on consecutive calls that
share the device's scratch, the second with a smaller input."""
import sys

import pytest

from geliosphere_amd import abi

pytestmark = pytest.mark.gpu

ALL_WEIGHTS = (abi.W_POWERLAW, abi.W_YAMADA, abi.W_WEBER, abi.W_USOSKIN)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(a, b):
    assert a.n_particles == b.n_particles
    assert a.n_events == b.n_events and a.n_outside == b.n_outside
    assert a.total_steps == b.total_steps
    assert a.n_nan == b.n_nan and a.n_step_cap == b.n_step_cap
    assert a.n_acc_inexact == b.n_acc_inexact and a.n_acc_overflow == b.n_acc_overflow
    assert bytes(a) == bytes(b)


CASES = [
    # (label, model, first N, second N, extra parameters)
    ("1D B-p, 1.5e6 then 3e4 (three scratch launches, then one)", abi.MODEL_1D_BP, 1_500_000, 30_000, {}),
    ("1D F-p, 2e7 then 1e6", abi.MODEL_1D_FP, 20_000_000, 1_000_000, {}),
    ("2D Geliosphere, 2e4 then 3e3", abi.MODEL_2D_GELIOSPHERE, 20_000, 3_000, {"tilt_deg": 45.0}),
]


@pytest.mark.parametrize("label,model,n1,n2,extra", CASES, ids=[c[0] for c in CASES])
def test_noalloc_equals_plain(engine, label, model, n1, n2, extra):
    for n in (n1, n2):
        p = abi.make_params(model, N=n, seed=97, weights=ALL_WEIGHTS, **extra)
        say("no-alloc vs plain:", label, "N", n)
        ref, _, _ = engine.run_plain(p)
        got = engine.run_plain_noalloc(p)
        same(got, ref)
        assert got.n_particles == n


def test_noalloc_after_other_calls_and_rejects_oversized_histograms(engine):
    """A call after batches and record runs on the same device (the scratch
    keeps the last call's contents: each call clears what it uses)."""
    p = abi.make_params(abi.MODEL_2D_SOLARPROP, N=5_000, seed=3)
    ref, _, _ = engine.run_plain(p, records=True, rec_cap=10)
    engine.run_batch([abi.make_params(abi.MODEL_1D_BP, N=2_000, seed=4)])
    a = engine.run_plain_noalloc(p)
    b = engine.run_plain_noalloc(p)
    same(a, ref)
    same(b, ref)
    q = abi.make_params(abi.MODEL_1D_BP, N=10)
    q.n_bins = abi.GS_MAX_BINS + 1
    s = abi.GsSpectrum()
    from geliosphere_amd.engine import lib
    assert lib().gs_run_plain_noalloc(engine._ctx, q, s) == abi.GS_ERR_INVALID_ARG
