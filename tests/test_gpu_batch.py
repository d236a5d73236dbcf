"""Batched Monte Carlo through the C ABI (gs_run_batch).

Every simulation of a batch must come out identical, bit for bit, to the
same simulation run alone with gs_run_plain: particles are keyed by
(seed, id), sums are exact integers, and a batch only changes which lane
integrates which particle and when.  The reference runs the rows of a batch
one after the other (Algorithm/src/BatchRun.cpp:50-78); its results per row
are what gs_run_plain gives.
"""
import math
import random
import sys

import pytest

from geliosphere_amd import abi

pytestmark = pytest.mark.gpu

ALL_WEIGHTS = (abi.W_POWERLAW, abi.W_YAMADA, abi.W_USOSKIN, abi.W_BURGER)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sample_sims(n, seed, n_scale=1.0):
    """A mixed batch in the spirit of the reference's batch CSV rows
    (year, month, K0, V, dt, N, r, theta, model): random model, K0, V or
    tilt, r, theta, dt and N; random binning and weights exercise the
    per-simulation histogram layout."""
    rng = random.Random(seed)
    sims = []
    for i in range(n):
        model = rng.randrange(4)
        two_d = model >= 2
        K0 = 10 ** rng.uniform(22, 23)
        V = rng.uniform(300, 800)
        dt = rng.choice((50.0, 100.0, 500.0)) if not two_d else rng.choice((50.0, 100.0))
        base = {0: 60_000, 1: 30_000, 2: 2_500, 3: 2_000}[model]
        N = int(base * n_scale * 10 ** rng.uniform(-1, 0))
        nb = rng.choice((10, 30, 64))
        k = rng.randint(1, 4)
        p = abi.make_params(model, K0_cm2s=K0, V_kms=V, dt=dt, N=N, seed=1000 + i,
                            r_inj=rng.uniform(1.0, 30.0),
                            theta_deg=rng.uniform(10.0, 170.0), tilt_deg=rng.uniform(5.0, 75.0),
                            polarity=rng.choice((1.0, -1.0)),
                            edges=abi.log_edges(nb), weights=tuple(rng.sample(ALL_WEIGHTS, k)))
        sims.append(p)
    return sims


def assert_batch_equals_plain(engine, sims, outs):
    for i, (p, b) in enumerate(zip(sims, outs)):
        a, _, _ = engine.run_plain(p)
        if bytes(a) != bytes(b):
            say("simulation %d (model %d, N=%d) differs: steps %d vs %d, events %d vs %d"
                % (i, p.model, p.n_particles, a.total_steps, b.total_steps, a.n_events, b.n_events))
        assert bytes(a) == bytes(b), "simulation %d" % i


def test_batch_mixed_models_bit_identical_to_plain(engine):
    sims = sample_sims(24, seed=1)
    sims.append(abi.make_params(abi.MODEL_1D_BP, N=0, seed=3))          # empty simulation
    sims.append(abi.make_params(abi.MODEL_2D_GELIOSPHERE, N=1, seed=4))  # one particle
    say("batch of", len(sims), "models", [p.model for p in sims])
    outs = engine.run_batch(sims)
    st = engine.last_stats()
    say("batch kernel_ms %.2f total_ms %.2f" % (st.kernel_ms, st.total_ms))
    assert st.active_cycles == sum(o.total_steps for o in outs)
    assert_batch_equals_plain(engine, sims, outs)


def test_batch_graph_replay_with_new_parameters(engine):
    """Same layout twice (the cached graph runs again) with new seeds and
    physical parameters: still identical to plain runs."""
    a = sample_sims(8, seed=2)
    b = [abi.clone_params(p, seed=p.seed + 77, K0=p.K0 * 1.3, dt=p.dt) for p in a]
    outs_a = engine.run_batch(a)
    outs_b = engine.run_batch(b)
    outs_a2 = engine.run_batch(a)
    assert [bytes(x) for x in outs_a] == [bytes(x) for x in outs_a2]
    assert_batch_equals_plain(engine, b, outs_b)


def test_batch_more_sims_than_one_launch_holds(engine):
    """1100 tiny simulations of one model need two trajectory launches
    (1024 simulations per launch)."""
    sims = [abi.make_params(abi.MODEL_1D_BP, N=20 + (i % 7), seed=50_000 + i,
                            K0_cm2s=5e22 * (1 + (i % 5) / 10.0)) for i in range(1100)]
    outs = engine.run_batch(sims)
    assert_batch_equals_plain(engine, sims[:40] + sims[1000:1100], outs[:40] + outs[1000:1100])
    assert all(o.n_particles == p.n_particles for o, p in zip(outs, sims))


def test_batch_budget_splits_and_runs_large_sims_alone(engine):
    sims = sample_sims(10, seed=3)
    engine.set_batch_budget(1 << 16)
    try:
        outs = engine.run_batch(sims)
    finally:
        engine.set_batch_budget(1 << 29)
    assert_batch_equals_plain(engine, sims, outs)


def test_batch_fp_many_crossings(engine):
    """1D F-p with a short step re-crosses 1 AU often; a batch whose event
    region is too small for a simulation runs that simulation again alone."""
    sims = [abi.make_params(abi.MODEL_1D_FP, N=400_000, seed=21, dt=50.0),
            abi.make_params(abi.MODEL_1D_BP, N=30_000, seed=22)]
    outs = engine.run_batch(sims)
    say("F-p events", outs[0].n_events, "for", sims[0].n_particles, "particles")
    assert_batch_equals_plain(engine, sims, outs)


def test_batch_errors(engine):
    good = abi.make_params(abi.MODEL_1D_BP, N=100, seed=1)
    bad = abi.clone_params(good, dt=-1.0)
    assert engine.run_batch_status([good, bad]) == abi.GS_ERR_INVALID_ARG
    assert engine.run_batch_status([]) == abi.GS_OK
    with pytest.raises(Exception):
        engine.set_batch_budget(10)


def test_autotune_picks_a_valid_shape_and_keeps_results(engine):
    p = abi.make_params(abi.MODEL_1D_BP, N=200_000, seed=31, weights=ALL_WEIGHTS)
    ref, _, _ = engine.run_plain(p)
    try:
        bpcu, rate = engine.autotune(p)
        say("autotune 1D B-p: %d blocks/CU, %.3e steps/s" % (bpcu, rate))
        assert 1 <= bpcu <= 16 and rate > 0 and math.isfinite(rate)
        again, _, _ = engine.run_plain(p)
    finally:
        engine.set_launch(abi.MODEL_1D_BP, 0, 0)
    assert bytes(ref) == bytes(again)
