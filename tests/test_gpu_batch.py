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

ALL_WEIGHTS = (abi.W_POWERLAW, abi.W_YAMADA, abi.W_WEBER, abi.W_USOSKIN, abi.W_BURGER)


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
    p = abi.make_params(abi.MODEL_1D_BP, N=200_000, seed=31, weights=ALL_WEIGHTS[:4])
    ref, _, _ = engine.run_plain(p)
    try:
        bpcu, rate = engine.autotune(p)
        say("autotune 1D B-p: %d blocks/CU, %.3e steps/s" % (bpcu, rate))
        assert 1 <= bpcu <= 16 and rate > 0 and math.isfinite(rate)
        again, _, _ = engine.run_plain(p)
    finally:
        engine.set_launch_shape(abi.MODEL_1D_BP, abi.KERNEL_PLAIN)
    assert bytes(ref) == bytes(again)


def test_config4_sampled_batch_matches_oracle(engine, oracle):
    """Sixteen simulations drawn by the batch benchmark's sampler (four of
    each model; K0 log-uniform 1e22-1e23 cm^2/s, V 300-800 km/s or tilt
    5-75 deg, r 1-80 AU, theta 10-170 deg, dt in {50, 100, 500} s, A +-1),
    at particle counts the oracle finishes in seconds, run as one batch and
    compared with the oracle at the same seed: 1D simulations bin for bin
    (counts and steps identical, sums to 1e-9); 2D simulations, whose
    trajectories amplify last-bit differences of sin/cos/atan, with
    identical per-bin counts (the bin is the injection energy), steps within
    2 % and per-bin mean weights within 3 SE.  The reference runs such rows
    one after the other (Algorithm/src/BatchRun.cpp:26-79)."""
    from geliosphere_amd.batch import sample_one
    from tests.parity import compare_spectra
    rng = random.Random(4040)
    n_small = {abi.MODEL_1D_FP: 200_000, abi.MODEL_1D_BP: 1_500, abi.MODEL_2D_SOLARPROP: 150,
               abi.MODEL_2D_GELIOSPHERE: 150}
    w1d = (abi.W_POWERLAW, abi.W_YAMADA, abi.W_WEBER, abi.W_USOSKIN)
    w2d = (abi.W_POWERLAW, abi.W_WEBER, abi.W_USOSKIN, abi.W_BURGER)
    sims = []
    for i in range(16):
        model = (abi.MODEL_1D_FP, abi.MODEL_1D_BP, abi.MODEL_2D_SOLARPROP, abi.MODEL_2D_GELIOSPHERE)[i % 4]
        p = sample_one(rng, 9000 + i, model=model)
        p.n_particles = n_small[model]
        w = w1d if model < 2 else w2d
        p.n_weights = len(w)
        for k, kind in enumerate(w):
            p.weights[k] = kind
        sims.append(p)
    outs = engine.run_batch(sims)
    for i, (p, g) in enumerate(zip(sims, outs)):
        rc, o, _, _ = oracle.run(p, n_threads=0)
        assert rc == 0
        say("sim %2d %-18s r %5.1f dt %3.0f  steps gpu %d oracle %d  events %d %d"
            % (i, abi.MODEL_NAMES[p.model], p.r_inj, p.dt, g.total_steps, o.total_steps, g.n_events, o.n_events))
        assert g.n_particles == o.n_particles == p.n_particles
        assert [g.count[b] for b in range(30)] == [o.count[b] for b in range(30)]
        assert g.n_events == o.n_events and g.n_outside == o.n_outside
        if p.model < 2:
            assert g.total_steps == o.total_steps
            for k in range(4):
                for b in range(30):
                    assert g.sum_w[k][b] == pytest.approx(o.sum_w[k][b], rel=1e-9, abs=0.0)
        else:
            assert abs(g.total_steps - o.total_steps) <= 0.02 * o.total_steps
            for k in range(4):
                zs, chi2, dof, pval = compare_spectra(g, o, k=k, min_count=5)
                assert all(abs(z) < 3.0 for z in zs), (i, k, zs)


def test_batch_records_equal_plain_records(engine):
    """gs_run_batch_records returns each simulation's exit records as
    gs_run_plain does (sorted by particle id), including a simulation that
    runs alone because the event budget is small, a simulation asking no
    records, and a record capacity below the event count."""
    sims = sample_sims(6, seed=8)
    sims.append(abi.make_params(abi.MODEL_1D_BP, N=70_000, seed=81))   # alone under a 2^16 budget
    caps = [2 * p.n_particles + 16 for p in sims]
    caps[1] = 0                       # no records for this one
    caps[2] = 7                       # fewer than its events
    engine.set_batch_budget(1 << 16)
    try:
        outs, recs, nrec = engine.run_batch_records(sims, caps)
    finally:
        engine.set_batch_budget(1 << 29)
    for i, p in enumerate(sims):
        s, r, n = engine.run_plain(p, records=True, rec_cap=max(2 * p.n_particles, 16))
        assert bytes(s) == bytes(outs[i]) and nrec[i] == n, i
        if caps[i] == 0:
            assert recs[i] is None
            continue
        want = r[:caps[i]]
        assert len(recs[i]) == min(caps[i], n), i
        assert [bytes(a) for a in recs[i]] == [bytes(b) for b in want], i


SHAPES = ((128, 0, 16), (512, 1, 256), (256, 2, 32), (128, 3, 128))


@pytest.mark.parametrize("model", [abi.MODEL_1D_BP, abi.MODEL_1D_FP, abi.MODEL_2D_GELIOSPHERE])
def test_results_independent_of_launch_shape(engine, model):
    """Workgroup size, blocks per CU and ids per queue grab change which lane
    integrates which particle and when, never a result: plain runs and
    batches under every shape equal the default shape's bit for bit."""
    n = {abi.MODEL_1D_BP: 200_000, abi.MODEL_1D_FP: 2_000_000, abi.MODEL_2D_GELIOSPHERE: 20_000}[model]
    p = abi.make_params(model, N=n, seed=61, weights=ALL_WEIGHTS[:4])
    sims = [p, abi.clone_params(p, seed=62, n_particles=n // 3, r_inj=7.0)]
    ref, _, _ = engine.run_plain(p)
    refb = engine.run_batch(sims)
    for shape in SHAPES:
        engine.set_launch_shape(model, abi.KERNEL_PLAIN, *shape)
        engine.set_launch_shape(model, abi.KERNEL_BATCH, *shape)
        try:
            got, _, _ = engine.run_plain(p)
            blk = engine.last_stats().block_threads
            gotb = engine.run_batch(sims)
            used = engine.launch_shape(model, abi.KERNEL_PLAIN)
        finally:
            engine.set_launch_shape(model, abi.KERNEL_PLAIN)
            engine.set_launch_shape(model, abi.KERNEL_BATCH)
        say("shape", shape, "-> used", used)
        assert blk == shape[0] and used[0] == shape[0] and used[2] == shape[2]
        assert bytes(got) == bytes(ref), shape
        assert [bytes(x) for x in gotb] == [bytes(x) for x in refb], shape


def test_autotuner_searches_shapes_and_caches_them(engine, tmp_path):
    """The autotuner searches workgroup size, blocks per CU and ids per grab
    for a model's plain and batch kernels, applies the best, and its cache
    file carries the shapes to a new context on the same device."""
    from geliosphere_amd.engine import Engine
    p = abi.make_params(abi.MODEL_1D_BP, N=300_000, seed=71)
    ref, _, _ = engine.run_plain(p)
    pilots = [abi.make_params(abi.MODEL_1D_BP, N=20_000 + 5_000 * i, seed=72 + i, r_inj=1.0 + 10 * i)
              for i in range(4)]
    try:
        sh, rate = engine.autotune_shape([p], abi.KERNEL_PLAIN, 1)
        shb, rateb = engine.autotune_shape(pilots, abi.KERNEL_BATCH, 1)
        say("tuned plain", sh, "%.3e" % rate, "batch", shb, "%.3e" % rateb)
        assert sh[0] in (128, 256, 512) and sh[1] >= 1 and sh[2] in (16, 32, 64, 128, 256)
        assert shb[0] in (128, 256, 512) and shb[1] >= 1 and rate > 0 and rateb > 0
        assert engine.launch_shape(abi.MODEL_1D_BP, abi.KERNEL_PLAIN) == sh
        again, _, _ = engine.run_plain(p)
        assert bytes(again) == bytes(ref)
        path = str(tmp_path / "autotune.txt")
        engine.autotune_save(path)
        lines = [l.split() for l in open(path) if not l.startswith("#")]
        assert sorted(l[2] for l in lines) == ["batch", "plain"] and all(l[3] == "1" for l in lines)
        with Engine(0) as e2:
            assert e2.autotune_load(path) == 2
            assert e2.launch_shape(abi.MODEL_1D_BP, abi.KERNEL_PLAIN) == sh
            assert e2.launch_shape(abi.MODEL_1D_BP, abi.KERNEL_BATCH) == shb
    finally:
        engine.set_launch_shape(abi.MODEL_1D_BP, abi.KERNEL_PLAIN)
        engine.set_launch_shape(abi.MODEL_1D_BP, abi.KERNEL_BATCH)


def test_autotune_cache_file_rules(engine, tmp_path):
    """The cache applies only well-formed lines of this device (architecture
    name and CU count) for the plain and batch kernels, and saving keeps the
    lines of other devices, including a last line without a newline."""
    from geliosphere_amd.engine import Engine
    path = tmp_path / "cache.txt"
    with Engine(0) as e:
        arch_line = None
        # learn this device's key from the save of one tuned shape
        p = abi.make_params(abi.MODEL_1D_BP, N=50_000, seed=3)
        e.autotune_shape([p], abi.KERNEL_PLAIN, 1)
        e.autotune_save(str(path))
        for l in open(path):
            if not l.startswith("#"):
                arch_line = l.split()
        arch, cus = arch_line[0], arch_line[1]
    text = ("# header\n"
            "%s %s plain 3 512 1 128 1e10\n" % (arch, cus) +         # applies
            "%s %s weird 2 256 1 64 1e10\n" % (arch, cus) +          # unknown kernel: skipped
            "%s %s batch 2 256\n" % (arch, cus) +                     # malformed: skipped
            "gfx000 %s batch 1 128 2 16 1e10\n" % cus +               # another device: skipped
            "%s 1 batch 1 128 2 16 1e10" % arch)                      # another CU count, no newline
    path.write_text(text)
    with Engine(0) as e:
        assert e.autotune_load(str(path)) == 1
        assert e.launch_shape(abi.MODEL_2D_GELIOSPHERE, abi.KERNEL_PLAIN) == (512, 1, 128)
        e.autotune_shape([p], abi.KERNEL_PLAIN, 1)
        e.autotune_save(str(path))
    lines = open(path).read().splitlines()
    assert "gfx000 %s batch 1 128 2 16 1e10" % cus in lines
    assert "%s 1 batch 1 128 2 16 1e10" % arch in lines
    mine = [l.split() for l in lines if l.startswith(arch + " " + cus + " ")]
    assert sorted((l[2], l[3]) for l in mine if len(l) == 8 and l[2] in ("plain", "batch")) == \
        [("plain", "1"), ("plain", "3")]
    assert "%s %s batch 2 256" % (arch, cus) in lines      # lines it cannot read are kept as they are
