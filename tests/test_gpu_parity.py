"""Parity of the gfx950 engine (through the C ABI) with the oracle.

Three levels:
 1. same seed, small N: the GPU and the oracle integrate the same Philox
    noise for every particle id, so their events must agree particle by
    particle (to FP64 rounding of a different evaluation order) and the
    histograms bin for bin;
 2. independent seeds: every energy bin within 3 combined Monte Carlo
    standard errors and chi^2 p > 0.01 (the reference's own acceptance rule
    for its CPU path);
 3. engine properties: bit-exact replay, independence of launch geometry,
    exact additivity of particle shards, edge cases and errors.
"""
import math
import sys
import time

import pytest

from geliosphere_amd import abi
from tests.parity import compare_spectra

pytestmark = pytest.mark.gpu

MODELS = [abi.MODEL_1D_BP, abi.MODEL_1D_FP, abi.MODEL_2D_SOLARPROP, abi.MODEL_2D_GELIOSPHERE]
# Same-seed sizes: the oracle finishes each in a few seconds on 8 threads.
SAME_SEED_N = {abi.MODEL_1D_BP: 4000, abi.MODEL_1D_FP: 400000,
               abi.MODEL_2D_SOLARPROP: 1500, abi.MODEL_2D_GELIOSPHERE: 1200}
ORACLE_FLAG_PERTURB = 0x80000000   # oracle/oracle.h
ALL_WEIGHTS = (abi.W_POWERLAW, abi.W_YAMADA, abi.W_WEBER, abi.W_BURGER)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def record_diffs(g, o):
    """Per oracle event: the smallest max-relative difference (T_exit, r,
    theta, sumac) to a GPU event of the same particle (inf if none)."""
    gi = {}
    for r in g:
        gi.setdefault(r.id, []).append(r)
    out = []
    for r in o:
        best = float("inf")
        for c in gi.get(r.id, []):
            if rel(c.t_inj, r.t_inj) > 1e-14:
                continue
            d = max(rel(c.t_exit, r.t_exit), rel(c.r, r.r),
                    abs(c.theta - r.theta) / max(1.0, abs(r.theta)),
                    abs(c.sumac - r.sumac) / max(1.0, abs(r.sumac)))
            best = min(best, d)
        out.append(best)
    return out


TOLS = (1e-12, 1e-10, 1e-8, 1e-6, 1e-5, 1e-3)


def fractions_within(diffs):
    return [sum(d < t for d in diffs) / max(len(diffs), 1) for t in TOLS]


def same_seed_check(engine, oracle, p, label):
    """1D models: every event agrees to 1e-8 and step counts are identical
    (their steps use only +,-,*,/,sqrt).  2D models add sin, cos, acos,
    atan, whose GPU and libm results differ in the last bit, and their
    trajectories amplify last-bit differences (most in the Geliosphere polar
    branch).  There the GPU must stay as close to the oracle as the oracle
    stays to itself when every normal is perturbed by one ulp: the two
    distributions of per-particle differences must agree within binomial
    noise at every tolerance."""
    n = p.n_particles
    model = p.model
    say("same-seed parity", label, "N", n)
    t = time.time()
    gs, grecs, gn = engine.run_plain(p, records=True, rec_cap=4 * n + 64)
    longest = engine.last_stats().max_particle_steps
    tg = time.time() - t
    t = time.time()
    rc, os_, orecs, on = oracle.run(p, n_threads=0, records=True, rec_cap=4 * n + 64)
    say("  gpu %.3fs oracle %.3fs events gpu %d oracle %d" % (tg, time.time() - t, gn, on))
    assert rc == 0
    assert gs.n_particles == os_.n_particles == n
    fg = fractions_within(record_diffs(grecs, orecs))
    say("  gpu vs oracle    within " + " ".join("%.0e:%.4f" % (t, f) for t, f in zip(TOLS, fg)))
    say("  steps gpu %d oracle %d" % (gs.total_steps, os_.total_steps))
    if model in (abi.MODEL_1D_BP, abi.MODEL_1D_FP):
        assert fg[2] == 1.0 and gn == on and gs.total_steps == os_.total_steps
        if n <= 4000:   # the longest trajectory, particle by particle on the oracle
            assert longest == max(oracle.particle(p, p.first_particle + i)[1] for i in range(n))
        for b in range(p.n_bins):
            assert gs.count[b] == os_.count[b]
            for k in range(p.n_weights):
                assert gs.sum_w[k][b] == pytest.approx(os_.sum_w[k][b], rel=1e-9, abs=0.0)
        return
    q = abi.clone_params(p, flags=p.flags | ORACLE_FLAG_PERTURB)
    rc, ps, precs, pn = oracle.run(q, n_threads=0, records=True, rec_cap=4 * n + 64)
    fp = fractions_within(record_diffs(precs, orecs))
    say("  oracle vs 1-ulp within " + " ".join("%.0e:%.4f" % (t, f) for t, f in zip(TOLS, fp)))
    for a, b in zip(fg, fp):
        margin = 3.5 * math.sqrt(max(b * (1 - b), 1.0 / n) / n) + 0.01
        assert a >= b - margin
    assert gn == on
    assert rel(gs.total_steps, os_.total_steps) < 0.01


@pytest.mark.parametrize("model", MODELS)
def test_same_seed_particle_parity(engine, oracle, model):
    p = abi.make_params(model, N=SAME_SEED_N[model], seed=20240607, weights=ALL_WEIGHTS, tilt_deg=45.0)
    same_seed_check(engine, oracle, p, abi.MODEL_NAMES[model])


# Other injection modes, shard offsets, injection points, polarities and the
# reference GPU's SolarProp drift variant, particle by particle.
SAME_SEED_CASES = [
    ("1D B-p grid injection, shard offset, r=5 AU, dt=100 s",
     dict(model=abi.MODEL_1D_BP, N=3000, injection=abi.INJ_GRID, grid_n=997, first=1_000_003, r_inj=5.0, dt=100.0,
          weights=(abi.W_USOSKIN, abi.W_WEBER, abi.W_POWERLAW, abi.W_YAMADA))),
    ("1D F-p grid injection, V=600 km/s",
     dict(model=abi.MODEL_1D_FP, N=1_000_000, injection=abi.INJ_GRID, grid_n=5000, V_kms=600.0)),
    ("SolarProp bins, reference GPU drift, A<0, tilt 75, theta 60, r=3 AU",
     dict(model=abi.MODEL_2D_SOLARPROP, N=1200, injection=abi.INJ_SOLARPROP_BINS,
          flags=abi.FLAG_SOLARPROP_GPU_DRIFT, polarity=-1.0, tilt_deg=75.0, theta_deg=60.0, r_inj=3.0)),
    ("Geliosphere bins, A<0, polar injection theta 1 deg",
     dict(model=abi.MODEL_2D_GELIOSPHERE, N=1000, injection=abi.INJ_SOLARPROP_BINS, polarity=-1.0,
          tilt_deg=20.0, theta_deg=1.0)),
    ("Geliosphere log-uniform, r=30 AU, theta 150, shard offset",
     dict(model=abi.MODEL_2D_GELIOSPHERE, N=1000, r_inj=30.0, theta_deg=150.0, tilt_deg=60.0, first=777_777)),
    ("Geliosphere polar cap, theta 1 deg, delta0 = 0.003 sin(1 deg) (the polar terms are exact derivatives)",
     dict(model=abi.MODEL_2D_GELIOSPHERE, N=1000, theta_deg=1.0, tilt_deg=30.0,
          delta0=0.003 * math.sin(math.radians(1.0)))),
]


@pytest.mark.parametrize("label,kw", SAME_SEED_CASES, ids=[c[0] for c in SAME_SEED_CASES])
def test_same_seed_particle_parity_configs(engine, oracle, label, kw):
    kw = dict(kw)
    model = kw.pop("model")
    weights = kw.pop("weights", ALL_WEIGHTS)
    p = abi.make_params(model, seed=424242, weights=weights, **kw)
    same_seed_check(engine, oracle, p, label)


STAT_CASES = [
    # model, GPU N, oracle N, extra params
    (abi.MODEL_1D_BP, 2_000_000, 60_000, {}),
    (abi.MODEL_2D_SOLARPROP, 200_000, 3000, {"tilt_deg": 10.0, "polarity": 1.0}),
    (abi.MODEL_2D_SOLARPROP, 200_000, 3000, {"tilt_deg": 75.0, "polarity": -1.0}),
    (abi.MODEL_2D_GELIOSPHERE, 200_000, 3000, {"tilt_deg": 45.0, "polarity": 1.0}),
    (abi.MODEL_2D_GELIOSPHERE, 200_000, 3000, {"tilt_deg": 10.0, "polarity": -1.0}),
]


@pytest.mark.parametrize("model,ng,no,extra", STAT_CASES)
def test_independent_seed_statistical_parity(engine, oracle, model, ng, no, extra):
    pg = abi.make_params(model, N=ng, seed=1001, **extra)
    po = abi.make_params(model, N=no, seed=2002, **extra)
    say("stat parity", abi.MODEL_NAMES[model], extra, "gpu N", ng, "oracle N", no)
    gs, _, _ = engine.run_plain(pg)
    rc, os_, _, _ = oracle.run(po, n_threads=0)
    assert rc == 0
    zs, chi2, dof, pval = compare_spectra(gs, os_, min_count=10)
    say("  bins %d chi2 %.2f p %.4f max|z| %.2f" % (dof, chi2, pval, max(abs(z) for z in zs)))
    assert dof >= 20
    assert pval > 0.01
    assert max(abs(z) for z in zs) < 3.0 + 0.5   # 3 SE; 0.5 slack for the max of ~30 bins


def test_fp_statistical_parity(engine, oracle):
    """1D F-p: about 3 in 10^5 particles reach 1 AU, and a particle that does
    usually crosses it several times, so crossings are clustered.  Compare the
    fraction of particles with at least one crossing (binomial)."""
    pg = abi.make_params(abi.MODEL_1D_FP, N=50_000_000, seed=77)
    po = abi.make_params(abi.MODEL_1D_FP, N=3_000_000, seed=78)
    gs, grecs, gn = engine.run_plain(pg, records=True, rec_cap=200_000)
    rc, os_, orecs, on = oracle.run(po, n_threads=0, records=True, rec_cap=200_000)
    assert gn <= 200_000 and on <= 200_000
    kg, ko = len({r.id for r in grecs}), len({r.id for r in orecs})
    rg, ro = kg / pg.n_particles, ko / po.n_particles
    se = math.sqrt(rg * (1 - rg) / pg.n_particles + ro * (1 - ro) / po.n_particles)
    say("F-p crossing particles gpu %d/%d oracle %d/%d  z %.2f  events/particle gpu %.2f oracle %.2f"
        % (kg, pg.n_particles, ko, po.n_particles, (rg - ro) / se, gn / max(kg, 1), on / max(ko, 1)))
    assert ko > 30
    assert abs(rg - ro) < 3.0 * se


REPLAY_N = {abi.MODEL_1D_BP: 300_000, abi.MODEL_1D_FP: 2_000_000,
            abi.MODEL_2D_SOLARPROP: 20_000, abi.MODEL_2D_GELIOSPHERE: 20_000}


@pytest.mark.parametrize("model", MODELS)
def test_replay_is_bit_exact(engine, model):
    """Same seed, same launch: spectra and every event identical, bit for bit."""
    p = abi.make_params(model, N=REPLAY_N[model], seed=5, weights=ALL_WEIGHTS)
    a, ra, na = engine.run_plain(p, records=True, rec_cap=200_000)
    b, rb, nb = engine.run_plain(p, records=True, rec_cap=200_000)
    assert bytes(a) == bytes(b)
    assert na == nb and [bytes(x) for x in ra] == [bytes(x) for x in rb]


@pytest.mark.parametrize("model", [abi.MODEL_1D_BP, abi.MODEL_2D_GELIOSPHERE])
def test_result_independent_of_launch_geometry(engine, model):
    p = abi.make_params(model, N=200_000 if model == abi.MODEL_1D_BP else 20_000, seed=6,
                        weights=ALL_WEIGHTS)
    ref, _, _ = engine.run_plain(p)
    engine.set_launch(model, 0, 1)
    try:
        one, _, _ = engine.run_plain(p)
    finally:
        engine.set_launch(model, 0, 0)
    assert bytes(ref) == bytes(one)


def test_shards_add_exactly(engine):
    """Particles are keyed by id, and sums are exact integers, so two shards
    of the id range add up to the whole run bit for bit (gs_spectrum_merge,
    the reduction a run split across GPUs would use)."""
    n = 100_000
    p = abi.make_params(abi.MODEL_1D_BP, N=n, seed=8, weights=ALL_WEIGHTS)
    whole, _, _ = engine.run_plain(p)
    a, _, _ = engine.run_plain(abi.clone_params(p, n_particles=37_000))
    b, _, _ = engine.run_plain(abi.clone_params(p, n_particles=n - 37_000, first_particle=37_000))
    for bin_ in range(p.n_bins):
        assert a.count[bin_] + b.count[bin_] == whole.count[bin_]
        for k in range(len(ALL_WEIGHTS)):
            for s in (0, 1):
                assert abi.acc_int(a, k, s, bin_) + abi.acc_int(b, k, s, bin_) == abi.acc_int(whole, k, s, bin_)
    assert a.total_steps + b.total_steps == whole.total_steps
    # the library's reduction of the shards is the whole run, bit for bit
    from geliosphere_amd.engine import merge_spectra
    rc, m = merge_spectra([b, a])
    assert rc == abi.GS_OK and bytes(m) == bytes(whole)


def test_exact_sums_convert_correctly(engine):
    from fractions import Fraction
    p = abi.make_params(abi.MODEL_1D_BP, N=20_000, seed=9, weights=ALL_WEIGHTS)
    s, _, _ = engine.run_plain(p)
    lsb = Fraction(2) ** abi.GS_ACC_LSB_EXP
    for k in range(len(ALL_WEIGHTS)):
        for b in range(p.n_bins):
            assert s.sum_w[k][b] == float(abi.acc_int(s, k, 0, b) * lsb)
            assert s.sum_w2[k][b] == float(abi.acc_int(s, k, 1, b) * lsb)


def test_zero_and_one_particle(engine, oracle):
    s, _, n = engine.run_plain(abi.make_params(abi.MODEL_1D_BP, N=0))
    assert s.n_particles == 0 and s.n_events == 0 and n == 0
    p = abi.make_params(abi.MODEL_2D_SOLARPROP, N=1, seed=3, first=12345)
    s, recs, n = engine.run_plain(p, records=True, rec_cap=4)
    rec_o, steps, status = oracle.particle(p, 12345)
    assert s.n_particles == 1 and n == len(rec_o) == 1 and recs[0].id == 12345
    assert s.total_steps == steps
    assert recs[0].t_exit == pytest.approx(rec_o[0].t_exit, rel=1e-9)


def test_step_cap_counts_and_stops(engine, oracle):
    p = abi.make_params(abi.MODEL_1D_BP, N=2000, seed=4, max_steps=500)
    s, _, _ = engine.run_plain(p)
    rc, o, _, _ = oracle.run(p, n_threads=0)
    assert s.n_step_cap == o.n_step_cap > 0
    assert s.total_steps == o.total_steps


def test_record_capacity_overflow_reports_total(engine):
    p = abi.make_params(abi.MODEL_1D_BP, N=5000, seed=4)
    s, recs, n = engine.run_plain(p, records=True, rec_cap=100)
    assert n == s.n_events == 5000 and len(recs) == 100


def test_invalid_params_rejected(engine):
    p = abi.make_params(abi.MODEL_1D_BP, N=10)
    p.dt = -1.0
    assert engine.run_plain_status(p) == abi.GS_ERR_INVALID_ARG


def test_outside_range_events_counted(engine, oracle):
    edges = abi.log_edges(10, 0.1, 10.0)
    p = abi.make_params(abi.MODEL_1D_BP, N=3000, seed=12, edges=edges)
    s, _, _ = engine.run_plain(p)
    rc, o, _, _ = oracle.run(p, n_threads=0)
    assert s.n_outside == o.n_outside
    assert s.n_outside + sum(s.count[b] for b in range(10)) == s.n_events


def test_overflow_rerun_keeps_earlier_chunks_statistics(engine, oracle):
    """Plain runs cut into launches of 3000 particles whose event region holds
    2000 records: a launch overflows (nearly every B-p particle exits), runs
    again as 1500 particles, and must start again from the statistics of the
    launches before it.  The result equals the one-launch run bit for bit and
    the oracle's counts."""
    p = abi.make_params(abi.MODEL_1D_BP, N=20_000, seed=31, weights=ALL_WEIGHTS)
    ref, _, nref = engine.run_plain(p)
    engine.set_plain_limits(3000, 2000)
    try:
        cut, _, ncut = engine.run_plain(p)
        st = engine.last_stats()
    finally:
        engine.set_plain_limits(0, 0)
    assert bytes(cut) == bytes(ref) and ncut == nref
    assert cut.n_particles == 20_000 and st.max_particle_steps > 0
    rc, o, _, _ = oracle.run(p, n_threads=0)
    assert cut.total_steps == o.total_steps and cut.n_events == o.n_events
    assert [cut.count[b] for b in range(30)] == [o.count[b] for b in range(30)]
