"""The baseline's 2D sweep under parity, and the 2D models in their
deterministic regime.

Config 3: SolarProp-like and Geliosphere, tilt 10/45/75 deg, polarity +1 and
-1 (12 runs), r = 1 AU, theta = 90 deg, K0 = 5e22 cm^2/s, V = 400 km/s,
dt = 50 s, log-uniform 0.01-100 GeV, 30 log bins, N = 1e7 on the GPU.  Each
run is compared with the oracle two ways:
 * independent seeds: the GPU spectrum (N = 1e7) against the oracle's at
   N = 40,000 (tests/golden/config3_oracle_spectra.json, written by
   tests/golden/make_config3_fixtures.py): every bin within 3 combined
   standard errors and chi^2 p > 0.01 for the power-law LIS, and within
   3.5 SE for the Webber-Higbie, Usoskin and Burger weights;
 * same seed, per particle (N = 800), as in test_gpu_parity.

The reference reaches tilt and polarity only through its month/year tables
and settings (Input/src/InputValidation.cpp:136-141).

Deterministic regime: with no diffusion (K0 = 0, or 1e5 cm^2/s for the
Geliosphere model, which divides by K_tt), A < 0 protons on the equator drift
out along the current sheet faster than the wind and leave on a smooth,
noise-free path.  There the GPU and the oracle must agree particle by
particle to 1e-9, and every 2D weight's bin sums to 1e-8 (the oracle's exit
energies on this path are pinned to the continuum limit, and its LIS to the
published spectra, in tests/test_oracle_pins.py).
"""
import json
import math
import os
import sys

import pytest

from geliosphere_amd import abi
from tests.test_gpu_parity import same_seed_check

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "config3_oracle_spectra.json")
WEIGHTS_2D = (abi.W_POWERLAW, abi.W_WEBER, abi.W_USOSKIN, abi.W_BURGER)
N_GPU = 10_000_000
CASES = [(m, t, pol) for m in (abi.MODEL_2D_SOLARPROP, abi.MODEL_2D_GELIOSPHERE)
         for t in (10.0, 45.0, 75.0) for pol in (1.0, -1.0)]
IDS = ["%s-tilt%d-%s" % ("SP" if m == abi.MODEL_2D_SOLARPROP else "GE", t, "Apos" if pol > 0 else "Aneg")
       for m, t, pol in CASES]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load_fixture():
    with open(FIXTURE) as f:
        return json.load(f)


def fixture_case(fx, model, tilt, pol):
    for c in fx["cases"]:
        if c["model"] == model and c["tilt_deg"] == tilt and c["polarity"] == pol:
            return c
    raise KeyError((model, tilt, pol))


def bin_mean_se(n, sw, sw2):
    m = sw / n
    return m, math.sqrt(max(sw2 / n - m * m, 0.0) / n)


@pytest.mark.parametrize("model,tilt,pol", CASES, ids=IDS)
def test_config3_statistical_parity(engine, model, tilt, pol):
    from scipy import stats
    fx = load_fixture()
    ref = fixture_case(fx, model, tilt, pol)
    assert fx["weights"] == list(WEIGHTS_2D)
    p = abi.make_params(model, N=N_GPU, seed=777_001, tilt_deg=tilt, polarity=pol, weights=WEIGHTS_2D)
    say("config 3", abi.MODEL_NAMES[model], "tilt", tilt, "polarity", pol, "GPU N", N_GPU, "oracle N", fx["N"])
    g, _, _ = engine.run_plain(p)
    st = engine.last_stats()
    say("  GPU %.2f s, %.3e particle-steps/s" % (st.kernel_ms * 1e-3, g.total_steps / (st.kernel_ms * 1e-3)))
    assert g.n_particles == N_GPU and g.n_nan == 0 and g.n_step_cap == 0
    # exit fraction per bin (binomial): the same particles leave
    for b in range(30):
        po = ref["count"][b] / fx["N"]
        pg = g.count[b] / N_GPU
        se = math.sqrt(po * (1 - po) / fx["N"] + pg * (1 - pg) / N_GPU) + 1e-12
        assert abs(pg - po) < 4.0 * se, (b, pg, po)
    for k, kind in enumerate(WEIGHTS_2D):
        zs = []
        for b in range(30):
            no, ng = ref["count"][b], g.count[b]
            if no < 50 or ng < 50:
                continue
            mo, so = bin_mean_se(no, ref["sum_w"][k][b], ref["sum_w2"][k][b])
            mg, sg = bin_mean_se(ng, g.sum_w[k][b], g.sum_w2[k][b])
            zs.append((mg - mo) / math.hypot(so, sg))
        chi2 = sum(z * z for z in zs)
        pval = stats.chi2.sf(chi2, len(zs))
        say("  weight %d: bins %d chi2 %.1f p %.3f max|z| %.2f" % (kind, len(zs), chi2, pval, max(map(abs, zs))))
        assert len(zs) >= 25
        if kind == abi.W_POWERLAW:
            assert pval > 0.01
            assert max(abs(z) for z in zs) < 3.0 + 0.5   # 3 SE; 0.5 slack for the max of ~30 bins
        else:
            assert max(abs(z) for z in zs) < 3.5 + 0.5


@pytest.mark.parametrize("model,tilt,pol", CASES, ids=IDS)
def test_config3_same_seed(engine, oracle, model, tilt, pol):
    p = abi.make_params(model, N=800, seed=606_000 + int(tilt) * 10 + (pol > 0), tilt_deg=tilt,
                        polarity=pol, weights=WEIGHTS_2D)
    same_seed_check(engine, oracle, p, "config 3 %s tilt %d polarity %+d" % (abi.MODEL_NAMES[model], tilt, pol))


@pytest.mark.parametrize("model", [abi.MODEL_2D_SOLARPROP, abi.MODEL_2D_GELIOSPHERE])
def test_sheet_drift_regime_particle_and_bin_parity(engine, oracle, model):
    K0 = 0.0 if model == abi.MODEL_2D_SOLARPROP else 1e5
    p = abi.make_params(model, K0_cm2s=K0, polarity=-1.0, tilt_deg=10.0, theta_deg=90.0, N=96,
                        injection=abi.INJ_GRID, grid_n=96, T_min=2.0, T_max=50.0, seed=5,
                        weights=WEIGHTS_2D, edges=abi.log_edges(12, 1.5, 60.0))
    g, grecs, gn = engine.run_plain(p, records=True, rec_cap=200)
    rc, o, orecs, on = oracle.run(p, n_threads=0, records=True, rec_cap=200)
    say("sheet-drift regime", abi.MODEL_NAMES[model], "events gpu", gn, "oracle", on,
        "steps", g.total_steps, o.total_steps)
    assert rc == 0 and gn == on == 96
    assert g.total_steps == o.total_steps
    worst = 0.0
    for a, b in zip(grecs, orecs):
        assert a.id == b.id and a.t_inj == b.t_inj
        worst = max(worst, abs(a.t_exit - b.t_exit) / b.t_exit, abs(a.r - b.r) / b.r)
        assert abs(a.theta - b.theta) < 1e-9
    say("  worst relative difference %.2e" % worst)
    assert worst < 1e-9
    for b in range(12):
        assert g.count[b] == o.count[b]
        for k in range(4):
            assert g.sum_w[k][b] == pytest.approx(o.sum_w[k][b], rel=1e-8, abs=0.0)
            assert g.sum_w2[k][b] == pytest.approx(o.sum_w2[k][b], rel=2e-8, abs=0.0)
