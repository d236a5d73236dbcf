"""Scheduling order of plain runs (gs_set_order).

The order decides which lane integrates which particle and when, never what
comes out: particles are keyed by id and the histogram sums are exact, so
id order and energy order (the default: lowest injection energy, i.e. the
longest trajectories of the backward models, first) must give bit-identical
spectra and records for every injection mode, shard offset and model.
"""
import pytest

from geliosphere_amd import abi
from geliosphere_amd import engine as E

pytestmark = pytest.mark.gpu

ALL_WEIGHTS = (abi.W_POWERLAW, abi.W_YAMADA, abi.W_USOSKIN, abi.W_BURGER)

CASES = [
    ("1D B-p log-uniform", dict(model=abi.MODEL_1D_BP, N=300_000)),
    ("1D B-p grid, shard", dict(model=abi.MODEL_1D_BP, N=200_000, injection=abi.INJ_GRID, grid_n=7777,
                               first=123_457)),
    ("1D F-p (id order kept)", dict(model=abi.MODEL_1D_FP, N=400_000)),
    ("SolarProp log-uniform", dict(model=abi.MODEL_2D_SOLARPROP, N=30_000)),
    ("Geliosphere SolarProp bins", dict(model=abi.MODEL_2D_GELIOSPHERE, N=20_000,
                                        injection=abi.INJ_SOLARPROP_BINS)),
    ("Geliosphere fixed energy", dict(model=abi.MODEL_2D_GELIOSPHERE, N=8_000, injection=abi.INJ_FIXED,
                                      T_min=1.0)),
]


def utilisation(st):
    return st.active_cycles / st.lane_cycles if st.lane_cycles else 0.0


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_order_does_not_change_results(engine, name, kw):
    kw = dict(kw)
    model = kw.pop("model")
    p = abi.make_params(model, seed=31, weights=ALL_WEIGHTS, **kw)
    engine.set_order(E.ORDER_ID)
    try:
        a, ra, na = engine.run_plain(p, records=True, rec_cap=400_000)
        ua = utilisation(engine.last_stats())
    finally:
        engine.set_order(E.ORDER_ENERGY)
    b, rb, nb = engine.run_plain(p, records=True, rec_cap=400_000)
    ub = utilisation(engine.last_stats())
    print("%-28s lane utilisation: id order %.3f, energy order %.3f" % (name, ua, ub))
    assert bytes(a) == bytes(b)
    assert na == nb and len(ra) == len(rb)
    assert [bytes(x) for x in ra] == [bytes(x) for x in rb]


def test_order_mode_is_validated(engine):
    with pytest.raises(E.GsError):
        engine.set_order(7)
    engine.set_order(E.ORDER_ENERGY)


def test_batch_order_does_not_change_results(engine):
    """The batch path orders each simulation's particles the same way; its
    spectra must not depend on the mode either."""
    from geliosphere_amd import batch
    sims = batch.sample_batch(24, seed=77, n_scale=0.003)
    engine.set_order(E.ORDER_ID)
    try:
        a = engine.run_batch(sims)
        ua = utilisation(engine.last_stats())
    finally:
        engine.set_order(E.ORDER_ENERGY)
    b = engine.run_batch(sims)
    ub = utilisation(engine.last_stats())
    print("batch of 24 lane utilisation: id order %.3f, energy order %.3f" % (ua, ub))
    assert [bytes(x) for x in a] == [bytes(x) for x in b]
