"""End-to-end physics check against measured data: the reference's Ulysses
comparison (Visualization/create_ulysses_Geliosphere_flux.py) through the
drop-in CLI.  For each of the 67 Carrington rotations of 1994-1998 a
Geliosphere run at Ulysses' position takes the rotation's K0, tilt and
polarity from the measured-value table (-m/-y); its Ulysses-spectrum flux
above 0.25 GeV is set beside the Ulysses KET P190 proton flux averaged over
the rotation (geliosphere_amd/data/ulysses_1994_1998.csv, made from the reference's
data files by make_ulysses_table.py).

The model is not fitted to these data, so the bar is physical, not
numerical: the modelled flux within a factor 2 of the measured one (median
over rotations), and the two time series ranked alike (Spearman > 0.6)
through Ulysses' fast latitude scan and the solar-minimum years.
"""
import pytest

pytestmark = pytest.mark.gpu


def test_geliosphere_tracks_ulysses_ket_p190(engine, tmp_path):
    from geliosphere_amd import ulysses
    out = ulysses.run(str(tmp_path / "uly"), millions=1, seed=11)
    s = ulysses.summary(out)
    print(s)
    assert s["rotations"] == 67
    assert 0.5 < s["median_ratio_model_to_ket"] < 2.0
    assert s["spearman"] > 0.6
