"""The band bar of the DF end-to-end tests (fixtures.within_band)."""
import fixtures as F


def test_band_is_two_sided_where_the_reference_agrees_with_itself():
    assert F.within_band(0.5004, [0.5, 0.5001], 1e-3)
    assert F.within_band(0.4992, [0.5, 0.5001], 1e-3)
    assert not F.within_band(0.5021, [0.5, 0.5001], 1e-3)
    assert not F.within_band(0.4985, [0.5, 0.5001], 1e-3)


def test_band_is_one_sided_where_the_reference_disagrees_with_itself():
    assert F.within_band(0.503, [0.376, 0.498], 1e-3)
    assert F.within_band(0.3755, [0.376, 0.498], 1e-3)
    assert not F.within_band(0.3745, [0.376, 0.498], 1e-3)
