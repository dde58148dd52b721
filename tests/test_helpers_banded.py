"""The guard-band island comparison used by the FP32 tests: it tolerates islands
whose extremum lies inside the band around a detection threshold and rejects a
missing island, a moved extremum or a flux error beyond it."""
import pytest

from tests.helpers import islands_agree_banded


def _isl(sign, x, y, val):
    return {"sign": sign, "extremum_x_idx": x, "extremum_y_idx": y, "extremum_val": val}


BASE = [_isl(1, 10, 10, 9.0), _isl(1, 40, 7, 5.3), _isl(-1, 3, 3, -8.0), _isl(1, 60, 60, 5.02)]


def test_identical_lists():
    assert islands_agree_banded(BASE, BASE, 5.0, -5.0, 0.05) == (3, 1)


def test_near_threshold_island_on_one_side_only():
    # 5.02 is inside the band (5.0 + 0.05): missing on one side is allowed
    assert islands_agree_banded(BASE[:3], BASE, 5.0, -5.0, 0.05) == (3, 1)


@pytest.mark.parametrize("mutate", ["drop", "move", "flux"])
def test_mistakes_beyond_the_band_fail(mutate):
    other = [dict(i) for i in BASE]
    if mutate == "drop":
        del other[1]
    elif mutate == "move":
        other[2]["extremum_x_idx"] += 1
    else:
        other[0]["extremum_val"] += 0.2
    with pytest.raises(AssertionError):
        islands_agree_banded(BASE, other, 5.0, -5.0, 0.05)
