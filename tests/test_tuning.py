"""Library tuning state (qlt_tuning_get / qlt_tuning_set), host side only."""
import ctypes

import pytest

from tensortoolkit_amd import autotune as T
from tensortoolkit_amd import capi


@pytest.fixture
def builtin_tuning():
    T.set_tuning(None)
    yield
    T.set_tuning(None)


def test_default_is_the_builtin_rule(builtin_tuning):
    d = T.get_tuning()
    assert all(v == -1 for row in d["tile_class"] for v in row)
    assert d["slice_target"] == [0] * capi.QLT_TILE_CLASSES
    assert T.overrides(d) == 0


def test_set_get_round_trip_and_reset(builtin_tuning):
    d = T.get_tuning()
    d["tile_class"][3][5] = 2
    d["tile_class"][15][15] = 0
    d["slice_target"][1] = 512
    d["slice_target"][4] = -1
    T.set_tuning(d)
    assert T.get_tuning() == d
    assert T.overrides(d) == 4
    T.set_tuning(None)
    assert T.overrides(T.get_tuning()) == 0


@pytest.mark.parametrize("field,index,value", [("tile_class", (0, 0), 4), ("tile_class", (2, 1), -2),
                                               ("slice_target", (0,), -5)])
def test_invalid_entries_are_rejected(builtin_tuning, field, index, value):
    t = T.tuning_from_dict(T.get_tuning())
    if field == "tile_class":
        t.tile_class[index[0]][index[1]] = value
    else:
        t.slice_target[index[0]] = value
    with pytest.raises(capi.QltError):
        capi.check(capi.lib().qlt_tuning_set(ctypes.byref(t)))
    assert T.overrides(T.get_tuning()) == 0      # unchanged


def test_json_round_trip(builtin_tuning, tmp_path):
    d = T.get_tuning()
    d["tile_class"][0][1] = 3
    d["slice_target"][2] = 4096
    path = tmp_path / "tuning.json"
    T.save(str(path), d)
    assert T.get_tuning()["tile_class"][0][1] == -1     # saving installs nothing
    assert T.load(str(path)) == d
    assert T.get_tuning() == d


def test_malformed_dict_is_rejected():
    with pytest.raises(ValueError):
        T.tuning_from_dict({"tile_class": [[-1] * 15] * 16, "slice_target": [0] * 6})
    with pytest.raises(ValueError):
        T.tuning_from_dict({"tile_class": [[-1] * 16] * 16, "slice_target": [0] * 5})
