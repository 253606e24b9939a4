"""The launch-shape cache file (erdst_gpu.tuning) with a stand-in context:
a key not in the file is tuned once and stored; a key in the file is
applied with set_tuning and not tuned again."""
import json
from types import SimpleNamespace

import pytest

from erdst_gpu import tuning


class FakeContext:
    def __init__(self, shapes):
        self.shapes, self.tuned, self.applied = shapes, 0, []

    def autotune(self, reps=0):
        self.tuned += 1
        return self.shapes

    def set_tuning(self, shapes):
        self.applied.append(shapes)


def test_key_names_what_the_shapes_depend_on():
    topo = SimpleNamespace(natom=90000, nmol=30000)
    p = dict(ms=(72, 72, 72), splodr=4, cltype=2, boxshp=1, slttype=2)
    k = tuning.tuning_key(topo, p, [5], 1000)
    assert k == ("atoms=90000 molecules=30000 grid=72x72x72 order=4 cltype=2 boxshp=1 slttype=2 "
                 "sites=5 insertions=1000")
    assert tuning.tuning_key(topo, p, [5], 500) != k
    assert tuning.tuning_key(topo, p | {"cltype": 3}, [5], 1000) != k


def test_cache_tunes_once_then_applies(tmp_path):
    path = str(tmp_path / "shapes.json")
    shapes = {"chunk": 1024, "hist_slices": 8, "pair_split": [2]}
    a = FakeContext(shapes)
    got, found = tuning.tune_cached(a, "k1", path)
    assert (got, found, a.tuned, a.applied) == (shapes, False, 1, [])
    b = FakeContext({"chunk": 256, "hist_slices": 1, "pair_split": [1]})
    got, found = tuning.tune_cached(b, "k1", path)
    assert (got, found, b.tuned, b.applied) == (shapes, True, 0, [shapes])
    c = FakeContext({"chunk": 512, "hist_slices": 0, "pair_split": [0]})
    tuning.tune_cached(c, "k2", path)
    assert sorted(json.load(open(path))) == ["k1", "k2"]


def test_cache_file_must_hold_an_object(tmp_path):
    path = tmp_path / "shapes.json"
    path.write_text("[1, 2]")
    with pytest.raises(ValueError):
        tuning.tune_cached(FakeContext({}), "k", str(path))


def test_driver_tuning_cache_needs_autotune(tmp_path, capsys):
    from erdst_gpu import erdst
    with pytest.raises(SystemExit) as e:
        erdst.main([str(tmp_path), "--tuning-cache", str(tmp_path / "shapes.json")])
    assert e.value.code == 2
    assert "--tuning-cache needs --autotune" in capsys.readouterr().err
