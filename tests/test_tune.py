"""The tuner of the build's tunables (disccofan_amd/tune.py): its knob tables and output on the CPU,
a short tuning run on the GPU (every candidate's output must agree bit for bit with the default's)."""
import numpy as np
import pytest

from disccofan_amd import synth, tune


def test_knob_tables():
    assert tune.knobs_for(False, False, True) is tune.KNOBS_TILED
    assert tune.KNOBS_TILED["DCF_ZIP_SEGMENTS"][0] == "64"
    for is3d, batched in ((False, False), (True, False), (False, True)):
        knobs = tune.knobs_for(is3d, batched)
        for name, values in knobs.items():
            assert name.startswith("DCF_")
            assert len(values) == len(set(values)) >= 2
        assert ("DCF_LEVEL_TILE2D" in knobs) == (not is3d and not batched)
        assert ("DCF_LEVEL_TILE3D" in knobs) == (is3d and not batched)
    # the defaults come first: the library's own (zones 2, records 0, targets 1, tiles 64x64 / 16^3)
    assert tune.KNOBS_2D["DCF_ZONES"][0] == "2" and tune.KNOBS_2D["DCF_LEVEL_TILE2D"][0] == "64x64x1"
    assert tune.KNOBS_2D["DCF_RECORDS"][0] == "0" and tune.KNOBS_2D["DCF_TARGETS"][0] == "1"
    assert tune.KNOBS_3D["DCF_LEVEL_TILE3D"][0] == "16x16x16"
    for shape in tune.KNOBS_2D["DCF_LEVEL_TILE2D"] + tune.KNOBS_3D["DCF_LEVEL_TILE3D"]:
        a, b, d = (int(v) for v in shape.split("x"))
        assert a * b * d <= 1 << 20
    # zone tiles fit the labelling's LDS tables (4096 pixels)
    for shape in tune.KNOBS_2D["DCF_ZONE_TILE2D"] + tune.KNOBS_BATCH["DCF_ZONE_TILE2D"]:
        a, b = (int(v) for v in shape.split("x"))
        assert a * b <= 4096
    assert all(int(v) ** 3 <= 4096 for v in tune.KNOBS_3D["DCF_ZONE_TILE3D"])


def test_env_line():
    assert tune.env_line({"DCF_ZONES": "1", "DCF_RECORDS": "0"}) == "export DCF_RECORDS=0; export DCF_ZONES=1;"


@pytest.mark.gpu
def test_tune_single_image_and_batch():
    img = synth.vnoise(384, 320)
    best, s, table = tune.tune([img], [300.0], connectivity=8, repeats=1,
                               knobs={"DCF_ZONES": ["2", "0"], "DCF_LEVEL_TILE2D": ["64x64x1", "32x32x1"]})
    assert s > 0 and len(table) == 3
    assert best["DCF_ZONES"] in ("2", "0") and best["DCF_LEVEL_TILE2D"] in ("64x64x1", "32x32x1")
    shapes = synth.batch_shapes(6, 64, 200)
    imgs = [synth.disks(H, W, is16, seed) for (W, H, is16, seed, lam) in shapes]
    lams = [lam for (W, H, is16, seed, lam) in shapes]
    best, s, table = tune.tune(imgs, lams, connectivity=4, batched=True, repeats=1)
    assert len(table) == 1 + sum(len(v) - 1 for v in tune.KNOBS_BATCH.values())
    assert set(best) == set(tune.KNOBS_BATCH)
    # the tiled build's merge schedule (every candidate's parent array equal to the default's)
    best, s, table = tune.tune([img], [300.0], connectivity=8, repeats=1, grid=(1, 3),
                               knobs={"DCF_ZIP_SEGMENTS": ["64", "1", "4"]})
    assert len(table) == 3 and best["DCF_ZIP_SEGMENTS"] in ("64", "1", "4")


def _fake_runner(monkeypatch, cost, outputs=None):
    """replace the GPU runner: the time of a setting is cost(setting), its output outputs(setting)"""
    def run(self, setting):
        out = outputs(setting) if outputs else [np.zeros(3)]
        return cost(setting), out
    monkeypatch.setattr(tune._Runner, "run", run)


def test_coordinate_descent_keeps_the_best_of_each_knob(monkeypatch):
    knobs = {"A": ["0", "1", "2"], "B": ["x", "y"]}
    table = {("0", "x"): 5.0, ("1", "x"): 3.0, ("2", "x"): 4.0, ("1", "y"): 2.0, ("0", "y"): 9.0, ("2", "y"): 9.0}
    _fake_runner(monkeypatch, lambda st: table[(st["A"], st["B"])])
    best, s, rows = tune.tune([np.zeros((4, 4), np.uint8)], [1.0], knobs=knobs)
    assert best == {"A": "1", "B": "y"} and s == 2.0
    # the default, two other values of A, then one other value of B with A at its best
    assert [r[0] for r in rows] == [{"A": "0", "B": "x"}, {"A": "1", "B": "x"}, {"A": "2", "B": "x"}, {"A": "1", "B": "y"}]


def test_a_setting_that_changes_the_output_is_refused(monkeypatch):
    knobs = {"A": ["0", "1"]}
    _fake_runner(monkeypatch, lambda st: 1.0, lambda st: [np.full(3, int(st["A"]))])
    with pytest.raises(RuntimeError, match="changed the output"):
        tune.tune([np.zeros((4, 4), np.uint8)], [1.0], knobs=knobs)


def test_cli_sample_sizes(monkeypatch):
    """--size defaults per workload: a 2048^2 image, a 256^3 volume (2048^3 would pass the 2^30-voxel limit), batch
    sides up to 2048; an explicit --size is taken as given"""
    import argparse
    calls = []
    monkeypatch.setattr(tune.synth, "vnoise", lambda H, W: calls.append(("vnoise", H, W)) or np.zeros((1, 1), np.uint16))
    monkeypatch.setattr(tune.synth, "spheres", lambda D, H, W: calls.append(("spheres", D, H, W)) or np.zeros((1, 1, 1), np.uint8))
    monkeypatch.setattr(tune.synth, "batch_shapes", lambda n, lo, hi: calls.append(("batch", n, lo, hi)) or [])
    for wl, size in (("vnoise", 0), ("spheres", 0), ("batch", 0), ("spheres", 64), ("tiled", 0)):
        tune._workload(argparse.Namespace(workload=wl, size=size, count=8, lo=64))
    assert calls == [("vnoise", 2048, 2048), ("spheres", 256, 256, 256), ("batch", 8, 64, 2048), ("spheres", 64, 64, 64),
                     ("vnoise", 4096, 4096)]
    assert 256 ** 3 < (1 << 30)
