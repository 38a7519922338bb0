"""Host logic of the batched engine's autotuner: candidate schedules respect the memory budget and cover every
instance exactly once; the tile cache round-trips.  No device calls."""
import json

from _pkg import egs

at = egs.autotune


def test_schedules_partition_under_budget():
    sizes = [8, 8, 8, 1, 1, 1, 4, 4, 2]
    for name, groups in at.schedules(sizes, 10).items():
        flat = sorted(i for g in groups for i in g)
        assert flat == list(range(len(sizes))), name
        for g in groups:
            assert sum(sizes[i] for i in g) <= 10 or len(g) == 1, (name, g)
    ffd = at.schedules(sizes, 10)["first_fit_decreasing"]
    assert len(ffd) == 4                                   # 8+2, 8+1+1, 8+1, 4+4: first-fit decreasing packing


def test_tile_cache_roundtrip(tmp_path):
    path = str(tmp_path / "c" / "tiles.json")
    at._store(path, {"k": [12, 14]})
    assert at._load(path) == {"k": [12, 14]}
    assert json.load(open(path)) == {"k": [12, 14]}
    assert at._load(str(tmp_path / "missing.json")) == {}
