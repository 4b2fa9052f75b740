"""Host-side behaviour of the batch tuner's cache file (no device calls)."""
import json

import pytest

pytest.importorskip("kernelkit")
from kernelkit.autotune import BatchTuner, LaunchConfig  # noqa: E402


def test_cache_store_and_reload(tmp_path):
    path = tmp_path / "tune.json"
    t = BatchTuner(cache_file=str(path))
    t._store("k1", (LaunchConfig(2, 1024, 0.5),))
    t._store("k2", (LaunchConfig(1, 0, 1.0), LaunchConfig(8, 2048, 2.0)))
    assert not list(tmp_path.glob("*.tmp"))
    t2 = BatchTuner(cache_file=str(path))
    assert t2._cache == t._cache
    assert t2._cache["k2"][1] == LaunchConfig(8, 2048, 2.0)


@pytest.mark.parametrize("text", ["{not json", "[1, 2]", '{"k": [{"x": 1}]}'])
def test_unreadable_cache_is_ignored(tmp_path, text):
    path = tmp_path / "tune.json"
    path.write_text(text)
    t = BatchTuner(cache_file=str(path))
    assert t._cache == {}
    t._store("k", (LaunchConfig(4, 0, 0.1),))      # replaces the bad file
    assert list(json.loads(path.read_text())) == ["k"]


def test_env_names_cache_file(tmp_path, monkeypatch):
    path = tmp_path / "env.json"
    monkeypatch.setenv("KK_TUNE_CACHE", str(path))
    t = BatchTuner()
    assert t.cache_file == str(path)
    t._store("k", (LaunchConfig(1, 0, 0.2),))
    assert path.exists()
