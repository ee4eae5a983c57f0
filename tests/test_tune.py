"""Autotuner cache (no GPU): keys, store/lookup through the JSON file, clear."""

import json
from types import SimpleNamespace


def test_tune_cache_round_trip(tmp_path, monkeypatch):
    from redbirdpy_amd import tune

    path = tmp_path / "sub" / "tune.json"
    monkeypatch.setenv("RB_TUNE_CACHE", str(path))
    monkeypatch.setattr(tune, "_device_name", lambda dev: "TestGPU")
    tune.clear()
    sysm = SimpleNamespace(mesh=SimpleNamespace(dev=None, nn=1000, nnz=14000), nsys=1, is_complex=True)
    assert tune.lookup(sysm, 64, "mixed") is None
    tune.store(sysm, 64, "mixed", {"lds": False, "grid": 256, "check_every": 8})
    assert tune.lookup(sysm, 64, "mixed") == {"lds": False, "grid": 256, "check_every": 8}
    # a different shape or precision is another entry
    assert tune.lookup(sysm, 32, "mixed") is None
    assert tune.lookup(sysm, 64, "fp64") is None
    data = json.loads(path.read_text())
    assert data == {"TestGPU|n=1000|nnz=14000|k=64|nsys=1|complex|mixed": {"lds": False, "grid": 256,
                                                                            "check_every": 8}}
    # a fresh load (another process) reads the file
    monkeypatch.setattr(tune, "_CACHE", None)
    assert tune.lookup(sysm, 64, "mixed")["grid"] == 256
    tune.clear()
    assert not path.exists()
    assert tune.lookup(sysm, 64, "mixed") is None


def test_tune_key_names_every_field():
    from redbirdpy_amd import tune

    assert tune.key("G", 5, 7, 64, 4, False, "fp64") == "G|n=5|nnz=7|k=64|nsys=4|real|fp64"
