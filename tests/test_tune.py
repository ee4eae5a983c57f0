"""Autotuner cache (no GPU): keys, store/lookup through the JSON file, clear; the preconditioner
settings it applies."""

import json
from types import SimpleNamespace

import pytest


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


def test_applied_sets_and_restores_each_preconditioner():
    """tune.applied sets the process-wide preconditioner a choice names (host-side settings, no GPU
    work) and restores the previous one; every name the bench reports has a setting."""
    from redbirdpy_amd import _abi, tune

    L = _abi.lib()
    assert set(tune.PRECOND_NAMES) == set(tune.PRECONDS)
    before = (L.rb_krylov_get_precond(), L.rb_krylov_get_sweep_grid())
    for pc in tune.PRECONDS:
        with tune.applied({"precond": pc, "grid": 256}):
            assert L.rb_krylov_get_precond() == pc
            assert L.rb_krylov_get_sweep_grid() == 256
        assert (L.rb_krylov_get_precond(), L.rb_krylov_get_sweep_grid()) == before
    with tune.applied({"grid": 512}):  # a choice without a preconditioner leaves it alone
        assert L.rb_krylov_get_precond() == before[0]
    with pytest.raises(_abi.RbError):
        _abi.call("rb_krylov_set_precond", max(tune.PRECONDS) + 1)
    assert L.rb_krylov_get_precond() == before[0]
