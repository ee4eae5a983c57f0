"""Autotuner on the device: tune_forward measures the launch choices of a 64-column FD solve and
caches the fastest; solves with the cached choice, and with a forced non-default one (the
LDS-staged SpMM off, the sweep's grid capped at 256 workgroups, polls every 8 iterations), agree
with the oracle's SuperLU solve."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_tune_forward_caches_a_choice_and_results_stay_exact(tmp_path, monkeypatch):
    import copy
    from oracle import rb_oracle as orc
    import redbirdpy_amd as rb
    from redbirdpy_amd import _abi, _gpu, tune, workloads
    from redbirdpy_amd.forward import forward_system

    monkeypatch.setenv("RB_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.delenv("RB_AUTOTUNE", raising=False)
    tune.clear()
    cfg = workloads.fd_box(2.0)
    c1, _ = rb.meshprep(copy.deepcopy(cfg))
    c2, _ = orc.meshprep(copy.deepcopy(cfg))
    d2, _ = orc.runforward(c2)
    out = tune.tune_forward(c1, precision="mixed", tol=1e-12)
    assert set(out) >= {"lds", "grid", "check_every", "seconds_per_solve", "candidates"}
    assert out["grid"] in tune.GRID_CAPS and out["check_every"] in tune.CHECK_EVERY
    assert len(out["candidates"]) >= len(tune.CHECK_EVERY)
    assert _abi.lib().rb_krylov_get_sweep_grid() == 0  # restored after every candidate
    mesh, sysm, B, k, *_ = forward_system(c1, (1,), "mixed")
    assert tune.lookup(sysm, k, "mixed") == {key: out[key] for key in ("lds", "grid", "check_every", "precond")}
    d1, _ = rb.runforward(c1, precision="mixed", tol=1e-12)
    assert np.max(np.abs(d1 - d2)) / np.max(np.abs(d2)) < 1e-9
    other = _abi.RB_PRECOND_POLY if out["precond"] == _abi.RB_PRECOND_JACOBI else _abi.RB_PRECOND_JACOBI
    tune.store(sysm, k, "mixed", {"lds": False, "grid": 256, "check_every": 8, "precond": other})
    d3, _ = rb.runforward(c1, precision="mixed", tol=1e-12)
    assert np.max(np.abs(d3 - d2)) / np.max(np.abs(d2)) < 1e-9
    assert _abi.lib().rb_krylov_get_sweep_grid() == 0
    tune.clear()
