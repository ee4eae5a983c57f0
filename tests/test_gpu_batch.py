"""forward_batch: uneven meshes (point sources + Fourier wide-field patterns) in one
multi-RHS solve, each mesh against the oracle's own runforward and jac."""

import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(omega):
    from redbirdpy_amd import workloads

    return workloads.batch_boxes(count=5, seed=5, sides=(8.0, 20.0), spacing=(1.2, 2.5), nsrc=(2, 9),
                                 nn_range=(100, 4000), kf=2, omega=omega)


@pytest.mark.parametrize("omega,precision,tol", [(0.0, "fp64", 1e-9), (2 * np.pi * 1e8, "fp64", 1e-9),
                                                 (2 * np.pi * 1e8, "mixed", 1e-9)])
def test_forward_batch_matches_oracle(omega, precision, tol):
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc

    raw = _batch(omega)
    cfgs = [rb.meshprep(copy.deepcopy(c))[0] for c in raw]
    sizes = [c["node"].shape[0] for c in cfgs]
    assert len(set(sizes)) > 1  # uneven
    res = rb.forward_batch(cfgs, precision=precision, jacobian=True)
    for c0, r in zip(raw, res):
        oc, sd = orc.meshprep(copy.deepcopy(c0))
        d, phi = orc.runforward(oc)
        assert r.detval.shape == d.shape and r.phi.shape == phi.shape
        assert np.max(np.abs(r.detval - d)) <= tol * np.max(np.abs(d))
        assert np.linalg.norm(r.phi - phi) / np.linalg.norm(phi) <= tol
        Jo, _ = orc.jac(sd, phi, oc["deldotdel"], oc["elem"], oc["evol"])
        assert np.linalg.norm(r.jac - Jo) / np.linalg.norm(Jo) <= 1e-8


def test_forward_batch_rejects_mixed_omega():
    import redbirdpy_amd as rb

    a, b = _batch(0.0)[:2]
    b["omega"] = 1e8
    cfgs = [rb.meshprep(copy.deepcopy(c))[0] for c in (a, b)]
    with pytest.raises(NotImplementedError):
        rb.forward_batch(cfgs)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_forward_batch_tune_caches_a_choice_and_detector_values_stay(precision, tmp_path, monkeypatch):
    """BatchForward.tune times the batch's solve (its first column group) and caches the choice;
    the later solves under it give the same detector values to the solve tolerance."""
    import redbirdpy_amd as rb
    from redbirdpy_amd import tune
    from redbirdpy_amd.forward import BatchForward

    monkeypatch.setenv("RB_TUNE_CACHE", str(tmp_path / "tune.json"))
    cfgs = [rb.meshprep(copy.deepcopy(c))[0] for c in _batch(2 * np.pi * 1e8)]
    bf = BatchForward(cfgs)
    ref = bf.solve(precision, 1e-11, fields=False)
    out = bf.tune(precision, 1e-11)
    assert out["precond"] in tune.PRECONDS and out["check_every"] in tune.CHECK_EVERY
    assert (tmp_path / "tune.json").exists()
    res = bf.solve(precision, 1e-11, fields=False)
    for a, b in zip(res, ref):
        assert np.max(np.abs(a.detval - b.detval)) <= 1e-9 * np.max(np.abs(b.detval))
