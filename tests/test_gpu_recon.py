"""Gauss-Newton reconstruction on the GPU against the oracle's runrecon."""

import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(h=5.0, n_lattice=4):
    from redbirdpy_amd import workloads
    from oracle import rb_oracle as orc

    cfg0, bg, truth, extin, E = workloads.chrome_box(h=h, n_lattice=n_lattice)
    ct, _ = orc.meshprep(copy.deepcopy(cfg0))
    ct["param"] = {k: v.copy() for k, v in truth.items()}
    ct["prop"] = orc.updateprop(ct, extin=extin)
    det, _ = orc.runforward(ct)
    rng = np.random.default_rng(40)
    y0 = {wv: det[wv] * (1 + 0.005 * rng.standard_normal(det[wv].shape)) for wv in workloads.WAVELENGTHS}
    return cfg0, bg, extin, y0


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-7), ("mixed", 1e-6)])
def test_runrecon_chromophores_match_oracle(precision, tol):
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc

    cfg0, bg, extin, y0 = _problem()
    c1, sd1 = rb.meshprep(copy.deepcopy(cfg0))
    c2, sd2 = orc.meshprep(copy.deepcopy(cfg0))
    m = sum(v.size for v in y0.values())
    lam = 1e-3 / m
    r1 = {"param": {k: v.copy() for k, v in bg.items()}}
    r2 = {"param": {k: v.copy() for k, v in bg.items()}}
    out1 = rb.runrecon(c1, r1, y0, sd1, maxiter=3, lambda_=lam, extin=extin, precision=precision, report=False)
    out2 = orc.runrecon(c2, r2, y0, sd2, maxiter=3, lambda_=lam, extin=extin)
    for k in bg:
        a, b = r1["param"][k], r2["param"][k]
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        assert err < tol, (k, err)
    # the misfit sum is a difference of near-equal numbers at the noise floor
    # (0.5% noise): a detector-value error e becomes ~e/0.005 in it, and the
    # Krylov solves guarantee a 1e-12 relative residual, not a 1e-12 field
    np.testing.assert_allclose(out1[1], out2[1], rtol=1e-7 if precision == "fp64" else 1e-5)
    assert out1[1][-1] < out1[1][0]  # the misfit decreases


def test_reginv_drop_in():
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc

    rng = np.random.default_rng(2)
    for m, n in ((40, 300), (300, 40)):
        A = rng.standard_normal((m, n))
        b = rng.standard_normal(m)
        np.testing.assert_allclose(rb.reginv(A, b, 0.1), orc.reginv(A, b, 0.1), rtol=1e-9, atol=1e-12)


def _fd_chrome(h=5.0, small=False):
    """Frequency-domain chromophore data.  small: a 20x20x10 mm box (h 2 mm) with 4+4 optodes
    whose detector values all lie within ~1e-5 of the largest, so that every pair is resolved
    by the solver's relative residual (log-amplitude rows need each value to relative accuracy;
    in the 60 mm box the farthest pairs are ~1e-15 of the peak, below any iterative tolerance)."""
    from conftest import grid_optodes
    from redbirdpy_amd import workloads
    from oracle import rb_oracle as orc

    if small:
        p1 = (20, 20, 10)
        cfg0, bg, truth, extin, E = workloads.chrome_box(h=2.0, box=p1, n_lattice=2)
        cfg0["srcpos"] = grid_optodes(p1, 2)
        cfg0["detpos"] = grid_optodes(p1, 2, margin=7.0)
        near = np.sum((cfg0["node"] - [10, 10, 4]) ** 2, axis=1) < 16
        truth = {k: np.where(near, 2 * v, v) for k, v in bg.items()}
    else:
        cfg0, bg, truth, extin, E = workloads.chrome_box(h=h, n_lattice=4)
    cfg0["omega"] = 2 * np.pi * 1e8
    ct, _ = orc.meshprep(copy.deepcopy(cfg0))
    ct["param"] = {k: v.copy() for k, v in truth.items()}
    ct["prop"] = orc.updateprop(ct, extin=extin)
    det, _ = orc.runforward(ct)
    rng = np.random.default_rng(41)
    y0 = {wv: det[wv] * (1 + 0.005 * rng.standard_normal(det[wv].shape)) for wv in workloads.WAVELENGTHS}
    return cfg0, bg, extin, y0


@pytest.mark.parametrize("reform,precision,tol", [("real", "fp64", 1e-7), ("real", "mixed", 1e-6),
                                                  ("logphase", "fp64", 1e-6), ("logphase", "mixed", 1e-6)])
def test_runrecon_fd_chromophores_match_oracle(reform, precision, tol):
    """Frequency-domain data: [Re; Im] rows (reform 'real') or log-amplitude/phase rows.

    Log-amplitude rows are scaled by 1/|y|: the solver's 1e-12 relative
    residual becomes ~1e-12 * max|y|/|y| relative error of a weak detector
    value (the box's values span 1.8e5), so 'logphase' parity is ~1e-7."""
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc

    cfg0, bg, extin, y0 = _fd_chrome(small=reform == "logphase")
    c1, sd1 = rb.meshprep(copy.deepcopy(cfg0))
    c2, sd2 = orc.meshprep(copy.deepcopy(cfg0))
    m = 2 * sum(v.size for v in y0.values())
    # log-amplitude rows of the 16-pair box need more damping to stay a descent
    # (at 1e-3/m the oracle's own iteration diverges to negative concentrations)
    lam = 0.1 if reform == "logphase" else 1e-3 / m
    r1 = {"param": {k: v.copy() for k, v in bg.items()}}
    r2 = {"param": {k: v.copy() for k, v in bg.items()}}
    # the mixed forwards default to a 1e-10 residual; log rows amplify it by the
    # detector range, so they are solved to 1e-12 like the float64 ones
    out1 = rb.runrecon(c1, r1, y0, sd1, maxiter=3, lambda_=lam, extin=extin, precision=precision, reform=reform,
                       report=False, solvertol=1e-12)
    out2 = orc.runrecon(c2, r2, y0, sd2, maxiter=3, lambda_=lam, extin=extin, reform=reform)
    for k in bg:
        a, b = r1["param"][k], r2["param"][k]
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        assert err < tol, (k, err)
    np.testing.assert_allclose(out1[1], out2[1], rtol=1e-7 if (precision, reform) == ("fp64", "real") else 1e-5)
    assert out1[1][-1] < out1[1][0]


def _mua_problem(omega, dual):
    """Single-wavelength node-based mua on a 20x20x10 mm box (h 2 mm), optionally with a
    5 mm recon mesh over the same box (mapid/mapweight located by the oracle's tsearchn)."""
    from conftest import box_cfg, grid_optodes
    from redbirdpy_amd.mesh import kuhn_box
    from oracle import rb_oracle as orc

    p1 = (20, 20, 10)
    cfg0 = box_cfg(p1, 2.0, src=grid_optodes(p1, 3), det=grid_optodes(p1, 3, margin=6.0), omega=omega)
    ct, _ = orc.meshprep(copy.deepcopy(cfg0))
    nn = ct["node"].shape[0]
    prop = np.tile([0.01, 1.0, 0.0, 1.37], (nn, 1))
    near = np.sum((ct["node"] - [10, 10, 4]) ** 2, axis=1) < 16
    prop[near, 0] = 0.02
    ct["prop"] = prop
    y0, _ = orc.runforward(ct)
    recon = {}
    if dual:
        cnode, celem = kuhn_box([0, 0, 0], p1, 5.0)
        loc, bary = orc.tsearchn(cnode, celem, ct["node"])
        recon = {"node": cnode, "elem": celem, "mapid": loc, "mapweight": bary,
                 "prop": np.tile([0.01, 1.0, 0.0, 1.37], (cnode.shape[0], 1))}
    else:
        recon = {"prop": np.tile([0.01, 1.0, 0.0, 1.37], (nn, 1))}
    return cfg0, recon, y0


@pytest.mark.parametrize("omega,dual,reform", [(0.0, True, "real"), (2 * np.pi * 1e8, True, "real"),
                                               (2 * np.pi * 1e8, False, "logphase"),
                                               (2 * np.pi * 1e8, False, "real")])
def test_runrecon_mua_fd_and_dual_mesh_match_oracle(omega, dual, reform):
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc

    cfg0, recon, y0 = _mua_problem(omega, dual)
    c1, sd1 = rb.meshprep(copy.deepcopy(cfg0))
    c2, sd2 = orc.meshprep(copy.deepcopy(cfg0))
    r1, r2 = copy.deepcopy(recon), copy.deepcopy(recon)
    lam = 1e-3 / y0.size
    out1 = rb.runrecon(c1, r1, y0, sd1, maxiter=3, lambda_=lam, reform=reform, report=False)
    out2 = orc.runrecon(c2, r2, y0, sd2, maxiter=3, lambda_=lam, reform=reform)
    a, b = r1["prop"][:, 0], r2["prop"][:, 0]
    assert a.shape == (recon["prop"].shape[0],)
    assert np.max(np.abs(a - b)) / np.max(np.abs(b)) < 1e-7
    np.testing.assert_allclose(out1[1], out2[1], rtol=1e-7)
    assert out1[1][-1] < out1[1][0]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-7), ("mixed", 1e-6)])
def test_runrecon_mua_reim_matches_oracle(precision, tol):
    """FD data with [[Re, -Im], [Im, Re]] rows: the real part of the complex update."""
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc

    cfg0, recon, y0 = _mua_problem(2 * np.pi * 1e8, False)
    c1, sd1 = rb.meshprep(copy.deepcopy(cfg0))
    c2, sd2 = orc.meshprep(copy.deepcopy(cfg0))
    r1, r2 = copy.deepcopy(recon), copy.deepcopy(recon)
    lam = 1e-3 / y0.size
    out1 = rb.runrecon(c1, r1, y0, sd1, maxiter=3, lambda_=lam, reform="reim", report=False,
                       precision=precision, solvertol=1e-12)
    out2 = orc.runrecon(c2, r2, y0, sd2, maxiter=3, lambda_=lam, reform="reim")
    a, b = r1["prop"][:, 0], r2["prop"][:, 0]
    assert np.max(np.abs(a - b)) / np.max(np.abs(b)) < tol
    np.testing.assert_allclose(out1[1], out2[1], rtol=tol)
    assert out1[1][-1] < out1[1][0]
    # not the 'real' update: the imaginary rows couple to the second unknown block
    r3 = copy.deepcopy(recon)
    c3, sd3 = rb.meshprep(copy.deepcopy(cfg0))
    rb.runrecon(c3, r3, y0, sd3, maxiter=1, lambda_=lam, reform="real", report=False)
    r4 = copy.deepcopy(recon)
    c4, sd4 = rb.meshprep(copy.deepcopy(cfg0))
    rb.runrecon(c4, r4, y0, sd4, maxiter=1, lambda_=lam, reform="reim", report=False)
    assert not np.allclose(r3["prop"][:, 0], r4["prop"][:, 0], rtol=1e-3)


def test_runrecon_unsupported_forms_raise():
    import redbirdpy_amd as rb

    cfg0, recon, y0 = _mua_problem(2 * np.pi * 1e8, True)
    c1, sd1 = rb.meshprep(copy.deepcopy(cfg0))
    with pytest.raises(NotImplementedError):
        rb.runrecon(c1, copy.deepcopy(recon), y0, sd1, maxiter=1, reform="reim", report=False)
    cfg0, recon, y0 = _mua_problem(2 * np.pi * 1e8, False)
    c1, sd1 = rb.meshprep(copy.deepcopy(cfg0))
    with pytest.raises(NotImplementedError):
        rb.runrecon(c1, copy.deepcopy(recon), y0, sd1, maxiter=1, reform="complex", report=False)
