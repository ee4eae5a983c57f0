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
                                                  ("logphase", "fp64", 1e-6), ("logphase", "mixed", 1e-6),
                                                  ("reim", "fp64", 1e-7), ("reim", "mixed", 1e-6)])
def test_runrecon_fd_chromophores_match_oracle(reform, precision, tol):
    """Frequency-domain data: [Re; Im] rows (reform 'real'), log-amplitude/phase rows, or
    [[Re, -Im], [Im, Re]] rows on [Re; Im] unknowns ('reim': 4 wavelengths x 3 chromophore blocks,
    each block taking the real part of its complex Tikhonov step, recon.py:243-251).

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


def test_runrecon_fd_chromophores_complex_matches_oracle():
    """reform 'complex' (recon.py:192-198): the unconjugated complex normal equations with the
    complex blockscale 1/sqrt(sum J^2); the concentrations become complex, as in the reference.
    A second iteration would assemble complex optical properties, which the GPU refuses."""
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc

    cfg0, bg, extin, y0 = _fd_chrome()
    c1, sd1 = rb.meshprep(copy.deepcopy(cfg0))
    c2, sd2 = orc.meshprep(copy.deepcopy(cfg0))
    lam = 1e-3 / sum(v.size for v in y0.values())
    r1 = {"param": {k: v.copy() for k, v in bg.items()}}
    r2 = {"param": {k: v.copy() for k, v in bg.items()}}
    out1 = rb.runrecon(c1, r1, y0, sd1, maxiter=1, lambda_=lam, extin=extin, reform="complex", report=False)
    out2 = orc.runrecon(c2, r2, y0, sd2, maxiter=1, lambda_=lam, extin=extin, reform="complex")
    for k in bg:
        a, b = r1["param"][k], r2["param"][k]
        assert np.iscomplexobj(a) and np.iscomplexobj(b)
        d1, d2 = a - bg[k], b - bg[k]  # the complex update itself
        err = np.max(np.abs(d1 - d2)) / np.max(np.abs(d2))
        assert err < 1e-7, (k, err)
        assert np.max(np.abs(d2.imag)) > 1e-3 * np.max(np.abs(d2.real))  # a genuinely complex step
    np.testing.assert_allclose(out1[1], out2[1], rtol=1e-9)
    with pytest.raises(NotImplementedError):
        rb.runrecon(c1, r1, y0, sd1, maxiter=1, lambda_=lam, extin=extin, reform="complex", report=False)


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
    # a complex mua step added in place to the real prop array: numpy's casting error, as the
    # reference's recon['prop'][:, 0] += dx (recon.py:276) raises
    with pytest.raises(TypeError):
        rb.runrecon(c1, copy.deepcopy(recon), y0, sd1, maxiter=1, reform="complex", report=False)
    with pytest.raises(NotImplementedError):  # the reference's runrecon cannot run two modes (jac gets a dict)
        rb.runrecon(c1, copy.deepcopy(recon), y0, sd1, maxiter=1, rfcw=[1, 2], report=False)


def _label_problem():
    """Two tissue labels (0 background, 1 inclusion) on a 20x20x10 mm box; node labels
    in recon['seg'] compress the Jacobian to two columns (an overdetermined step)."""
    from conftest import box_cfg, grid_optodes
    from redbirdpy_amd.mesh import kuhn_box
    from oracle import rb_oracle as orc

    p1 = (20, 20, 10)
    node, elem = kuhn_box([0, 0, 0], p1, 2.0)
    cen = node[elem[:, :4] - 1].mean(axis=1)
    elab = (np.sum((cen - [10, 10, 4]) ** 2, axis=1) < 16).astype(int)
    nlab = (np.sum((node - [10, 10, 4]) ** 2, axis=1) < 16).astype(int)
    truth = np.array([[0.01, 1.0, 0.0, 1.37], [0.02, 1.0, 0.0, 1.37]])
    cfg0 = box_cfg(p1, 2.0, prop=truth, labels=elab, src=grid_optodes(p1, 3), det=grid_optodes(p1, 3, margin=6.0))
    ct, _ = orc.meshprep(copy.deepcopy(cfg0))
    y0, _ = orc.runforward(ct)
    cfg0["prop"] = np.array([[0.01, 1.0, 0.0, 1.37], [0.01, 1.0, 0.0, 1.37]])
    recon = {"prop": cfg0["prop"].copy(), "seg": nlab}
    return cfg0, recon, y0


@pytest.mark.parametrize("reg", ["none", "ltl", "lmat"])
def test_runrecon_label_compression_matches_oracle(reg):
    """Label-based mua (recon.py:210-218 compression, reginvover with an optional L^T L)."""
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc

    cfg0, recon, y0 = _label_problem()
    kw = {}
    if reg == "ltl":
        kw["ltl"] = np.diag([1.0, 2.0])
    elif reg == "lmat":
        kw["lmat"] = np.array([[1.0, -0.5], [-0.5, 1.0]])
    c1, sd1 = rb.meshprep(copy.deepcopy(cfg0))
    c2, sd2 = orc.meshprep(copy.deepcopy(cfg0))
    r1, r2 = copy.deepcopy(recon), copy.deepcopy(recon)
    out1 = rb.runrecon(c1, r1, y0, sd1, maxiter=3, lambda_=0.05, report=False, **kw)
    out2 = orc.runrecon(c2, r2, y0, sd2, maxiter=3, lambda_=0.05, **kw)
    np.testing.assert_allclose(r1["prop"], r2["prop"], rtol=1e-8)
    np.testing.assert_allclose(out1[1], out2[1], rtol=1e-8)
    assert r1["prop"][1, 0] > 0.0125  # the inclusion label moves toward its true 0.02
    assert out1[3][0]["mua"].shape == (2,)


def test_runrecon_node_prior_matches_oracle():
    """Underdetermined node-based mua with an explicit Laplace prior over two node groups:
    J' = J R^-1 (R from QR of L), dmu = R^-1 J'^T y (recon.py:225-233, :458-499)."""
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc

    cfg0, recon, y0 = _mua_problem(0.0, False)
    ct, _ = orc.meshprep(copy.deepcopy(cfg0))
    grp = (np.sum((ct["node"] - [10, 10, 4]) ** 2, axis=1) < 16).astype(int)
    L = rb.prior(grp, "laplace")
    c1, sd1 = rb.meshprep(copy.deepcopy(cfg0))
    c2, sd2 = orc.meshprep(copy.deepcopy(cfg0))
    r1, r2 = copy.deepcopy(recon), copy.deepcopy(recon)
    lam = 1e-3 / y0.size
    out1 = rb.runrecon(c1, r1, y0, sd1, maxiter=2, lambda_=lam, report=False, lmat=L)
    out2 = orc.runrecon(c2, r2, y0, sd2, maxiter=2, lambda_=lam, lmat=L)
    a, b = r1["prop"][:, 0], r2["prop"][:, 0]
    assert np.max(np.abs(a - b)) / np.max(np.abs(b - 0.01)) < 1e-7
    np.testing.assert_allclose(out1[1], out2[1], rtol=1e-7)
    # the prior changes the step
    r3 = copy.deepcopy(recon)
    c3, sd3 = rb.meshprep(copy.deepcopy(cfg0))
    rb.runrecon(c3, r3, y0, sd3, maxiter=1, lambda_=lam, report=False)
    r4 = copy.deepcopy(recon)
    c4, sd4 = rb.meshprep(copy.deepcopy(cfg0))
    rb.runrecon(c4, r4, y0, sd4, maxiter=1, lambda_=lam, report=False, lmat=L)
    assert not np.allclose(r3["prop"][:, 0], r4["prop"][:, 0], rtol=1e-4)


@pytest.mark.parametrize("kw", [{"lmat": np.eye(1)}, {"ltl": np.eye(1)}])
def test_runrecon_reference_prior_cases_match_oracle(kw):
    """The reference's own prior cases (test/test_recon.py:723-767): a label table with a
    1x1 lmat or ltl on a node Jacobian, where the prior's size matches no unknown block
    and the label rows take the first node updates (recon.py:255-266)."""
    import redbirdpy_amd as rb
    from conftest import box_cfg
    from oracle import rb_oracle as orc

    cfg0 = box_cfg((20, 20, 10), 2.0)
    ct, sdt = orc.meshprep(copy.deepcopy(cfg0))
    y0, _ = orc.runforward(ct, sd=sdt)
    y0 = y0 * 0.95
    recon = {"prop": cfg0["prop"].copy(), "lambda": 0.1, "seg": np.arange(1)}
    c1, sd1 = rb.meshprep(copy.deepcopy(cfg0))
    c2, sd2 = orc.meshprep(copy.deepcopy(cfg0))
    r1, r2 = copy.deepcopy(recon), copy.deepcopy(recon)
    out1 = rb.runrecon(c1, r1, y0, sd1, maxiter=1, report=False, **kw)
    out2 = orc.runrecon(c2, r2, y0, sd2, maxiter=1, **kw)
    np.testing.assert_allclose(r1["prop"], r2["prop"], rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(out1[1], out2[1], rtol=1e-8)
