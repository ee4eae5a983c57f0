"""The drop-in contract around the hot path, against the oracle:

* solver failure is visible (SolverWarning from runforward / runrecon /
  forward_batch, the flag from femsolve) and NaN/Inf inputs are refused
  before any launch;
* runrecon's 5th return value is the reference's Jmua (the jacchrome dict of
  chromophore Jacobians, or {'mua': J});
* public reginvover / reginvunder with a prior (LTL / invR, whole or per block);
* the reference's update whitelist (aa3 is not updated) and lmat-over-ltl rule;
* forward_batch on meshes whose RHS norms differ by orders of magnitude.
"""

import copy
import warnings

import numpy as np
import pytest

from conftest import box_cfg, grid_optodes

pytestmark = pytest.mark.gpu


def test_iteration_limit_warns_and_flags():
    import redbirdpy_amd as rb
    from redbirdpy_amd._gpu import SolverWarning

    cfg, _ = rb.meshprep(box_cfg((20, 20, 10), 2.0))
    with pytest.warns(SolverWarning, match="iteration limit"):
        rb.runforward(cfg, maxiter=1)
    A = rb.femlhs(cfg, cfg["deldotdel"])
    b = np.zeros((A.shape[0], 2))
    b[5, 0] = b[77, 1] = 1.0
    x, flag = rb.femsolve(A, b, method="cg", maxiter=1)
    assert flag == 1
    with warnings.catch_warnings():
        warnings.simplefilter("error", SolverWarning)
        rb.runforward(cfg)  # converges: no warning


def test_nan_inputs_rejected_before_launch():
    import redbirdpy_amd as rb

    cfg, _ = rb.meshprep(box_cfg((20, 20, 10), 2.0))
    bad = copy.deepcopy(cfg)
    bad["prop"] = bad["prop"].copy()
    bad["prop"][1, 0] = np.nan
    with pytest.raises(ValueError, match="NaN"):
        rb.runforward(bad)
    bad = copy.deepcopy(cfg)
    bad["srcpos"] = np.array([[10.0, np.inf, 0.0]])
    with pytest.raises(ValueError, match="srcpos"):
        rb.runforward(bad)


def _chrome_problem():
    from redbirdpy_amd import workloads
    from oracle import rb_oracle as orc

    cfg0, bg, truth, extin, E = workloads.chrome_box(h=5.0, n_lattice=4)
    ct, _ = orc.meshprep(copy.deepcopy(cfg0))
    ct["param"] = {k: v.copy() for k, v in truth.items()}
    ct["prop"] = orc.updateprop(ct, extin=extin)
    det, _ = orc.runforward(ct)
    return cfg0, bg, extin, det


def test_runrecon_returns_reference_jmua():
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc

    cfg0, bg, extin, y0 = _chrome_problem()
    c1, sd1 = rb.meshprep(copy.deepcopy(cfg0))
    c2, sd2 = orc.meshprep(copy.deepcopy(cfg0))
    lam = 1e-3 / sum(v.size for v in y0.values())
    out1 = rb.runrecon(c1, {"param": copy.deepcopy(bg)}, y0, sd1, maxiter=1, lambda_=lam, extin=extin,
                       report=False)
    out2 = orc.runrecon(c2, {"param": copy.deepcopy(bg)}, y0, sd2, maxiter=1, lambda_=lam, extin=extin)
    J1, J2 = out1[4], out2[4]
    assert list(J1.keys()) == list(J2.keys()) == ["hbo", "hbr", "water"]
    for k in J2:
        a = J1[k]
        assert a.shape == J2[k].shape == (4 * 64, c1["node"].shape[0])
        assert np.linalg.norm(a - J2[k]) / np.linalg.norm(J2[k]) < 1e-8
    # the same matrices as rb.jacchrome over rb.jac of the returned fields
    phi = out1[6]
    Jw = {wv: rb.jac(sd1[wv], phi[wv], c1["deldotdel"], c1["elem"], c1["evol"], return_elem=False)[0]
          for wv in phi}
    E = np.array([[extin[wv][c] for c in ("hbo", "hbr", "water")] for wv in phi])
    Jc = rb.jacchrome(Jw, ["hbo", "hbr", "water"], extin=E)
    for k in Jc:
        np.testing.assert_allclose(J1[k], Jc[k], rtol=1e-12, atol=1e-14 * np.abs(Jc[k]).max())
    # one wavelength, node mua: {'mua': J}
    cfg = box_cfg((20, 20, 10), 2.0, src=grid_optodes((20, 20, 10), 2), det=grid_optodes((20, 20, 10), 2, margin=6))
    c3, sd3 = rb.meshprep(copy.deepcopy(cfg))
    y, phi3 = rb.runforward(c3)
    nn = c3["node"].shape[0]
    r = {"prop": np.tile([0.012, 1.0, 0, 1.37], (nn, 1))}
    out3 = rb.runrecon(c3, r, y, sd3, maxiter=1, lambda_=1e-3, report=False)
    assert list(out3[4].keys()) == ["mua"] and out3[4]["mua"].shape == (16, nn)


def test_reginv_drop_in_with_priors():
    import redbirdpy_amd as rb
    from scipy.linalg import qr
    from oracle import rb_oracle as orc

    rng = np.random.default_rng(5)
    # overdetermined, whole and per-block L^T L, with a zero column and a zero row
    A = rng.standard_normal((60, 8))
    A[:, 3] = 0
    A[7] = 0
    b = rng.standard_normal(60)
    L = np.eye(8) + 0.3 * rng.standard_normal((8, 8))
    for LTL in (L.T @ L, (L[:4, :4].T @ L[:4, :4])):
        np.testing.assert_allclose(rb.reginvover(A, b, 0.2, LTL), orc.reginvover(A, b, 0.2, LTL), rtol=1e-9,
                                   atol=1e-12)
    # underdetermined, whole and per-block R^-1
    A = rng.standard_normal((12, 90))
    b = rng.standard_normal(12)
    for n in (90, 30):
        Lm = np.eye(n) + 0.2 * rng.standard_normal((n, n))
        invR = np.linalg.inv(np.triu(qr(Lm)[1]))
        blocks = {"hbo": (12, 30), "hbr": (12, 30), "water": (12, 30)}
        got = rb.reginvunder(A, b, 0.1, invR, blocks)
        np.testing.assert_allclose(got, orc.reginvunder(A, b, 0.1, invR, blocks), rtol=1e-8, atol=1e-11)
    # complex (reform 'complex'): unconjugated normal equations
    Ac = rng.standard_normal((10, 40)) + 1j * rng.standard_normal((10, 40))
    bc = rng.standard_normal(10) + 1j * rng.standard_normal(10)
    np.testing.assert_allclose(rb.reginv(Ac, bc, 0.3), orc.reginv(Ac, bc, 0.3), rtol=1e-9, atol=1e-12)


def _label_problem():
    from oracle import rb_oracle as orc

    cfg0 = box_cfg((12, 12, 6), 1.5, src=grid_optodes((12, 12, 6), 2, margin=3),
                   det=grid_optodes((12, 12, 6), 2, margin=4.5))
    c, sd = orc.meshprep(copy.deepcopy(cfg0))
    e0 = c["elem"][:, :4] - 1
    eseg = (c["node"][e0].mean(axis=1)[:, 0] > 6).astype(int)
    nseg = (c["node"][:, 0] > 6).astype(int)
    cfg0["seg"] = eseg
    ct = copy.deepcopy(c)
    ct["seg"] = eseg
    ct["prop"] = np.array([[0.01, 1.0, 0, 1.37], [0.016, 1.0, 0, 1.37]])
    y0, _ = orc.runforward(ct)
    return cfg0, nseg, y0


def test_lmat_takes_precedence_over_ltl_like_the_oracle():
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc

    cfg0, nseg, y0 = _label_problem()
    L = np.array([[1.0, -0.9], [-0.9, 1.0]])
    outs = {}
    for name, kw in (("lmat", {"lmat": L}), ("both", {"lmat": L, "ltl": 5 * np.eye(2)})):
        for mod in (rb, orc):
            c, sd = mod.meshprep(copy.deepcopy(cfg0))
            rec = {"prop": np.array([[0.01, 1.0, 0, 1.37], [0.01, 1.0, 0, 1.37]]), "seg": nseg}
            extra = {"report": False} if mod is rb else {}
            mod.runrecon(c, rec, y0, sd, maxiter=2, lambda_=1e-2, **kw, **extra)
            outs[name, mod.__name__] = rec["prop"][:, 0].copy()
    for name in ("lmat", "both"):
        np.testing.assert_allclose(outs[name, "redbirdpy_amd"], outs[name, "oracle.rb_oracle"], rtol=1e-7)
    np.testing.assert_allclose(outs["both", "redbirdpy_amd"], outs["lmat", "redbirdpy_amd"], rtol=1e-12)


def test_aa3_keeps_its_value_like_the_reference():
    import redbirdpy_amd as rb
    from redbirdpy_amd import workloads
    from oracle import rb_oracle as orc

    cfg0, bg, truth, extin, E = workloads.chrome_box(h=5.0, n_lattice=4)
    for wv in extin:
        extin[wv]["aa3"] = 1e-3
    bg = dict(bg, aa3=np.full(len(bg["hbo"]), 2.0))
    truth = dict(truth, aa3=np.full(len(bg["hbo"]), 2.0))
    ct, _ = orc.meshprep(copy.deepcopy(cfg0))
    ct["param"] = copy.deepcopy(truth)
    ct["prop"] = orc.updateprop(ct, extin=extin)
    y0, _ = orc.runforward(ct)
    lam = 1e-3 / sum(v.size for v in y0.values())
    c1, sd1 = rb.meshprep(copy.deepcopy(cfg0))
    c2, sd2 = orc.meshprep(copy.deepcopy(cfg0))
    r1, r2 = {"param": copy.deepcopy(bg)}, {"param": copy.deepcopy(bg)}
    out1 = rb.runrecon(c1, r1, y0, sd1, maxiter=2, lambda_=lam, extin=extin, report=False)
    orc.runrecon(c2, r2, y0, sd2, maxiter=2, lambda_=lam, extin=extin)
    np.testing.assert_array_equal(r1["param"]["aa3"], bg["aa3"])
    assert "aa3" in out1[3][0] and "aa3" in out1[4]  # its step and Jacobian are still reported
    for k in ("hbo", "hbr", "water"):
        err = np.max(np.abs(r1["param"][k] - r2["param"][k])) / np.max(np.abs(r2["param"][k]))
        assert err < 1e-7, (k, err)


def test_forward_batch_meshes_with_unequal_rhs_norms():
    """A wide-field (unit-L1 pattern) mesh next to point-source meshes: every mesh's own
    relative residual stays within tol (each mesh's part of each column is scaled alike)."""
    import redbirdpy_amd as rb
    from redbirdpy_amd import workloads
    from oracle import rb_oracle as orc

    raw = workloads.batch_boxes(count=3, seed=9, sides=(8.0, 20.0), spacing=(1.2, 2.5), nsrc=(2, 4),
                                nn_range=(100, 4000), kf=0)
    wide = workloads.widefield_box(2.0, "fourier", 2, 2, box=(20.0, 20.0, 10.0))
    wide["srcparam1"] = np.array([10.0, 0.0, 0.0, 2.0])
    wide["srcparam2"] = np.array([0.0, 10.0, 0.0, 2.0])
    wide["srcpos"] = np.array([[5.0, 5.0, 0.0]])
    wide["detpos"] = np.array([[8.0, 12.0, 0.0], [12.0, 8.0, 0.0]])
    raw = raw + [wide]
    cfgs = [rb.meshprep(copy.deepcopy(c))[0] for c in raw]
    res = rb.forward_batch(cfgs)
    for c0, r in zip(raw, res):
        oc, _ = orc.meshprep(copy.deepcopy(c0))
        d, phi = orc.runforward(oc)
        assert np.linalg.norm(r.phi - phi) / np.linalg.norm(phi) < 1e-9
        assert np.max(np.abs(r.detval - d)) <= 1e-9 * np.max(np.abs(d))
