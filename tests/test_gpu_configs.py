"""The benchmark configurations at their own shapes, against the oracle.

* the plain CW forward on the 2 mm box (one point source, two point
  detectors): float64 GPU forward vs the oracle's direct (SuperLU) solve;
* the adjoint Jacobian of 32 x 32 = 1024 source-detector pairs on the 2 mm
  box (the Jacobian kernel's full 32 x 32 source x detector tile): the whole
  matrix vs the oracle's jac, the autograd gradient vs rb.jac, and central
  differences on 64 seeded nodes (default_rng(3)) at a 1e-6 relative step.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_plain_cw_forward_matches_direct_solve():
    import redbirdpy_amd as rb
    from redbirdpy_amd import workloads
    from oracle import rb_oracle as orc

    c1, sd1 = rb.meshprep(workloads.cw_box(2.0))
    c2, sd2 = orc.meshprep(workloads.cw_box(2.0))
    assert c1["node"].shape[0] == 15376 and c1["elem"].shape[0] == 81000
    det, phi = rb.runforward(c1)
    odet, ophi = orc.runforward(c2)
    assert det.shape == (2, 1) and phi.shape == (15376, 3)
    assert np.linalg.norm(phi - ophi) / np.linalg.norm(ophi) < 1e-10
    assert np.max(np.abs(det - odet) / np.abs(odet)) < 1e-10
    assert np.all(det > 0)
    # reciprocity: the detector fields evaluated at the source give the same values
    A = orc.femlhs(c2, c2["deldotdel"])
    rhs, _, _, _ = orc.femrhs(c2, sd2)
    np.testing.assert_allclose(rhs[:, 0].toarray().ravel() @ phi[:, 1:], det[:, 0], rtol=1e-10)
    assert np.linalg.norm(A @ phi - rhs.toarray()) / np.linalg.norm(rhs.toarray()) < 1e-11


@pytest.fixture(scope="module")
def jac_case():
    import redbirdpy_amd as rb
    from redbirdpy_amd import workloads
    from oracle import rb_oracle as orc

    cfg, sd = rb.meshprep(workloads.jac_box(2.0))
    ocfg, osd = orc.meshprep(workloads.jac_box(2.0))
    det, phi = rb.runforward(cfg)
    odet, ophi = orc.runforward(ocfg)
    oJ, _ = orc.jac(osd, ophi, ocfg["deldotdel"], ocfg["elem"], ocfg["evol"])
    return rb, cfg, sd, det, phi, oJ


def test_jacobian_1024_pairs_full_matrix(jac_case):
    rb, cfg, sd, det, phi, oJ = jac_case
    assert sd.shape == (1024, 3) and phi.shape[1] == 64
    Jn, Je = rb.jac(sd, phi, cfg["deldotdel"], cfg["elem"], cfg["evol"])
    assert Jn.shape == (1024, 15376) and Je.shape == (1024, 81000)
    assert np.linalg.norm(Jn - oJ) / np.linalg.norm(oJ) < 1e-8
    assert np.max(np.abs(Jn - oJ)) < 1e-8 * np.max(np.abs(oJ))
    assert np.all(Jn <= 1e-10 * np.max(np.abs(Jn)))  # CW sensitivities are not positive
    np.testing.assert_allclose(Jn.sum(axis=1), Je.sum(axis=1), rtol=1e-10)  # 0.25 * 4 nodes per tet


def test_jacobian_1024_pairs_autograd_and_finite_differences(jac_case):
    import torch

    rb, cfg, sd, det, phi, oJ = jac_case
    nn = cfg["node"].shape[0]
    Jn, _ = rb.jac(sd, phi, cfg["deldotdel"], cfg["elem"], cfg["evol"], return_elem=False)
    model = rb.autograd.MuaForward(cfg, tol=1e-14)
    mua = torch.full((nn,), 0.01, dtype=torch.float64, device="cuda", requires_grad=True)
    d = rb.autograd.detval_mua(cfg, mua, model=model)
    assert d.shape == (32, 32)
    rng = np.random.default_rng(3)
    nodes = rng.choice(nn, 64, replace=False)
    w = torch.as_tensor(np.random.default_rng(31).standard_normal((32, 32)), device="cuda")
    (w * d).sum().backward()
    g = mua.grad.cpu().numpy()
    ref = w.cpu().numpy().ravel() @ Jn  # row d * Nsrc + s of sd is detval[d, s]
    assert np.max(np.abs(g - ref)) <= 1e-10 * np.max(np.abs(ref))
    h = 1e-6 * 0.01
    for n in nodes:
        vals = []
        for sgn in (+1, -1):
            m = mua.detach().clone()
            m[n] += sgn * h
            vals.append((w * rb.autograd.detval_mua(cfg, m, model=model)).sum().item())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(g[n] - fd) <= 1e-5 * np.max(np.abs(g)) + 1e-5 * abs(fd), (n, g[n], fd)
