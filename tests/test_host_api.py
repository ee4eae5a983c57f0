"""Host-side helpers of the drop-in API (no GPU): prior, addnoise, musp2sasp, exports."""

import copy
import warnings

import numpy as np
import pytest


def test_prior_laplace_and_helmholtz_rows():
    import redbirdpy_amd as rb

    seg = np.array([3, 1, 1, 2, 2, 2, 7])
    for pt, extra in (("laplace", 0.0), ("helmholtz", 2.5)):
        L = rb.prior(seg, pt, {"beta": 2.5})
        assert np.allclose(L, L.T) and np.all(np.diag(L) == 1)
        for i in range(len(seg)):
            g = np.flatnonzero(seg == seg[i])
            off = [L[i, j] for j in g if j != i]
            assert all(v == pytest.approx(-1.0 / (len(g) + extra)) for v in off)
            assert np.count_nonzero(L[i]) == (len(g) if len(g) > 1 else 1)
    assert rb.prior(seg, "") is None and rb.prior(seg, "unknown") is None


def test_prior_comp_links_close_compositions():
    import redbirdpy_amd as rb

    S = np.array([[1.0, 0.0], [0.95, 0.05], [0.0, 1.0], [0.5, 0.5]])
    C = rb.prior(S, "comp", {"alpha": 0.2, "beta": 1.0}).toarray()
    assert np.allclose(C, C.T) and np.all(np.diag(C) == 1)
    # members 0 and 1 differ by 0.1 < alpha*nc = 0.4: -alpha - 0.1/2 = -0.25, normalised by |rowsum| = 0.25
    assert C[0, 1] == pytest.approx(-1.0)
    assert C[0, 2] == 0 and C[1, 3] == 0


def test_musp2sasp_round_trip():
    import redbirdpy_amd as rb

    lam = np.array([690.0, 830.0])
    sa, sp = rb.musp2sasp(1.7 * (lam / 500) ** (-0.9), lam)
    assert sa == pytest.approx(1.7, rel=1e-13) and sp == pytest.approx(0.9, rel=1e-13)
    with pytest.raises(ValueError):
        rb.musp2sasp(np.array([1.0]), np.array([690.0]))


def test_addnoise_seeded_and_scaled():
    import redbirdpy_amd as rb

    d = np.full(20000, 4.0)
    a = rb.addnoise(d, 20.0)
    assert np.array_equal(a, rb.addnoise(d, 20.0))  # seeded
    # shot noise std = sqrt(|d|) * 10^(-20/20) = 0.2
    assert np.std(a - d) == pytest.approx(0.2, rel=0.05)
    with warnings.catch_warnings(record=True):
        warnings.simplefilter("always")
        assert np.array_equal(rb.addnoise(d, np.inf), d)
    z = rb.addnoise(d * (1 + 1j), 30.0 + 40.0j, 60.0)
    assert np.iscomplexobj(z) and not np.allclose(z, d * (1 + 1j))


def test_reference_exports_present():
    import redbirdpy_amd as rb

    for name in ("prior", "addnoise", "HAS_ISO2MESH", "musp2sasp", "setmesh", "meshinterp", "src2bc", "meshprep",
                 "runforward", "runrecon", "femsolve", "jac", "jacchrome", "femlhs", "femrhs", "femgetdet", "sdmap",
                 "getoptodes", "getreff", "getbulk", "updateprop", "extinction", "reginv", "matreform", "matflat",
                 "syncprop", "run"):
        assert hasattr(rb, name), name


def test_extinction_updateprop_jacchrome_match_oracle_default_table():
    """The built-in chromophore table path (property.py:30-213, forward.py:572-589): extinction with
    interpolation and extrapolation, updateprop with its water/lipids defaults on label and node
    params, and jacchrome's wavelength-stacked E[w, c] J_w, against the oracle."""
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc
    from redbirdpy_amd.mesh import kuhn_box

    wvs = ["650", "690", "785", "808", "830", "905", "1050"]
    chs = ["hbo", "hbr", "water", "lipids", "aa3"]
    np.testing.assert_allclose(rb.extinction(wvs, chs)[0], orc.extinction(wvs, chs)[0], rtol=1e-14)
    node, elem = kuhn_box([0, 0, 0], [4, 4, 2], 1.0)
    nn = node.shape[0]
    rng = np.random.default_rng(4)
    tab = np.array([[0, 0, 1, 1], [0.01, 1.0, 0.0, 1.37], [0.02, 1.2, 0.0, 1.37]])
    for param in ({"hbo": rng.uniform(10, 30, nn), "hbr": rng.uniform(5, 15, nn)},
                  {"hbo": np.array([20.0, 40.0]), "hbr": np.array([10.0, 12.0]), "aa3": np.array([1.0, 2.0])},
                  {"hbo": rng.uniform(10, 30, nn), "hbr": rng.uniform(5, 15, nn),
                   "scatamp": np.full(nn, 1.5), "scatpow": np.full(nn, 1.1)}):
        c1 = {"node": node, "elem": elem, "prop": {w: tab.copy() for w in wvs[:3]}, "param": copy.deepcopy(param)}
        c2 = copy.deepcopy(c1)
        p1, p2 = rb.updateprop(c1), orc.updateprop(c2)
        assert list(p1) == list(p2) == wvs[:3]
        for w in wvs[:3]:
            np.testing.assert_allclose(p1[w], p2[w], rtol=1e-14)
        assert c1["param"]["water"] == 0.23 and c1["param"]["lipids"] == 0.58
    J = {w: rng.standard_normal((6, nn)) for w in wvs[:3]}
    Jc1 = rb.jacchrome(J, ["hbo", "hbr", "water"])
    Jc2 = orc.jacchrome(J, ["hbo", "hbr", "water"])
    for k in Jc2:
        assert Jc1[k].shape == (18, nn)
        np.testing.assert_allclose(Jc1[k], Jc2[k], rtol=1e-15)
    with pytest.raises(ValueError):
        rb.jacchrome(J["650"], ["hbo"])


@pytest.mark.parametrize("multi", [False, True])
def test_sdmap_maxdist_and_exclusions_brute_force(multi):
    """sdmap's excludesrc / excludedet (keyword or cfg field) and maxdist (utility.py:719-796)
    against a brute-force table: excluded optodes leave the grid, pair row = d*Ns' + s over the
    kept ones, a point pair is active iff its distance is below maxdist, wide-field pairs always
    are; multi-wavelength cfgs give one 4-column table per wavelength.  Product and oracle."""
    import redbirdpy_amd as rb
    from oracle import rb_oracle as orc

    rng = np.random.default_rng(5)
    src = rng.uniform(0, 40, (6, 3))
    det = rng.uniform(0, 40, (5, 3))
    cfg = {"srcpos": src, "detpos": det, "prop": np.zeros((2, 4)),
           "widesrc": np.zeros((10, 2)), "excludedet": [3]}  # nodes x 2 patterns
    if multi:
        cfg["prop"] = {"690": np.zeros((2, 4)), "830": np.zeros((2, 4))}
    maxdist = 25.0
    ks, kd = [0, 2, 3, 5], [0, 1, 2, 4]  # excludesrc [1, 4] by keyword, excludedet [3] from cfg
    srcs = ks + [6, 7]  # the two wide-field sources follow the point sources
    want = []
    for d in kd:
        for s in srcs:
            act = 1.0 if s >= 6 else float(np.linalg.norm(src[s] - det[d]) < maxdist)
            want.append([s, 8 + d, act])
    want = np.array(want)
    assert 0 < want[:, 2].sum() < len(want)  # the distance cut bites
    for mod in (rb, orc):
        sd = mod.sdmap(cfg, maxdist=maxdist, excludesrc=[1, 4])
        tabs = sd.values() if multi else [sd]
        if multi:
            assert list(sd.keys()) == ["690", "830"]
        for t in tabs:
            assert t.shape == (len(want), 4 if multi else 3)
            np.testing.assert_array_equal(t[:, :3], want)
            if multi:
                assert np.all(t[:, 3] == 1)
    # no cut: every kept pair active
    assert np.all(rb.sdmap({**cfg, "prop": np.zeros((2, 4))})[:, 2] == 1)
