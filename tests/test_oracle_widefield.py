"""Pin the oracle's wide-field sources (src2bc, utility.py:252-669) without iso2mesh.

The reference's own wide-field tests (test/test_widefield.py) need iso2mesh
and hold no numeric vectors; their checks (shapes, per-pattern L1
normalisation, non-negativity, mapping fields, pattern counts, the
_inpolygon known answers) are repeated here, and the plane cut that iso2mesh
performs is pinned by what geometry fixes:

* every cut point lies on the plane and on its edge, with its weight in [0, 1];
* a plane crossing the whole box cuts it in polygons whose areas add up to
  the box's cross-section (a polygon with corners out of cyclic order, a
  dropped or doubled polygon, a wrong area all change the sum);
* the cut triangles lie in the sink plane at depth z0 = 1/mus', so the
  depth moment sum_i rhs_i z_i of a normalised column equals z0 exactly
  (P1 weights reproduce linear functions): a weight given to the wrong end
  of its edge moves it off z0;
* the x/y moments of a uniform rectangle's column are its centre, up to
  the faces cut by the rectangle's border;
* a Fourier pattern with kx = ky = 1, M = 1, phase 0 is the uniform source.
"""

import numpy as np
import pytest

from oracle import rb_oracle as orc
from redbirdpy_amd import workloads


def test_inpolygon_known_answers():
    # test/test_widefield.py:617-649
    px, py = np.array([0, 1, 1, 0, 0]), np.array([0, 0, 1, 1, 0])
    r = orc.inpolygon(np.array([0.5, 0.5, 1.5, -0.5]), np.array([0.5, 1.5, 0.5, 0.5]), px, py)
    assert r.tolist() == [True, False, False, False]
    px, py = np.array([0, 1, 0.5, 0]), np.array([0, 0, 1, 0])
    r = orc.inpolygon(np.array([0.5, 0.5, 0.1]), np.array([0.3, 0.8, 0.5]), px, py)
    assert r.tolist() == [True, True, False]


# planes n.p = c that stay inside 0 < z < 30 over the whole 60 x 60 footprint
@pytest.mark.parametrize("normal,c", [((0, 0, 1), 0.99), ((0, 0.2, 1), 20.0), ((0.3, -0.1, 1), 20.0)])
def test_qmeshcut_geometry(normal, c):
    cfg = workloads.cw_box(3.0)
    cfg, _ = orc.meshprep(cfg)
    node = cfg["node"]
    n = np.asarray(normal, float)
    plane = np.concatenate([n, [-c]]) / np.linalg.norm(n)
    pos, fd, eid, nid = orc.qmeshcut(cfg["elem"], node, plane)
    w = nid[:, 2] - 1.0
    ends = nid[:, :2].astype(int) - 1
    assert np.all((w >= 0) & (w <= 1))
    np.testing.assert_allclose(pos @ plane[:3] + plane[3], 0, atol=1e-12)
    np.testing.assert_allclose(pos, w[:, None] * node[ends[:, 0]] + (1 - w[:, None]) * node[ends[:, 1]], atol=1e-12)
    # the plane crosses the box through its four side faces: cross-section = 60*60 / cos(tilt)
    tri = fd[:, 2] == fd[:, 3]
    q = np.flatnonzero(~tri)
    f = np.vstack([fd[tri, :3], fd[q][:, [0, 1, 2]], fd[q][:, [0, 2, 3]]])
    area = orc.elemvolume(pos, f).sum()
    assert area == pytest.approx(3600.0 / abs(plane[2]), rel=1e-12)
    # every crossing tet contributes one polygon, and nothing else does
    d = node @ plane[:3] + plane[3]
    up = d[cfg["elem"][:, :4] - 1] >= 0
    assert len(eid) == int(np.sum((up.sum(1) > 0) & (up.sum(1) < 4)))


def _col_moments(cfg, col):
    node = cfg["node"]
    return col.sum(), col @ node[:, 0], col @ node[:, 1], col @ node[:, 2]


@pytest.mark.parametrize("h", [2.0, 3.0])
def test_planar_source_moments_and_fields(h):
    cfg, sd = orc.meshprep(workloads.widefield_box(h))
    w = cfg["widesrc"]
    assert w.shape == (cfg["node"].shape[0], 1)
    assert cfg["srcpos"].shape[0] == 0 and cfg["srcpos0"].shape == (1, 3)
    assert np.array_equal(cfg["wfsrcmapping"], [[0, 0, 0]])
    assert np.all(w >= 0)
    s, mx, my, mz = _col_moments(cfg, w[:, 0])
    assert s == pytest.approx(1.0, rel=1e-14)
    assert mz == pytest.approx(1.0 / (0.01 + 1.0), rel=1e-13)
    assert mx == pytest.approx(40.0, abs=0.05 * h) and my == pytest.approx(40.0, abs=0.05 * h)
    # only nodes of edges crossing the sink plane carry flux
    z = cfg["node"][:, 2]
    assert np.all(w[z > h + 1e-9, 0] == 0)
    # sd: the one wide source against both point detectors (rhs columns 0 | 1, 2)
    assert sd[:, :2].tolist() == [[0, 1], [0, 2]]


def test_fourier_and_pattern_sources():
    base, _ = orc.meshprep(workloads.widefield_box(2.0))
    cfg, _ = orc.meshprep(workloads.widefield_box(2.0, "fourier", kx=2, ky=3))
    w = cfg["widesrc"]
    assert w.shape[1] == 6 and np.array_equal(cfg["wfsrcmapping"], [[0, 0, 5]])
    np.testing.assert_allclose(w.sum(0), 1.0, rtol=1e-14)
    np.testing.assert_allclose(w[:, 0], base["widesrc"][:, 0], rtol=1e-14, atol=1e-18)  # i = j = 0: uniform
    assert not np.allclose(w[:, 1], w[:, 0])
    # i = 1 (x frequency 1, phase 0): 0.5 (1 + cos 2 pi bx) peaks at the rectangle's x edges
    mx1 = w[:, 3] @ cfg["node"][:, 0]
    assert abs(mx1 - 40.0) < 0.5
    # image patterns: the left half of a 1x2 image lights x in [30, 40]
    pat = np.zeros((1, 2, 2))
    pat[0, 0, 0] = 1.0
    pat[0, 1, 1] = 1.0
    cp, _ = orc.meshprep(workloads.widefield_box(2.0, "pattern", pattern=pat))
    wp = cp["widesrc"]
    assert wp.shape[1] == 2
    assert wp[:, 0] @ cp["node"][:, 0] == pytest.approx(35.0, abs=0.3)
    assert wp[:, 1] @ cp["node"][:, 0] == pytest.approx(45.0, abs=0.3)
    ones, _ = orc.meshprep(workloads.widefield_box(2.0, "pattern", pattern=np.ones((3, 4))))
    np.testing.assert_allclose(ones["widesrc"][:, 0], base["widesrc"][:, 0], rtol=1e-14, atol=1e-18)


def test_surface_source_and_weight():
    cfg = workloads.widefield_box(2.0)
    cfg["iscolimated"] = False
    cfg["srcweight"] = 2.5
    cfg, _ = orc.meshprep(cfg)
    w = cfg["widesrc"][:, 0]
    z = cfg["node"][:, 2]
    assert w.sum() == pytest.approx(2.5, rel=1e-14)
    assert np.all(w[z > 0] == 0) and np.all(w >= 0)
    _, mx, my, mz = _col_moments(cfg, w / 2.5)
    assert mz == 0 and abs(mx - 40) < 0.1 and abs(my - 40) < 0.1


def test_wide_detector_and_multiwavelength():
    cfg = workloads.cw_box(3.0)
    cfg.update(dettype="planar", detparam1=np.array([10.0, 0, 0, 0]), detparam2=np.array([0, 10.0, 0, 0]))
    cfg, sd = orc.meshprep(cfg)
    assert cfg["widedet"].shape == (cfg["node"].shape[0], 1)
    assert cfg["detpos"].shape == (1, 3) and cfg["detpos0"].shape == (2, 3)
    assert sd[:, :2].tolist() == [[0, 1], [0, 2]]
    # two wavelengths: one column per wavelength, each sunk by its own 1/mus'
    mw = workloads.widefield_box(3.0)
    mw["prop"] = {"690": np.array([[0, 0, 1, 1], [0.01, 1.0, 0, 1.37]]),
                  "830": np.array([[0, 0, 1, 1], [0.02, 0.5, 0, 1.37]])}
    mw, sdm = orc.meshprep(mw)
    assert mw["widesrc"].shape[1] == 2 and set(mw["wfsrcmapping"]) == {"690", "830"}
    zs = mw["widesrc"].T @ mw["node"][:, 2]
    np.testing.assert_allclose(zs, [1 / 1.01, 1 / 0.52], rtol=1e-13)


def test_fourier_face_values_closed_form():
    # srcdir = +z: the rotation is the identity and a face centroid (cx, cy)
    # has rectangle coordinates ((cx - 30)/20, (cy - 30)/20); pattern i*ky + j
    # is 0.5 (1 + M cos(2 pi (i bx + j by) + phi0)) with M = 1 - frac(ky
    # entry), phi0 = 2 pi frac(kx entry)
    cfg, _ = orc.meshprep(workloads.cw_box(2.0))
    p1 = np.array([20.0, 0, 0, 2.25])
    p2 = np.array([0, 20.0, 0, 3.4])
    bc, npat, pface, parea, pnode, nodeid, sink = orc._planar_source(
        cfg, np.array([30.0, 30, 0]), np.array([0, 0, 1.0]), p1, p2, "fourier", None, 1 / 1.01)
    assert npat == 6 and sink
    c = orc.meshcentroid(pnode, pface)
    lit = bc[0] != 0
    inside = (c[:, 0] > 30) & (c[:, 0] < 50) & (c[:, 1] > 30) & (c[:, 1] < 50)
    assert np.array_equal(lit, inside)
    bx, by = (c[lit, 0] - 30) / 20, (c[lit, 1] - 30) / 20
    M, phi0 = 1 - 0.4, 0.25 * 2 * np.pi
    for i in range(2):
        for j in range(3):
            ref = 0.5 * (1 + M * np.cos(2 * np.pi * (i * bx + j * by) + phi0))
            np.testing.assert_allclose(bc[i * 3 + j, lit], ref, rtol=1e-12, atol=1e-14)
