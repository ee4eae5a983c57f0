"""The LDS-staged SpMM (k_cg_spmm_lds through rb_krylov_solve_planned) inside the Krylov solves it
serves: 64-column solves of one system with complex64 inner iterations (mixed) or float64 values
(plain).  Checked against the oracle (SuperLU), against the same solve with the gather SpMM
(RB_LDS_SPMM=0), and on an irregular mesh whose rows span more than one 16-record group."""

import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _relmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _forward_both(cfg, precision, monkeypatch):
    import redbirdpy_amd as rb
    from redbirdpy_amd import _gpu

    c1, _ = rb.meshprep(copy.deepcopy(cfg))
    assert _gpu.mesh_for(c1).spmm_plan() is not None
    d1, f1 = rb.runforward(c1, precision=precision, tol=1e-12)
    it1 = rb.forward.last_info.iters
    monkeypatch.setenv("RB_LDS_SPMM", "0")
    c2, _ = rb.meshprep(copy.deepcopy(cfg))
    d2, f2 = rb.runforward(c2, precision=precision, tol=1e-12)
    it2 = rb.forward.last_info.iters
    monkeypatch.delenv("RB_LDS_SPMM")
    return (d1, f1, it1), (d2, f2, it2)


@pytest.mark.parametrize("h", [2.0, 1.5])
def test_fd_mixed_64col_against_oracle_and_gather(h, monkeypatch):
    """The benchmark's solve shape (FD box, 32 sources + 32 detectors, complex64 COCG with float64
    reliable updates) at two mesh sizes."""
    from oracle import rb_oracle as orc
    from redbirdpy_amd import workloads

    cfg = workloads.fd_box(h)
    (d1, f1, it1), (d2, f2, it2) = _forward_both(cfg, "mixed", monkeypatch)
    c3, _ = orc.meshprep(copy.deepcopy(cfg))
    d3, f3 = orc.runforward(c3)
    assert f1.shape[1] == 64
    assert _rel(f1, f3) < 1e-10 and _relmax(d1, d3) < 1e-8
    assert _rel(f1, f2) < 1e-10 and abs(it1 - it2) <= 3


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cw_64col_against_oracle_and_gather(precision, monkeypatch):
    """Real solves of 64 columns (the Jacobian box's forward): plain float64 (512-byte rows), and
    mixed with float32 inner solves (256-byte rows, four to a copy, on the same plan)."""
    from oracle import rb_oracle as orc
    from redbirdpy_amd import workloads

    cfg = workloads.jac_box(2.0)
    (d1, f1, it1), (d2, f2, it2) = _forward_both(cfg, precision, monkeypatch)
    c3, _ = orc.meshprep(copy.deepcopy(cfg))
    d3, f3 = orc.runforward(c3)
    assert _rel(f1, f3) < 1e-10 and _relmax(d1, d3) < 1e-9
    assert _rel(f1, f2) < (1e-11 if precision == "fp64" else 1e-10) and abs(it1 - it2) <= 3


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_irregular_mesh_64col(precision, monkeypatch):
    """Rows of 5 to ~40 entries (several 8-record groups per row, blocks closed by the record
    limit) on a centroid-split box; complex for mixed, real for fp64."""
    from oracle import rb_oracle as orc
    from redbirdpy_amd.mesh import kuhn_box
    from conftest import box_cfg

    p1 = (16, 16, 8)
    node, elem = kuhn_box([0, 0, 0], p1, 2.0)
    cen = node[elem[:, :4] - 1].mean(axis=1)
    split = np.flatnonzero(np.sum((cen - [8.0, 8.0, 4.0]) ** 2, axis=1) < 12.0)
    keep = np.setdiff1d(np.arange(len(elem)), split)
    ids = node.shape[0] + 1 + np.arange(len(split))
    quads = []
    for (a, b, c, d), x in zip(elem[split, :4], ids):
        quads += [(a, b, c, x), (a, b, x, d), (a, x, c, d), (x, b, c, d)]
    xs, ys = np.meshgrid(np.linspace(2, 14, 8), np.linspace(2, 14, 4))
    src = np.column_stack([xs.ravel(), ys.ravel(), np.zeros(32)])
    omega = 2 * np.pi * 1e8 if precision == "mixed" else 0.0
    cfg = box_cfg(p1, 2.0, src=src, det=src + [1.0, 1.0, 0.0], omega=omega)
    cfg["node"] = np.vstack([node, cen[split]])
    cfg["elem"] = np.vstack([elem[keep, :4], np.array(quads)]).astype(int)
    cfg["seg"] = np.ones(len(cfg["elem"]), dtype=int)
    (d1, f1, it1), (d2, f2, it2) = _forward_both(cfg, precision, monkeypatch)
    c3, _ = orc.meshprep(copy.deepcopy(cfg))
    d3, f3 = orc.runforward(c3)
    assert _rel(f1, f3) < 1e-10 and _relmax(d1, d3) < 1e-8
    assert _rel(f1, f2) < 1e-10 and abs(it1 - it2) <= 3


@pytest.mark.parametrize("nw", [2, 4, 6])
def test_cw_mixed_wavelength_pairs_against_oracle_and_gather(nw, monkeypatch):
    """Mixed solves of nw real 64-column systems (one per wavelength, the Gauss-Newton forward's
    shape): one LDS launch per pair of systems (float32 pair records, lanes 0-31 the first system's
    columns, 32-63 the second's), then the gated float64 residual over all 64-column chunks.  Every
    wavelength's fields and detector values against the oracle and against the gather SpMM."""
    from oracle import rb_oracle as orc
    from redbirdpy_amd import _gpu, workloads
    import redbirdpy_amd as rb

    cfg = workloads.jac_box(2.0)
    cfg["prop"] = {str(650 + 30 * i): np.array([[0, 0, 1, 1], [0.006 + 0.004 * i, 0.8 + 0.1 * i, 0, 1.37]])
                   for i in range(nw)}
    c0, _ = rb.meshprep(copy.deepcopy(cfg))
    # one record set per pair, then one per pair for the pairs' A D^-1
    assert _gpu.lds_plan_sets(nw, 64, False, "mixed") == nw
    assert _gpu.mesh_for(c0).spmm_plan(nw).struct.recv_sets >= nw
    (d1, f1, it1), (d2, f2, it2) = _forward_both(cfg, "mixed", monkeypatch)
    c3, _ = orc.meshprep(copy.deepcopy(cfg))
    d3, f3 = orc.runforward(c3)
    assert sorted(d1) == sorted(d3)
    for wv in d3:
        assert f1[wv].shape[1] == 64
        assert _rel(f1[wv], f3[wv]) < 1e-10 and _relmax(d1[wv], d3[wv]) < 1e-8
        assert _rel(f1[wv], f2[wv]) < 1e-10
    assert abs(it1 - it2) <= 3
