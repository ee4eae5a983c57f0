"""The degree-1 polynomial preconditioner (rb_krylov_set_precond(RB_PRECOND_POLY)) in every solve
path it serves: mixed FD through the LDS-staged SpMM, float64 FD through the column sweep, float64
CW through the LDS SpMM, a mixed real single system, two complex systems in one solve, two and four
real systems in chunk pairs (LDS pair product, mixed), an LDS plan without the A D^-1 records, and the
row-per-group SpMM of a solve past the sweep's descriptor limit.  Each against the oracle's SuperLU
solves at the same tolerances as the Jacobi solves, and with fewer iterations than Jacobi."""

import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _relmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _both(cfg, precision, monkeypatch, tol=1e-12):
    import redbirdpy_amd as rb
    from redbirdpy_amd import tune

    monkeypatch.setenv("RB_TUNE_CACHE", "/nonexistent/dir/tune.json")  # no cached choices apply
    out = {}
    for pc in (0, 1, 2, 3):
        c1, _ = rb.meshprep(copy.deepcopy(cfg))
        with tune.applied({"precond": pc}):
            d, f = rb.runforward(c1, precision=precision, tol=tol)
        out[pc] = (d, f, rb.forward.last_info.iters)
    return out


def _oracle(cfg):
    from oracle import rb_oracle as orc

    c2, _ = orc.meshprep(copy.deepcopy(cfg))
    return orc.runforward(c2)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_poly_fd_64col(precision, monkeypatch):
    from redbirdpy_amd import workloads

    cfg = workloads.fd_box(2.0)
    out = _both(cfg, precision, monkeypatch)
    d3, f3 = _oracle(cfg)
    for pc in (0, 1, 2, 3):
        d, f, _ = out[pc]
        assert _rel(f, f3) < 1e-10 and _relmax(d, d3) < 1e-8
    assert out[1][2] < 0.8 * out[0][2]
    if precision == "mixed":  # complex64: the fused LDS product, where degrees 2 and 3 apply
        assert out[2][2] < 0.85 * out[1][2] and out[3][2] < out[2][2]
    else:  # complex128 runs the sweep: degree 1
        assert out[3][2] == out[2][2] == out[1][2]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_poly_cw_64col(precision, monkeypatch):
    from redbirdpy_amd import workloads

    cfg = workloads.jac_box(2.0)
    out = _both(cfg, precision, monkeypatch)
    d3, f3 = _oracle(cfg)
    for pc in (1, 2, 3):
        d, f, _ = out[pc]
        assert _rel(f, f3) < 1e-10 and _relmax(d, d3) < 1e-8
    assert out[1][2] < 0.8 * out[0][2]
    # real float64 and float32: the fused LDS product
    assert out[2][2] < 0.85 * out[1][2] and out[3][2] < out[2][2]


def test_poly_two_complex_systems_one_solve(monkeypatch):
    from redbirdpy_amd import workloads

    cfg = workloads.fd_box(2.0)
    cfg["prop"] = {"690": np.array([[0, 0, 1, 1], [0.008, 1.0, 0, 1.37], [0.02, 1.2, 0, 1.37]]),
                   "830": np.array([[0, 0, 1, 1], [0.012, 0.9, 0, 1.37], [0.03, 1.1, 0, 1.37]])}
    out = _both(cfg, "mixed", monkeypatch)
    d3, f3 = _oracle(cfg)
    d, f, _ = out[1]
    for wv in d3:
        assert _rel(f[wv], f3[wv]) < 1e-10 and _relmax(d[wv], d3[wv]) < 1e-8


@pytest.mark.parametrize("nw", [2, 4])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_poly_real_64_column_systems_in_pairs(precision, nw, monkeypatch):
    """Real 64-column systems (the Gauss-Newton forward's shape) run in chunk pairs under Jacobi.
    Under the polynomial preconditioner the mixed solve keeps the pairs with the LDS-staged pair
    product (the pairs' A D^-1 records, z for both systems of a pair in its epilogue), checked
    also against the one-column sweep (LDS SpMM off); float64 runs them in the one-column sweep."""
    from redbirdpy_amd import workloads

    cfg = workloads.jac_box(2.0)
    cfg["prop"] = {str(690 + 40 * i): np.array([[0, 0, 1, 1], [0.008 + 0.003 * i, 1.0 - 0.05 * i, 0, 1.37],
                                                [0.02, 1.2, 0, 1.37]]) for i in range(nw)}
    out = _both(cfg, precision, monkeypatch)
    d3, f3 = _oracle(cfg)
    for pc in (0, 1, 2, 3):
        d, f, _ = out[pc]
        for wv in d3:
            assert _rel(f[wv], f3[wv]) < 1e-10 and _relmax(d[wv], d3[wv]) < 1e-8
    assert out[1][2] < 0.8 * out[0][2]
    if precision == "mixed":  # degrees 2 and 3 on the pair product
        assert out[2][2] < 0.85 * out[1][2] and out[3][2] < out[2][2]
    else:
        assert out[3][2] == out[2][2] == out[1][2]
    if precision == "mixed":
        monkeypatch.setenv("RB_LDS_SPMM", "0")
        sweep = _both(cfg, precision, monkeypatch)
        d, f, it = sweep[1]
        for wv in d3:
            assert _rel(f[wv], out[1][1][wv]) < 1e-10
        assert abs(it - out[1][2]) <= 3


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_poly_lds_plan_without_the_scaled_records(precision, monkeypatch):
    """A plan with one record set (no A D^-1 records) runs the preconditioner's product through
    the LDS SpMM on y = D^-1 r and forms z in its own kernel; both forms converge to the same
    solution in the same number of iterations (give or take one)."""
    import torch
    import redbirdpy_amd as rb
    from redbirdpy_amd import _gpu, tune, workloads
    from redbirdpy_amd.forward import forward_system

    monkeypatch.setenv("RB_TUNE_CACHE", "/nonexistent/dir/tune.json")
    cfg = workloads.fd_box(2.0) if precision == "mixed" else workloads.jac_box(2.0)
    c1, _ = rb.meshprep(copy.deepcopy(cfg))
    mesh, sysm, B, k, *_ = forward_system(c1, (1,), precision)
    assert k == 64
    m = sysm.mesh
    two = m.spmm_plan(2)
    assert two is not None and two.struct.recv_sets >= 2
    one = _gpu._build_spmm_plan(m.row_ptr.cpu().numpy(), m.col_idx[: m.nnz].cpu().numpy(), m.dev)
    assert one.struct.recv_sets == 1
    res = {}
    with tune.applied({"precond": 1}):
        for name, plan in (("fused", two), ("split", one)):
            res[name] = _gpu.krylov_solve(m.row_ptr, m.col_idx, m.diag, m.nn, m.nnz, sysm.vals64, sysm.vals32, B, k,
                                          1, sysm.is_complex, precision, 1e-12, 4000, plan=plan)
    torch.cuda.synchronize()
    (xf, inf_f), (xs, inf_s) = res["fused"], res["split"]
    assert inf_f.flag == 0 and inf_s.flag == 0
    assert inf_f.relres < 1e-12 and inf_s.relres < 1e-12
    assert abs(inf_f.iters - inf_s.iters) <= 1
    assert (torch.linalg.norm(xf - xs) / torch.linalg.norm(xs)).item() < 1e-10


def test_poly_row_group_spmm(monkeypatch):
    """Past the sweep's descriptor limit (lowered for the test) the row-per-group SpMM serves both
    products of an iteration."""
    from redbirdpy_amd import workloads

    monkeypatch.setenv("RB_SPMM_FLAT_MAX_BYTES", "4096")
    monkeypatch.setenv("RB_LDS_SPMM", "0")
    cfg = workloads.fd_box(2.0)
    out = _both(cfg, "mixed", monkeypatch)
    d3, f3 = _oracle(cfg)
    d, f, _ = out[1]
    assert _rel(f, f3) < 1e-10 and _relmax(d, d3) < 1e-8


def test_precond_setting_round_trip():
    from redbirdpy_amd import _abi

    L = _abi.lib()
    assert L.rb_krylov_get_precond() == _abi.RB_PRECOND_JACOBI
    _abi.call("rb_krylov_set_precond", _abi.RB_PRECOND_POLY)
    assert L.rb_krylov_get_precond() == _abi.RB_PRECOND_POLY
    _abi.call("rb_krylov_set_precond", _abi.RB_PRECOND_JACOBI)
    with pytest.raises(_abi.RbError):
        _abi.call("rb_krylov_set_precond", 7)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("omega", [0.0, 2 * np.pi * 1e8])
def test_poly_femsolve_edge_columns_and_flags(omega, precision):
    """Under the polynomial preconditioner, through femsolve (column sweep and the separate z kernel:
    the columns are not 64 wide): a zero column stays zero, a warm start at the solution keeps it, and
    an iteration cap reports flag 1."""
    import redbirdpy_amd as rb
    from redbirdpy_amd import tune
    from oracle import rb_oracle as orc
    from conftest import box_cfg

    c2, sd2 = orc.meshprep(copy.deepcopy(box_cfg((16, 16, 8), 2.0, omega=omega)))
    A = orc.femlhs(c2, c2["deldotdel"])
    rhs, _, _, _ = orc.femrhs(c2, sd2)
    xr, _ = orc.femsolve(A, rhs)
    rhs_d = np.column_stack([rhs.toarray(), np.zeros(rhs.shape[0])])
    with tune.applied({"precond": 1}):
        x, flag = rb.femsolve(A, rhs_d, precision=precision)
        assert flag == 0 and np.all(x[:, -1] == 0)
        assert _rel(x[:, :-1], xr) < 1e-10
        x2, flag2 = rb.femsolve(A, rhs_d, precision=precision, x0=x)
        assert flag2 == 0 and _rel(x2[:, :-1], x[:, :-1]) < 1e-10 and np.all(x2[:, -1] == 0)
        _, flag3 = rb.femsolve(A, rhs_d, precision=precision, maxiter=3)
        assert flag3 == 1
