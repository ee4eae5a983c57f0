"""Linear solves (drop-in for redbirdpy.solver.femsolve on the hot path).

Every ``method`` of the reference (solver.py:385-948) maps onto a GPU
multi-RHS Krylov solver, all columns in one launch sequence: Jacobi CG for
real symmetric matrices and COCG (unconjugated) for complex symmetric ones,
and Jacobi BiCGSTAB for 'bicgstab', 'gmres' and 'qmr' (the reference's
general-matrix methods), for any matrix that is not symmetric, and to
finish a CG/COCG column that breaks down.  Direct methods ('auto',
'direct', 'superlu', 'pardiso', 'umfpack', 'cholmod') are emulated with a
tight relative-residual tolerance of 1e-12 (override with ``tol``);
iterative methods use the reference default tol=1e-10 and maxiter=1000.
Extra keyword ``precision='mixed'`` runs the CG/COCG inner iterations in
float32/complex64 with float64 residual refinement.
"""

__all__ = ["femsolve", "solverinfo"]

import numpy as np
import torch
from scipy import sparse

from . import _gpu

_DIRECT = {"auto", "direct", "superlu", "pardiso", "umfpack", "cholmod"}
_ITER = {"cg", "cg+amg", "gmres", "bicgstab", "qmr", "minres", "minres+amg", "blqmr"}
_GENERAL = {"bicgstab", "gmres", "qmr"}  # methods the reference applies to general matrices

last_info = None  # SolveInfo of the most recent femsolve / runforward solve


def _diag_slots(indptr, indices, n):
    rows = np.repeat(np.arange(n), np.diff(indptr))
    hit = np.flatnonzero(indices == rows)
    d = np.full(n, -1, dtype=np.int64)
    d[rows[hit]] = hit
    return d


def femsolve(Amat, rhs, method="auto", **kwargs):
    """Solve A x = b for one or many right-hand sides; returns (x, flag).

    Zero RHS columns give zero solutions, as in the reference
    (solver.py:569-581).  A must be symmetric (complex symmetric for FD).
    """
    global last_info
    if method not in _DIRECT and method not in _ITER:
        raise ValueError("Unknown solver: %s" % method)
    if sparse.issparse(rhs):
        rhs = rhs.toarray()
    rhs = np.asarray(rhs)
    oned = rhs.ndim == 1
    if oned:
        rhs = rhs.reshape(-1, 1)
    n, ncol = rhs.shape
    A = sparse.csr_matrix(Amat)
    A.sum_duplicates()
    A.sort_indices()
    is_complex = np.iscomplexobj(A.data) or np.iscomplexobj(rhs)
    asym = abs(A - A.T)
    symmetric = not (asym.nnz and asym.max() > 1e-12 * (abs(A).max() if A.nnz else 1.0))
    use_bicgstab = method in _GENERAL or not symmetric
    diag_ok = np.all(np.diff(A.indptr) > 0) and np.all(_diag_slots(A.indptr, A.indices, n) >= 0)
    if not diag_ok:
        raise ValueError("femsolve: every row needs a stored diagonal entry (Jacobi preconditioning)")
    if method in _DIRECT:
        tol = kwargs.get("tol", 1e-12)
        tol = min(tol, 1e-12)
        maxiter = kwargs.get("maxiter", max(4000, 4 * int(np.sqrt(n)) + 1000))
    else:
        tol = kwargs.get("tol", 1e-10)
        maxiter = kwargs.get("maxiter", 1000)
    precision = kwargs.get("precision", "fp64")
    dev = _gpu.device()
    rdt = torch.complex128 if is_complex else torch.float64
    row_ptr = torch.as_tensor(A.indptr.astype(np.int32), device=dev)
    col_idx = torch.as_tensor(A.indices.astype(np.int32), device=dev)
    diag = torch.as_tensor(_diag_slots(A.indptr, A.indices, n).astype(np.int32), device=dev)
    vals = torch.as_tensor(A.data.astype(complex if is_complex else float), device=dev).reshape(1, -1)
    x = np.zeros((n, ncol), dtype=complex if is_complex else float)
    x0 = kwargs.get("x0", None)
    # chunks of at most 512 columns; the column sweep's buffer descriptor (float64 vectors under
    # 2 GiB) may ask for narrower ones on large meshes
    step = max(1, 512 if use_bicgstab else _gpu.sweep_columns(n, min(ncol, 512), 1, is_complex))
    # the LDS-staged SpMM serves 64-column chunks with 8-byte or 4-byte inner values (complex64 or
    # float32 mixed, float64 plain)
    plan = None
    if not use_bicgstab and (precision == "mixed" or not is_complex) and (step == 64 or ncol % step == 64):
        import os

        if os.environ.get("RB_LDS_SPMM", "1") != "0":
            plan = _gpu.spmm_plan(A.indptr, A.indices, dev)
            if plan is not None:  # A and A D^-1 (the polynomial preconditioner's product)
                plan.ensure_sets(_gpu.lds_plan_sets(1, 64, is_complex, precision))
    flag = 0
    info = None
    for c0 in range(0, ncol, step):
        c1 = min(ncol, c0 + step)
        B = torch.as_tensor(np.ascontiguousarray(rhs[:, c0:c1]), device=dev).to(rdt)
        X0 = None
        if x0 is not None:
            x0a = np.asarray(x0).reshape(n, -1)
            X0 = torch.as_tensor(np.ascontiguousarray(x0a[:, c0:c1] if x0a.shape[1] > 1 else x0a), device=dev)
            X0 = X0.to(rdt).expand(n, c1 - c0).contiguous()
        if use_bicgstab:
            X, info = _gpu.bicgstab_solve(row_ptr, col_idx, diag, n, A.nnz, vals, B, is_complex, tol, maxiter, X0)
        else:
            X, info = _gpu.krylov_solve(row_ptr, col_idx, diag, n, A.nnz, vals, None, B, c1 - c0, 1, is_complex,
                                        precision, tol, maxiter, X0, plan=plan if c1 - c0 == 64 else None)
            if info.flag == 4:  # CG/COCG breakdown: BiCGSTAB from the last iterate
                X, info = _gpu.bicgstab_solve(row_ptr, col_idx, diag, n, A.nnz, vals, B, is_complex, tol, maxiter,
                                              X)
        x[:, c0:c1] = X.cpu().numpy()
        flag = max(flag, info.flag)
    last_info = info
    return (x.ravel() if oned else x), flag


def solverinfo():
    """Solver capabilities (solver.py:951-974 analogue)."""
    return {
        "direct_solver": "gpu-krylov",
        "has_pardiso": False,
        "has_umfpack": False,
        "has_cholmod": False,
        "has_amg": False,
        "has_blqmr": False,
        "complex_direct": "gpu-cocg",
        "complex_iterative": ["cocg", "bicgstab"],
        "general_iterative": ["bicgstab"],
        "symmetric_iterative": ["cg"],
        "precision": ["fp64", "mixed"],
        "cpu_count": 1,
    }
