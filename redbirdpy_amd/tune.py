"""Launch-parameter autotuner for the multi-RHS Krylov solves.

The reference splits its right-hand sides for the CPU solvers with ``rhsblock``
and ``nthread`` (redbirdpy solver.py:600-639, ``_blqmr_parallel`` :233-307).
Here one multi-RHS solve covers every column, and what is left to choose is
how the GPU runs it:

* ``lds``: the LDS-staged SpMM (its plan exists for 64-column solves of one
  system with 8-byte or 4-byte inner values) or the die-wide column sweep;
* ``grid``: the column sweep's workgroup cap (0 = one resident round of the
  device, else a multiple of 8; rb_krylov_set_sweep_grid);
* ``check_every``: iterations between the host's convergence polls (at most;
  the solver polls sooner when the residual decay says the columns are
  about to converge);
* ``precond``: Jacobi, or the degree-1 polynomial preconditioner
  (rb_krylov_set_precond): one more SpMM per iteration for ~1.8x fewer
  iterations on 3-D FEM systems (real 64-column systems keep the chunk-pair
  layout with the LDS-staged pair SpMM, else run in the one-column sweep), or
  the degree-2 and degree-3 ones on the fused LDS products (two or three more
  products for ~2.8x or ~3.7x fewer iterations than Jacobi);

``tune_system`` times each candidate on the system itself with full solves to
the requested tolerance (columns converge at different iterations and the
kernels skip converged 64-column chunks, so a fixed-iteration timing with
every column active ranks the SpMM paths differently: at 115k nodes it put
the 4-system LDS pairs 1% ahead of the sweep, whose Gauss-Newton forward
solves are 6% faster), first the SpMM choices, then the poll interval for
the fastest, and keeps the fastest.  Results are cached per (device name, nodes, nonzeros, columns,
systems, real/complex, precision) in a JSON file (``RB_TUNE_CACHE``, default
``~/.cache/redbirdpy_amd/tune.json``); ``_gpu.solve_system`` applies a cached
entry to every later solve of that shape, and with ``RB_AUTOTUNE=1`` tunes a
shape on its first solve.  Every candidate runs the same arithmetic: the
choice changes the launch shape, not the result beyond rounding.
"""

import json
import os
import threading
import time

import torch

from . import _abi

_LOCK = threading.Lock()
_CACHE = None
_CACHE_FILE = None

GRID_CAPS = (0, 512, 256)
CHECK_EVERY = (8, 16, 32)
PRECONDS = (_abi.RB_PRECOND_JACOBI, _abi.RB_PRECOND_POLY, _abi.RB_PRECOND_POLY2, _abi.RB_PRECOND_POLY3)
PRECOND_NAMES = {_abi.RB_PRECOND_JACOBI: "jacobi", _abi.RB_PRECOND_POLY: "poly", _abi.RB_PRECOND_POLY2: "poly2",
                 _abi.RB_PRECOND_POLY3: "poly3"}
DEFAULT = {"lds": True, "grid": 0, "check_every": 16, "precond": _abi.RB_PRECOND_JACOBI}


def cache_path():
    return os.environ.get("RB_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "redbirdpy_amd",
                                                            "tune.json")


def _load():
    global _CACHE, _CACHE_FILE
    path = cache_path()
    if _CACHE is None or _CACHE_FILE != path:
        try:
            with open(path) as f:
                data = json.load(f)
            _CACHE = data if isinstance(data, dict) else {}
        except (OSError, ValueError):
            _CACHE = {}
        _CACHE_FILE = path
    return _CACHE


def _save():
    path = cache_path()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = "%s.%d.tmp" % (path, os.getpid())
    with open(tmp, "w") as f:
        json.dump(_CACHE, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def key(device_name, n, nnz, k, nsys, is_complex, precision):
    return "%s|n=%d|nnz=%d|k=%d|nsys=%d|%s|%s" % (device_name, n, nnz, k, nsys, "complex" if is_complex else "real",
                                                  precision)


_NAMES = {}


def _device_name(dev):
    i = dev.index if dev.index is not None else torch.cuda.current_device()
    if i not in _NAMES:
        _NAMES[i] = torch.cuda.get_device_name(i)
    return _NAMES[i]


def system_key(sysm, k, precision):
    m = sysm.mesh
    return key(_device_name(m.dev), m.nn, m.nnz, k, sysm.nsys, sysm.is_complex, precision)


def lookup(sysm, k, precision):
    """The cached launch choices for this system shape, or None."""
    with _LOCK:
        return _load().get(system_key(sysm, k, precision))


def store(sysm, k, precision, choice):
    with _LOCK:
        _load()[system_key(sysm, k, precision)] = dict(choice)
        _save()


def clear():
    """Forget every cached choice (the file too)."""
    global _CACHE
    with _LOCK:
        _CACHE = {}
        try:
            os.remove(cache_path())
        except OSError:
            pass


class applied:
    """Context that sets the column sweep's grid cap and the preconditioner that the choice names
    for the solves inside it (settings it leaves out, or a choice of None, stay as they are)."""

    def __init__(self, choice):
        choice = choice or {}
        self.grid = int(choice["grid"]) if "grid" in choice else None
        self.precond = int(choice["precond"]) if "precond" in choice else None

    def __enter__(self):
        L = _abi.lib()
        self.prev = (int(L.rb_krylov_get_sweep_grid()), int(L.rb_krylov_get_precond()))
        if self.grid is not None:
            _abi.call("rb_krylov_set_sweep_grid", self.grid)
        if self.precond is not None:
            _abi.call("rb_krylov_set_precond", self.precond)
        return self

    def __exit__(self, *exc):
        _abi.call("rb_krylov_set_sweep_grid", self.prev[0])
        _abi.call("rb_krylov_set_precond", self.prev[1])
        return False


def _timed(fn, reps):
    fn()  # warm (workspace, plan, first-launch costs)
    torch.cuda.synchronize()
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best


def tune_system(sysm, B, k, precision="mixed", tol=1e-10, maxiter=2000, reps=2, save=True, X0=None):
    """Measure the launch choices for solving sysm (nsys systems x k columns of B) to tol and cache
    the fastest.  X0: the warm start the workload's solves begin from (a Gauss-Newton forward starts
    from the previous iteration's fields; warm and cold solves can rank the SpMM paths differently).
    Returns {"lds", "grid", "check_every", "precond", "seconds_per_solve", "candidates": [...]}."""
    from . import _gpu

    m = sysm.mesh
    sets = _gpu.lds_plan_sets(sysm.nsys, k, sysm.is_complex, precision)
    plan_ok = sets > 0 and m.spmm_plan(sets) is not None
    flat = _gpu.sweep_columns(m.nn, k, sysm.nsys, sysm.is_complex) == k
    spmms = []
    if plan_ok:
        spmms.append({"lds": True, "grid": 0})
    if flat or not plan_ok:
        spmms += [{"lds": False, "grid": g} for g in (GRID_CAPS if flat else (0,))]
    # real 64-column systems in pairs: the polynomial preconditioner keeps the pair layout with the
    # LDS-staged pair SpMM (mixed, a plan with the pairs' A D^-1 records), else it runs them in the
    # one-column sweep
    pairs = not sysm.is_complex and k == 64 and sysm.nsys % 2 == 0
    cands = [dict(c, precond=_abi.RB_PRECOND_JACOBI) for c in spmms]
    if pairs:
        cands += [{"lds": False, "grid": g, "precond": _abi.RB_PRECOND_POLY} for g in (GRID_CAPS if flat else (0,))]
        if plan_ok:
            cands += [{"lds": True, "grid": 0, "precond": pc}
                      for pc in (_abi.RB_PRECOND_POLY, _abi.RB_PRECOND_POLY2, _abi.RB_PRECOND_POLY3)]
    else:
        cands += [dict(c, precond=_abi.RB_PRECOND_POLY) for c in spmms]
        if plan_ok:  # degrees 2 and 3 exist on the fused LDS product only
            cands += [{"lds": True, "grid": 0, "precond": pc} for pc in (_abi.RB_PRECOND_POLY2, _abi.RB_PRECOND_POLY3)]
    def solve_time(ch):
        with applied(ch):
            return _timed(lambda: _gpu.solve_system(sysm, B, k, precision, tol, maxiter, X0=X0, tuned=ch), reps)

    res = [dict(c, check_every=16, seconds_per_solve=solve_time(dict(c, check_every=16))) for c in cands]
    best = min(res, key=lambda r: r["seconds_per_solve"])
    polls = [best]
    for ce in CHECK_EVERY:
        if ce != 16:
            ch = {"lds": best["lds"], "grid": best["grid"], "precond": best["precond"], "check_every": ce}
            polls.append(dict(ch, seconds_per_solve=solve_time(ch)))
    pbest = min(polls, key=lambda r: r["seconds_per_solve"])
    choice = {key: pbest[key] for key in ("lds", "grid", "check_every", "precond")}
    if save:
        store(sysm, k, precision, choice)
    return dict(choice, seconds_per_solve=pbest["seconds_per_solve"], candidates=res + polls[1:])


def tune_forward(cfg, rfcw=(1,), precision="mixed", tol=1e-12, maxiter=None, optodes=None, warm_from=None, **kw):
    """Tune the solves of cfg's forward run (forward_device's systems and right-hand sides): one
    entry for the whole multi-system solve, or, past 512 columns, for its one-system solves.
    warm_from: a cfg on the same mesh whose forward fields start each timed solve (as the next
    Gauss-Newton iteration starts from the last one's).  Returns tune_system's result."""
    import numpy as np
    from . import _gpu
    from .forward import forward_system

    from .forward import forward_device

    X0 = None
    if warm_from is not None:
        X0 = forward_device(warm_from, rfcw, "fp64", tol, maxiter, optodes=optodes, where="tune_forward").X
    mesh, sysm, B, k, systems, _, _ = forward_system(cfg, rfcw, precision, optodes)
    if maxiter is None:
        maxiter = max(4000, 4 * int(np.sqrt(mesh.nn)) + 1000)
    if len(systems) * k > 512:
        sysm = _gpu.System(mesh, sysm.vals64[:1], None if sysm.vals32 is None else sysm.vals32[:1], 1,
                           sysm.is_complex)
        B = B[:, :k].contiguous()
        X0 = None if X0 is None else X0[:, :k].contiguous()
    return tune_system(sysm, B, k, precision, tol, maxiter, X0=X0, **kw)
