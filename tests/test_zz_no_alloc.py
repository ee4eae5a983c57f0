"""rb_krylov_solve_noalloc (fixed per-device scratch, no allocation) against rb_krylov_solve.

Both entry points run the same kernels, so the solutions, iteration counts,
residuals and flags must agree bit for bit.  Two calls in a row per case, the
second on a smaller system, so the scratch holds the previous call's contents.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(p1, h, omega, nsys_wv=1):
    from conftest import box_cfg
    import redbirdpy_amd as rb
    from redbirdpy_amd import _gpu

    cfg = box_cfg(p1=p1, h=h, omega=omega)
    if nsys_wv > 1:
        cfg["prop"] = {str(700 + 50 * i): np.array([[0, 0, 1, 1], [0.01 + 0.002 * i, 1.0, 0, 1.37]])
                       for i in range(nsys_wv)}
    cfg, _ = rb.meshprep(cfg)
    mesh = _gpu.mesh_for(cfg)
    wvs = list(cfg["prop"].keys()) if isinstance(cfg["prop"], dict) else [""]
    sysm = _gpu.assemble(mesh, cfg, cfg["deldotdel"], [(wv, 1) for wv in wvs], want32=True)
    return mesh, sysm


def _solve(entry, mesh, sysm, B, k, precision, tol=1e-10, maxiter=3000):
    import torch
    from redbirdpy_amd import _abi
    from redbirdpy_amd._abi import ptr

    X = torch.empty_like(B)
    it, rel, fl = ctypes.c_int(0), ctypes.c_double(0), ctypes.c_int(0)
    args = [mesh.nn, mesh.nnz, ptr(mesh.row_ptr), ptr(mesh.col_idx), ptr(mesh.diag), sysm.nsys, k,
            int(sysm.is_complex), ptr(sysm.vals64), ptr(sysm.vals32), ptr(B), ptr(X), precision, tol, maxiter, 0, 16]
    if entry == "rb_krylov_solve":
        nb = ctypes.c_size_t(0)
        _abi.call("rb_krylov_workspace", mesh.nn, sysm.nsys, k, int(sysm.is_complex), precision, ctypes.byref(nb))
        work = torch.empty(int(nb.value), dtype=torch.uint8, device=B.device)
        args += [ptr(work), nb.value]
    args += [ctypes.byref(it), ctypes.byref(rel), ctypes.byref(fl), _abi.stream_ptr()]
    _abi.call(entry, *args)
    torch.cuda.synchronize()
    return X, it.value, rel.value, fl.value


def _rhs(mesh, C, is_complex, seed):
    import torch

    g = torch.Generator(device="cuda").manual_seed(seed)
    B = torch.randn((mesh.nn, C), generator=g, device="cuda", dtype=torch.float64)
    if is_complex:
        B = torch.complex(B, torch.randn((mesh.nn, C), generator=g, device="cuda", dtype=torch.float64))
    return B.contiguous()


@pytest.mark.parametrize("omega,precision", [(2 * np.pi * 1e8, 1), (0.0, 0), (0.0, 1)])
def test_noalloc_matches_workspace_entry(omega, precision):
    import torch

    # (box, spacing, wavelengths, columns per system): the second call is smaller
    calls = [((30, 30, 15), 1.0, 1, 64), ((20, 20, 10), 2.0, 1, 24)]
    if omega == 0 and precision == 1:
        calls = [((30, 30, 15), 1.5, 4, 64), ((20, 20, 10), 1.0, 2, 64)]
    for case, (p1, h, nw, k) in enumerate(calls):
        mesh, sysm = _system(p1, h, omega, nw)
        B = _rhs(mesh, sysm.nsys * k, sysm.is_complex, 7 + case)
        print("noalloc case omega=%g precision=%d call %d: nn=%d C=%d" % (omega, precision, case, mesh.nn,
                                                                          sysm.nsys * k), flush=True)
        Xa, ita, rela, fla = _solve("rb_krylov_solve_noalloc", mesh, sysm, B, k, precision)
        Xw, itw, relw, flw = _solve("rb_krylov_solve", mesh, sysm, B, k, precision)
        assert fla == flw == 0 and ita == itw and rela == relw
        assert torch.equal(Xa, Xw)
        assert rela <= 1e-10


def test_noalloc_rejects_oversize():
    import torch
    from redbirdpy_amd import _abi
    from redbirdpy_amd._abi import ptr

    # 200k nodes x 512 complex columns (mixed) needs far more than the fixed scratch: refused
    # on the host before any launch (the pointers are never dereferenced)
    t = torch.zeros(8, dtype=torch.float64, device="cuda")
    it, rel, fl = ctypes.c_int(0), ctypes.c_double(0), ctypes.c_int(0)
    rc = _abi.lib().rb_krylov_solve_noalloc(200000, 3000000, ptr(t), ptr(t), ptr(t), 1, 512, 1, ptr(t), ptr(t),
                                            ptr(t), ptr(t), 1, 1e-10, 100, 0, 16, ctypes.byref(it),
                                            ctypes.byref(rel), ctypes.byref(fl), _abi.stream_ptr())
    assert rc == -3
    assert b"scratch" in _abi.lib().rb_last_error()
