"""ctypes binding of librbhip.so (the C ABI declared in include/rbhip.h).

The library is built in-tree by ``__graft_entry__.build()``.  There is no
fallback: if the library or a GPU is missing, every compute entry point
raises.
"""

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librbhip.so")

RB_PROP_ELEM = 0
RB_PROP_NODE = 1
RB_PROP_VOLUME = 2
RB_PREC_FP64 = 0
RB_PREC_MIXED = 1
RB_PRECOND_JACOBI = 0
RB_PRECOND_POLY = 1
RB_PRECOND_POLY2 = 2
RB_PRECOND_POLY3 = 3
RB_F64, RB_C128, RB_F32, RB_C64 = 0, 1, 2, 3
RB_PROF_SPMM, RB_PROF_UPDATE, RB_PROF_PUPDATE, RB_PROF_RESID, RB_PROF_INIT, RB_PROF_POLY = 0, 1, 2, 3, 4, 5
RB_WF_SINK, RB_WF_FACES = 0, 1
RB_WF_UNIFORM, RB_WF_FOURIER, RB_WF_IMAGE = 0, 1, 2
RB_WF_FRAME = 34
# LDS-staged SpMM plan padding (include/rbhip.h)
RB_LDS_UPAD, RB_LDS_RPAD, RB_LDS_RECPAD = 256, 512, 1024
RB_LDS_LIMITS = (128, 768)  # (RB_LDS_UMAX, RB_LDS_RECMAX)


class SpmmPlan(ctypes.Structure):
    """rb_spmm_plan: device pointers of an LDS-staged SpMM plan (include/rbhip.h)."""

    _fields_ = [("nblk", ctypes.c_int64), ("nrec", ctypes.c_int64), ("umax", ctypes.c_int32),
                ("recmax", ctypes.c_int32), ("brow", ctypes.c_void_p),
                ("bu", ctypes.c_void_p), ("ulist", ctypes.c_void_p), ("rinfo", ctypes.c_void_p),
                ("rmap", ctypes.c_void_p), ("loff", ctypes.c_void_p), ("recv", ctypes.c_void_p),
                ("recv_sets", ctypes.c_int64)]

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t

# name -> argtypes (every function returns int unless listed in _RESTYPE)
_SIGS = {
    "rb_version": [],
    "rb_last_error": [],
    "rb_device_count": [ctypes.POINTER(_int)],
    "rb_zero": [_vp, _sz, _vp],
    "rb_incidence": [_vp, _i64, _int, _i64, _vp, _vp, _vp],
    "rb_csr_pattern": [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _i64, ctypes.POINTER(_i64), _vp, _vp],
    "rb_tet_prep": [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "rb_boundary_faces": [_i64, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_i64), _vp],
    "rb_face_area": [_i64, _vp, _vp, _vp, _vp],
    "rb_node_volume": [_i64, _vp, _vp, _vp, _vp, _vp],
    "rb_slot_map": [_vp, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    "rb_assemble": [_vp, _vp, _i64, _vp, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _int, _int, _vp, _vp, _vp,
                    _i64, _i64, _vp, _vp, _int, _vp, _vp, _vp],
    "rb_assemble_workspace": [_i64, _i64, _int, ctypes.POINTER(_sz)],
    "rb_assemble_ws": [_vp, _vp, _i64, _vp, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _int, _int, _vp, _vp, _vp,
                       _i64, _i64, _vp, _vp, _int, _vp, _vp, _vp, _vp, _sz, _vp],
    "rb_assemble_bin_count": [_i64, _i64, ctypes.POINTER(_i64)],
    "rb_assemble_bins": [_vp, _i64, _i64, _i64, _vp, _vp],
    "rb_tsearch": [_vp, _vp, _i64, _vp, _int, _dbl, _vp, _vp, _vp],
    "rb_rhs_points": [_vp, _vp, _vp, _int, _vp, _i64, _int, _vp, _vp],
    "rb_widefield": [_i64, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _int, _vp, _int, _int, _int, _int,
                     _vp, _int, _int, _vp, _vp, _vp],
    "rb_krylov_workspace": [_i64, _int, _int, _int, _int, ctypes.POINTER(_sz)],
    "rb_krylov_set_sweep_grid": [_i64],
    "rb_krylov_get_sweep_grid": [],
    "rb_krylov_set_precond": [_int],
    "rb_krylov_get_precond": [],
    "rb_krylov_solve": [_i64, _i64, _vp, _vp, _vp, _int, _int, _int, _vp, _vp, _vp, _vp, _int, _dbl, _int,
                        _int, _int, _vp, _sz, ctypes.POINTER(_int), ctypes.POINTER(_dbl), ctypes.POINTER(_int),
                        _vp],
    "rb_krylov_solve_noalloc": [_i64, _i64, _vp, _vp, _vp, _int, _int, _int, _vp, _vp, _vp, _vp, _int, _dbl, _int,
                                _int, _int, ctypes.POINTER(_int), ctypes.POINTER(_dbl), ctypes.POINTER(_int), _vp],
    "rb_spmm": [_i64, _i64, _vp, _vp, _int, _int, _int, _vp, _vp, _vp, _vp],
    "rb_spmm_plan_count": [_i64, _vp, _vp, _int, _int, ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                           ctypes.POINTER(_i64)],
    "rb_spmm_plan_fill": [_i64, _vp, _vp, _int, _int, _vp, _vp, _vp, _vp, _vp, _vp],
    "rb_krylov_solve_planned": [_i64, _i64, _vp, _vp, _vp, _int, _int, _int, _vp, _vp, _vp, _vp, _int, _dbl, _int,
                                _int, _int, _vp, _sz, _vp, ctypes.POINTER(_int), ctypes.POINTER(_dbl),
                                ctypes.POINTER(_int), _vp],
    "rb_bicgstab_workspace": [_i64, _int, _int, ctypes.POINTER(_sz)],
    "rb_bicgstab_solve": [_i64, _i64, _vp, _vp, _vp, _int, _int, _vp, _vp, _vp, _dbl, _int, _int, _int, _vp, _sz,
                          ctypes.POINTER(_int), ctypes.POINTER(_dbl), ctypes.POINTER(_int), _vp],
    "rb_getdet": [_int, _vp, _vp, _vp, _vp, _i64, _int, _int, _int, _vp, _vp],
    "rb_jac_grid": [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp, _i64, _int, _int, _int, _vp, _vp,
                    _vp, _vp, _i64, _int, _vp, _vp, _vp, _vp],
    "rb_jac_patch_q": [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp, _vp, _vp],
    "rb_jac_contract_grid": [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _int, _int, _int, _vp, _vp, _vp, _vp,
                             _vp, _vp, _vp],
    "rb_gram": [_i64, _i64, _vp, _i64, _int, _vp, _int, _vp, _vp],
    "rb_gram_workspace": [_i64, _i64, _int, ctypes.POINTER(_sz)],
    "rb_gram_mfma": [_i64, _i64, _vp, _i64, _int, _vp, _int, _vp, _vp, _vp, _sz, _vp],
    "rb_jt_rows": [_i64, _i64, _vp, _i64, _int, _int, _i64, _vp, _vp, _vp],
    "rb_jt_remap": [_i64, _vp, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp],
    "rb_jt_cols_workspace": [_i64, _i64, ctypes.POINTER(_sz)],
    "rb_jt_cols": [_i64, _i64, _vp, _i64, _int, _int, _i64, _vp, _int, _vp, _vp, _sz, _vp],
    "rb_jt_hv_workspace": [_i64, _i64, ctypes.POINTER(_sz)],
    "rb_potrf_workspace": [_i64, ctypes.POINTER(_sz)],
    "rb_potrf": [_i64, _vp, _i64, _vp, _sz, _vp, _vp],
    "rb_jt_hv": [_i64, _i64, _vp, _i64, _int, _i64, _vp, _vp, _dbl, _dbl, _vp, _vp, _sz, _vp],
    "rb_prof_enable": [_int],
    "rb_prof_reset": [],
    "rb_prof_read": [_int, ctypes.POINTER(_dbl), ctypes.POINTER(_i64)],
}
_RESTYPE = {"rb_last_error": ctypes.c_char_p, "rb_krylov_get_sweep_grid": _i64}

_lib = None


def exported_names():
    """Names of every entry point this binding expects the library to export."""
    return sorted(_SIGS)


def load(path=LIB_PATH):
    """Load the library and bind signatures (no GPU needed)."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            "librbhip.so is not built (%s); run __graft_entry__.build() first" % path
        )
    L = ctypes.CDLL(path)
    for name, args in _SIGS.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _RESTYPE.get(name, _int)
    if path == LIB_PATH:
        _lib = L
    return L


def lib():
    return load()


class RbError(RuntimeError):
    pass


def check(rc, name):
    if rc != 0:
        msg = lib().rb_last_error()
        raise RbError("%s failed (%d): %s" % (name, rc, msg.decode() if msg else ""))


def call(name, *args):
    """Call an entry point and raise RbError on a nonzero status."""
    rc = getattr(lib(), name)(*args)
    check(rc, name)
    return rc


def ptr(t):
    """Device pointer of a tensor (None -> NULL).

    A tensor with torch's lazy conjugate or negative bit set is refused: its
    memory holds the unconjugated (un-negated) values the kernel would read.
    """
    if t is None:
        return None
    if t.is_conj() or t.is_neg():
        raise ValueError("pass a tensor whose conjugation is resolved (resolve_conj() / resolve_neg())")
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


_checked = False


def require_gpu():
    """Raise unless a HIP device is visible and the library is loadable."""
    global _checked
    if _checked:
        return
    lib()
    if not torch.cuda.is_available():
        raise RbError("redbirdpy_amd needs a HIP GPU (torch.cuda.is_available() is False)")
    _checked = True
