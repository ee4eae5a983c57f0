"""The C-ABI library loads and exports every entry point include/rbhip.h declares (no GPU needed)."""

import os
import re
import subprocess

import pytest

from redbirdpy_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rbhip.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(rb_\w+)\s*\(", txt, flags=re.M)))


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_library_built():
    assert os.path.exists(_abi.LIB_PATH), "run __graft_entry__.build() first"


def test_every_declared_symbol_is_exported():
    exp = _exported()
    missing = [n for n in _declared() if n not in exp]
    assert not missing, missing


def test_binding_matches_header():
    assert set(_abi.exported_names()) == set(_declared())


def test_load_without_gpu_and_version():
    import re
    hdr = open(os.path.join(ROOT, "include", "rbhip.h")).read()
    want = int(re.search(r"#define RB_ABI_VERSION (\d+)", hdr).group(1))
    L = _abi.load()
    assert L.rb_version() == want
    assert isinstance(L.rb_last_error(), bytes)


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "redbirdpy_amd")
    for dp, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dp, f)).read()
                assert "oracle" not in re.findall(r"^\s*(?:from|import)\s+(\w+)", src, flags=re.M), f


def test_compute_entry_points_need_a_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(_abi.RbError):
        _abi.require_gpu()


def test_jt_products_reject_bad_shapes_before_any_launch():
    """rb_jt_cols mode 0 reads z[k][j / np] for every column j < M: M must be nw blocks of np, and
    np >= 1 (checked on the host; the fake pointers below are never dereferenced)."""
    import ctypes

    L = _abi.load()
    fake = ctypes.c_void_p(256)
    for K, M, nw, npr, mode in ((10, 8, 2, 0, 0), (10, 8, 1, 4, 0), (10, 8, 3, 4, 0), (10, 8, 2, 4, 2)):
        rc = L.rb_jt_cols(K, M, fake, M, _abi.RB_F64, nw, npr, fake, mode, fake, fake, 1 << 20, None)
        assert rc == -1, (K, M, nw, npr, mode)
    assert L.rb_jt_rows(10, 8, fake, 8, _abi.RB_F64, 2, 0, fake, fake, None) == -1
    assert L.rb_jt_rows(10, 8, fake, 4, _abi.RB_F64, 2, 4, fake, fake, None) == -1  # ldj < M
