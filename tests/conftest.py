import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built librbhip.so")


def _has_gpu():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason="no HIP GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


def box_cfg(p1=(20, 20, 10), h=2.0, prop=None, src=None, det=None, omega=0.0, labels=None):
    """Kuhn-box cfg with label properties (row 0 background, row 1.. tissues)."""
    from redbirdpy_amd.mesh import kuhn_box

    node, elem = kuhn_box([0, 0, 0], p1, h)
    cfg = {
        "node": node,
        "elem": elem,
        "prop": np.array([[0, 0, 1, 1], [0.01, 1.0, 0, 1.37]]) if prop is None else np.asarray(prop, float),
        "seg": np.ones(elem.shape[0], dtype=int) if labels is None else labels,
        "srcpos": np.array([[p1[0] / 2, p1[1] / 2, 0.0]]) if src is None else np.atleast_2d(src),
        "srcdir": np.array([[0.0, 0.0, 1.0]]),
        "detpos": np.array([[p1[0] / 2, p1[1] / 2 + 4, 0.0], [p1[0] / 2 + 4, p1[1] / 2, 0.0]]) if det is None
        else np.atleast_2d(det),
        "detdir": np.array([[0.0, 0.0, 1.0]]),
        "omega": omega,
    }
    return cfg


def grid_optodes(p1, n=3, z=0.0, margin=4.0):
    xs = np.linspace(margin, p1[0] - margin, n)
    ys = np.linspace(margin, p1[1] - margin, n)
    X, Y = np.meshgrid(xs, ys)
    return np.column_stack([X.ravel(), Y.ravel(), np.full(X.size, z)])
