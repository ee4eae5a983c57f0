"""redbirdpy_amd -- MI355X-native FEM forward / Jacobian / Gauss-Newton path.

A drop-in for the hot path of redbirdpy (fangq/redbirdpy v0.2.0): the same
function names, cfg/recon dicts and return values, with the compute in
hand-written HIP kernels for gfx950 (librbhip.so, C ABI in include/rbhip.h).
elem/face indices stay 1-based as in redbirdpy.

    import redbirdpy_amd as rb
    cfg, sd = rb.meshprep(cfg)
    detval, phi = rb.run(cfg)

rb.autograd.detval_mua(cfg, mua) gives the detector values as a torch
function of nodal mua whose backward is the adjoint Jacobian contraction.
"""

__version__ = "0.2.0"

from . import forward, recon, utility, property, solver, mesh, autograd
from .forward import *  # noqa: F401,F403
from .recon import *  # noqa: F401,F403
from .utility import *  # noqa: F401,F403
from .property import *  # noqa: F401,F403
from .solver import *  # noqa: F401,F403

__all__ = (
    forward.__all__
    + solver.__all__
    + recon.__all__
    + utility.__all__
    + property.__all__
    + ["run", "forward", "recon", "utility", "property", "solver", "mesh", "autograd"]
)


def run(cfg, recon_cfg=None, detphi0=None, sd=None, **kwargs):
    """Forward model when only cfg is given, reconstruction otherwise (__init__.py:83-112)."""
    if recon_cfg is None:
        return forward.runforward(cfg, **kwargs)
    if detphi0 is None:
        raise ValueError("detphi0 is required for reconstruction")
    return recon.runrecon(cfg, recon_cfg, detphi0, sd, **kwargs)
