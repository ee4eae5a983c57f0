"""rb_assemble_ws (per-element property records, codes binned per workgroup with their
contributions staged in LDS) against rb_assemble (one thread per slot, per-code gathers) on the
same device inputs: label, element and node properties (node tables with and without a
refractive-index column), CW and FD, one to nine systems (the slot pass takes the systems in
groups of four).  Both sum each slot's codes in the slot map's order; the binned form splits
the P1 mass weights into a per-code and a per-slot term, so values agree to rounding."""

import numpy as np
import pytest

from conftest import box_cfg

pytestmark = pytest.mark.gpu


def _cfg(branch, omega, nw):
    cfg = box_cfg((12, 10, 6), 2.0, omega=omega)
    ne, nn = cfg["elem"].shape[0], cfg["node"].shape[0]

    def table(i):
        rng = np.random.default_rng(11 + i)
        if branch == "elem":
            return np.column_stack([rng.uniform(0.005, 0.03, ne), rng.uniform(0.5, 1.5, ne), np.zeros(ne),
                                    rng.uniform(1.3, 1.45, ne)])
        if branch == "node":
            return np.column_stack([rng.uniform(0.005, 0.03, nn), rng.uniform(0.5, 1.5, nn), np.zeros(nn),
                                    rng.uniform(1.3, 1.45, nn)])
        if branch == "node_n1":  # no n column: one refractive index per system
            return np.column_stack([rng.uniform(0.005, 0.03, nn), rng.uniform(0.5, 1.5, nn),
                                    rng.uniform(0.0, 0.2, nn)])
        return np.array([[0, 0, 1, 1], [0.01 + 0.002 * i, 1.0, 0, 1.37], [0.02, 1.2 + 0.05 * i, 0.1, 1.37]])

    if branch == "label":
        lab = np.ones(ne, dtype=int)
        lab[np.random.default_rng(5).random(ne) < 0.3] = 2
        cfg["seg"] = lab
    elif branch.startswith("node"):
        cfg.pop("seg")
    cfg["prop"] = table(0) if nw == 1 else {str(600 + 10 * i): table(i) for i in range(nw)}
    return cfg


@pytest.mark.parametrize("nw", [1, 4, 9])
@pytest.mark.parametrize("omega", [0.0, 2 * np.pi * 1e8])
@pytest.mark.parametrize("branch", ["label", "elem", "node", "node_n1"])
def test_table_assembly_matches_per_code_assembly(branch, omega, nw):
    import torch
    import redbirdpy_amd as rb
    from redbirdpy_amd import _gpu

    if branch == "node_n1":
        # meshprep reads Reff from an n column; the 3-column tables replace the props after it
        c, _ = rb.meshprep(_cfg("node", omega, nw))
        c3 = _cfg(branch, omega, nw)
        c["prop"] = c3["prop"]
    else:
        c, _ = rb.meshprep(_cfg(branch, omega, nw))
    mesh = _gpu.mesh_for(c)
    wvs = list(c["prop"].keys()) if isinstance(c["prop"], dict) else [""]
    a = _gpu.assemble_inputs(mesh, c, c["deldotdel"], [(wv, 1) for wv in wvs])
    assert a.nsys == nw
    assert a.nlen == (1 if branch in ("label", "node_n1") else a.plen)
    v64, v32 = _gpu.assemble_call(mesh, a, want32=True)
    r64, _ = _gpu.assemble_call(mesh, a, entry="rb_assemble")
    torch.cuda.synchronize()
    assert v64.dtype == r64.dtype and v64.is_complex() == (omega > 0)
    for s in range(nw):
        scale = torch.max(torch.abs(r64[s])).item()
        assert scale > 0
        assert torch.max(torch.abs(v64[s] - r64[s])).item() <= 2e-15 * scale
    # the 32-bit copy is the rounding of the 64-bit values
    assert torch.equal(v32, v64.to(v32.dtype))
    # a second assembly into the kept workspace gives the same bits
    w64, _ = _gpu.assemble_call(mesh, a)
    assert torch.equal(w64, v64)


def test_table_assembly_refuses_a_short_workspace():
    import torch
    import redbirdpy_amd as rb
    from redbirdpy_amd import _gpu, _abi

    c, _ = rb.meshprep(_cfg("label", 0.0, 1))
    mesh = _gpu.mesh_for(c)
    a = _gpu.assemble_inputs(mesh, c, c["deldotdel"], [("", 1)])
    v64 = torch.empty((1, mesh.nnz), dtype=torch.float64, device=mesh.dev)
    work = torch.empty(64, dtype=torch.uint8, device=mesh.dev)
    P = _abi.ptr
    with pytest.raises(_abi.RbError, match="workspace"):
        _abi.call("rb_assemble_ws", P(mesh.slot_ptr), P(mesh.codes), mesh.nnz, P(mesh.el), mesh.ne, P(mesh.fc),
                  mesh.nf, mesh.nn, P(mesh.dd), P(mesh.evol), P(mesh.area), 1, a.kind, P(a.mua), P(a.musp),
                  P(a.nref), a.plen, a.nlen, P(a.om), P(a.bc), 0, P(v64), None, P(mesh.asm_bins()), P(work),
                  work.numel(),
                  _abi.stream_ptr())
