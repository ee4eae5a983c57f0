"""Benchmark: multi-RHS frequency-domain forward solves on one MI355X.

Workload (one "step" = one pass of the forward hot path): the 60x60x30 mm
box at 1 mm spacing (115,351 nodes, 648,000 tets, CSR nnz 1,670,851) with
a spherical inclusion, omega = 2 pi 100 MHz, 32 sources + 32 detectors.
Each step assembles the complex FEM matrix on the GPU, fills the 64 RHS
columns, solves all 64 columns in one multi-RHS COCG solve and reads the
32x32 detector values.  The mesh, its CSR pattern and the located optodes
are resident in HBM before the timed region.

value   : RHS-solves/s of the mixed path (complex64 COCG with float64
          reliable updates), relative residual 1e-12 per column;
extra   : the float64 path (complex128 COCG to 1e-12), the error of the
          mixed solution against it, and (unless --no-gn) Gauss-Newton
          iterations/s of the 4-wavelength chromophore reconstruction in
          both precisions.
autotune: before its timed steps each workload's launch choices (SpMM path,
          sweep grid, poll interval) are tuned on it (tune.py); --no-autotune
          keeps the defaults.
roofline: the SpMM kernel of the mixed path (q = A p with the p.q
          reduction, LDS-staged: each block of rows copies its neighbour
          rows of p into LDS once).  achieved = the SpMM's algorithmic
          bytes per launch / mean launch duration, measured with start/stop
          hipEvents carried by each dispatch on its stream over a second run
          of the same steps (the events are kept out of the timed steps,
          where they cost ~9 %).  The float64 true residual of the column
          chunks whose reliable update the previous iteration requested runs
          in its own gated launch right after (reported as resid64).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SPMM_KERNEL = "k_cg_spmm_lds<cplx<float>,LdsCfg<128,768,2,8>,false,false>"


def _args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--h", type=float, default=1.0, help="mesh spacing in mm (1.0 = the 115,351-node mesh)")
    p.add_argument("--tol-mixed", type=float, default=1e-12)
    p.add_argument("--tol-fp64", type=float, default=1e-12)
    p.add_argument("--skip-cpu", action="store_true")
    p.add_argument("--skip-fp64", action="store_true")
    p.add_argument("--no-gn", dest="gn", action="store_false", help="skip the Gauss-Newton iterations leg")
    p.add_argument("--gn-iters", type=int, default=5)
    p.add_argument("--gn-tol", type=float, default=1e-10,
                   help="forward solver tolerance of both Gauss-Newton legs (redbirdpy femsolve's default)")
    p.add_argument("--gn-precisions", default="mixed,fp64", help="comma list of GN legs to run")
    p.add_argument("--gn-reftol", type=float, default=1e-8,
                   help="relative residual of the mixed leg's refined Gauss-Newton step (its float64 normal equations)")
    p.add_argument("--no-autotune", dest="autotune", action="store_false",
                   help="solve with the default launch choices instead of tuning each workload first")
    p.add_argument("--no-kernels", action="store_true", help="skip the Jacobian, Gram and assembly kernel legs")
    p.add_argument("--only-assembly", action="store_true", help="only the assembly leg (for kernel profiles)")
    p.add_argument("--no-batch", dest="batch", action="store_false", help="skip the 64-mesh uneven batch leg")
    p.add_argument("--batch-meshes", type=int, default=64)
    p.add_argument("--gn-phases", action="store_true", help="add a synchronising per-phase timing run")
    p.add_argument("--cpu-h", type=float, default=1.3, help="mesh spacing of the bounded CPU sample")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only copy 512 MiB three times (reference kernel for the FETCH_SIZE/WRITE_SIZE scale)")
    return p.parse_args()


def _spmm_bytes(nn, nnz, C, sv):
    """Algorithmic HBM bytes of one q = A p launch: CSR once, p read once, q written once."""
    return nnz * (sv + 4) + (nn + 1) * 4 + 2 * nn * C * sv


def cpu_threads():
    """Host threads of the CPU baseline: every core this process may run on, capped by OMP_NUM_THREADS
    when the host sets it (the GPU box gives one GPU's share of its cores)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def _blas_threads(n):
    """Context that runs numpy/scipy BLAS on n threads (numpy is imported before the environment could say)."""
    from threadpoolctl import threadpool_limits

    return threadpool_limits(limits=n)


def cpu_baseline(h):
    """Oracle (reference algorithm, scipy SuperLU) on the same FD box at spacing h."""
    from redbirdpy_amd import workloads
    from oracle import rb_oracle as orc

    n = cpu_threads()
    cfg = workloads.fd_box(h)
    cfg, sd = orc.meshprep(cfg)
    rhs, loc, bary, _ = orc.femrhs(cfg, sd)
    with _blas_threads(n):
        t0 = time.perf_counter()
        A = orc.femlhs(cfg, cfg["deldotdel"])
        x, _ = orc.femsolve(A, rhs)
        det = orc.femgetdet(x, cfg, rhs, loc, bary)
        dt = time.perf_counter() - t0
    k = rhs.shape[1]
    return {
        "value": k / dt,
        "unit": "RHS-solves/s",
        "cores": n,
        "kind": "port",
        "sample": "oracle femlhs + SuperLU factor/solve of %d complex RHS + femgetdet on the same FD box at %.1f mm "
                  "(%d nodes, not the 115k mesh), one pass, %.1f s; BLAS on %d threads, the SuperLU factorisation "
                  "itself is single-threaded" % (k, h, cfg["node"].shape[0], dt, n),
        "seconds": dt,
        "nodes": int(cfg["node"].shape[0]),
    }


def cpu_gn_baseline(h=2.0, jac_pairs=128):
    """One Gauss-Newton iteration of the oracle (the reference's default CPU path) on the 4-wavelength,
    3-chromophore chromophore box at spacing h: runforward (femlhs + SuperLU + femgetdet per wavelength),
    the numpy jac per wavelength, jacchrome/matflat and reginvunder.  The jac phase is timed on
    jac_pairs of each wavelength's pairs and scaled to all of them (it is a loop over pairs); every
    other phase runs in full."""
    import copy
    from redbirdpy_amd import workloads
    from oracle import rb_oracle as orc

    n = cpu_threads()
    cfg0, bg, truth, extin, E = workloads.chrome_box(h)
    c, sd = orc.meshprep(copy.deepcopy(cfg0))
    c["param"] = {k: v.copy() for k, v in truth.items()}
    c["prop"] = orc.updateprop(c, extin=extin)
    wvs = list(c["prop"].keys())
    with _blas_threads(n):
        t0 = time.perf_counter()
        det, phi = orc.runforward(c)
        t_fwd = time.perf_counter() - t0
        Jm = {}
        t_jac = 0.0
        npairs = 0
        for wv in wvs:
            sdw = sd[wv] if isinstance(sd, dict) else sd
            act = sdw[sdw[:, 2] == 1] if sdw.shape[1] >= 3 else sdw
            npairs = act.shape[0]
            sub = act[:jac_pairs]
            t1 = time.perf_counter()
            Jn, _ = orc.jac(sub, phi[wv], c["deldotdel"], c["elem"], c["evol"])
            t_jac += (time.perf_counter() - t1) * npairs / sub.shape[0]
            Jm[wv] = np.tile(Jn, (npairs // sub.shape[0] + 1, 1))[:npairs]  # full-size stand-in for the step
        chs = list(truth.keys())
        Emat = np.array([[extin[wv][ch] for ch in chs] for wv in wvs])
        t2 = time.perf_counter()
        Jc = orc.jacchrome(Jm, chs, extin=Emat)
        A = orc.matflat(Jc)
        y = orc.flatten_detphi(det, sd, wvs, [1])
        y0 = y * (1 + 0.005 * np.random.default_rng(40).standard_normal(y.shape))
        A, mis, _ = orc.matreform(A, y0, y, "real")
        sc = 1.0 / np.sqrt(np.sum(A ** 2))  # the blockscale of recon.py:235-236
        m = A.shape[0]
        dmu = orc.reginv(A * sc, mis, 1e-3 / m) * sc
        t_step = time.perf_counter() - t2
    t_iter = t_fwd + t_jac + t_step
    return {
        "value": 1.0 / t_iter,
        "unit": "GN iters/s",
        "cores": n,
        "kind": "port",
        "sample": "one oracle Gauss-Newton iteration, 4 wavelengths x %d pairs x 3 chromophores, on the chromophore "
                  "box at %.1f mm (%d nodes, not the 115k mesh): runforward %.1f s (SuperLU, single-threaded), numpy "
                  "jac %.1f s (timed on %d pairs per wavelength, scaled to %d), jacchrome + matflat + reginvunder "
                  "(m=%d, N=%d) %.1f s; BLAS on %d threads" % (
                      npairs, h, c["node"].shape[0], t_fwd, t_jac, min(jac_pairs, npairs), npairs, m, A.shape[1],
                      t_step, n),
        "seconds": t_iter,
        "nodes": int(c["node"].shape[0]),
    }


def cpu_jac_baseline(phi, sd, cfg, pairs=32):
    """The oracle's numpy jac (the reference's path without numba) on the 1024-pair CW box given the
    fields, timed on `pairs` of the pairs and scaled to the full Jacobian."""
    from oracle import rb_oracle as orc

    n = cpu_threads()
    act = sd[sd[:, 2] == 1] if sd.shape[1] >= 3 else sd
    sub = act[:pairs]
    with _blas_threads(n):
        t0 = time.perf_counter()
        orc.jac(sub, phi, cfg["deldotdel"], cfg["elem"], cfg["evol"])
        dt = time.perf_counter() - t0
    full = dt * act.shape[0] / sub.shape[0]
    return {
        "value": 1.0 / full,
        "unit": "Jacobians/s",
        "cores": n,
        "kind": "port",
        "sample": "oracle numpy jac (node and element Jacobians) for %d of the %d pairs on the %d-node CW box, "
                  "%.1f s, scaled to all pairs" % (sub.shape[0], act.shape[0], phi.shape[0], dt),
        "seconds_full_jacobian": full,
    }


def _tuned(cfg, precision, tol, warm_from=None):
    """Tune cfg's forward solves (tune.tune_forward, cached for the run) and report the choice."""
    from redbirdpy_amd import tune

    t0 = time.perf_counter()
    out = tune.tune_forward(cfg, precision=precision, tol=tol, warm_from=warm_from)
    return {"lds": out["lds"], "grid": out["grid"], "check_every": out["check_every"],
            "precond": tune.PRECOND_NAMES[out["precond"]], "tuning_s": time.perf_counter() - t0,
            "ms_per_solve": {"%s, %s, poll %d" % ("lds" if c["lds"] else "sweep grid %d" % c["grid"],
                                                  tune.PRECOND_NAMES[c["precond"]], c["check_every"]):
                             c["seconds_per_solve"] * 1e3 for c in out["candidates"]}}


def gn_bench(a):
    """Gauss-Newton iterations/s: 4 wavelengths x 3 chromophores, 1024 pairs per wavelength.

    Data: FP64 GPU forward of the true concentrations with 0.5% seeded
    Gaussian noise.  Each leg starts from the background and runs
    a.gn_iters iterations; iterations/s = iterations / wall time.  Both legs
    solve the forward systems to the same relative residual (a.gn_tol), so
    their ratio measures precision alone.
    """
    import copy
    import torch
    import redbirdpy_amd as rb
    from redbirdpy_amd import workloads, _gpu
    cfg0, bg, truth, extin, E = workloads.chrome_box(a.h)
    ct, sd = rb.meshprep(copy.deepcopy(cfg0))
    ct["param"] = {k: v.copy() for k, v in truth.items()}
    ct["prop"] = rb.updateprop(ct, extin=extin)
    det, _ = rb.runforward(ct)
    rng = np.random.default_rng(40)
    y0 = {wv: det[wv] * (1 + 0.005 * rng.standard_normal(det[wv].shape)) for wv in workloads.WAVELENGTHS}
    m = sum(v.size for v in y0.values())
    out = {"m": m, "nodes": int(ct["node"].shape[0]), "iters": a.gn_iters, "solver_tol": a.gn_tol,
           "mixed_step_reftol": a.gn_reftol}
    if a.autotune:
        # the Gauss-Newton forward solves start from the previous iteration's fields: tuned warm-started
        # from the fields at the starting (background) concentrations
        cb = {k: v for k, v in ct.items()}
        cb["param"] = {k: v.copy() for k, v in bg.items()}
        cb["prop"] = rb.updateprop(cb, extin=extin)
        out["autotune"] = {prec: _tuned(ct, prec, a.gn_tol, warm_from=cb) for prec in a.gn_precisions.split(",")}
    results = {}
    for prec in a.gn_precisions.split(","):
        # one untimed iteration first (workspace and mesh-structure allocation)
        c, sd = rb.meshprep(copy.deepcopy(cfg0))
        r = {"param": {k: v.copy() for k, v in bg.items()}}
        rb.runrecon(c, r, y0, sd, maxiter=1, lambda_=1e-3 / m, extin=extin, precision=prec, report=False,
                    solvertol=a.gn_tol, reftol=a.gn_reftol)
        c, sd = rb.meshprep(copy.deepcopy(cfg0))
        r = {"param": {k: v.copy() for k, v in bg.items()}}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, resid, *_ = rb.runrecon(c, r, y0, sd, maxiter=a.gn_iters, lambda_=1e-3 / m, extin=extin,
                                   precision=prec, report=False, solvertol=a.gn_tol, reftol=a.gn_reftol)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        # phase breakdown from a separate, synchronising run (not the timed one)
        tm = {}
        if a.gn_phases:
            c2, sd2 = rb.meshprep(copy.deepcopy(cfg0))
            r2 = {"param": {k: v.copy() for k, v in bg.items()}}
            rb.runrecon(c2, r2, y0, sd2, maxiter=a.gn_iters, lambda_=1e-3 / m, extin=extin, precision=prec,
                        report=False, timing=tm, solvertol=a.gn_tol, reftol=a.gn_reftol)
        results[prec] = r
        out[prec] = {"gn_iters_per_s": a.gn_iters / dt, "s_per_iter": dt / a.gn_iters,
                     "resid": [float(x) for x in resid], "phases_s": tm}
    if "mixed" in results and "fp64" in results:
        errs = {k: float(np.max(np.abs(results["mixed"]["param"][k] - results["fp64"]["param"][k]))
                         / np.max(np.abs(results["fp64"]["param"][k]))) for k in bg}
        out["mixed"]["param_maxrel_err_vs_fp64"] = errs
        out["speedup_mixed_vs_fp64"] = out["mixed"]["gn_iters_per_s"] / out["fp64"]["gn_iters_per_s"]
    if not a.skip_cpu:
        out["cpu_baseline"] = cpu_gn_baseline()
    return out


def batch_bench(a):
    """The uneven batch: 64 seeded Kuhn meshes (2,000-200,000 nodes each), point sources plus 16
    Fourier patterns and as many detectors per mesh; GPU mesh preparation for all meshes
    (meshprep_batch), then every mesh's forward solve in one block-diagonal multi-RHS solve
    (BatchForward), float64 and mixed, tolerance 1e-10.  RHS-solves/s counts each mesh's own
    source and detector columns; the timed solve includes assembly, the Krylov solve and the
    detector values (the batch's mesh, optodes and right-hand sides stay resident)."""
    import ctypes
    import torch
    import redbirdpy_amd as rb
    from redbirdpy_amd import workloads, _abi, tune
    from redbirdpy_amd.forward import BatchForward

    cfgs = workloads.batch_boxes(count=a.batch_meshes)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prepped = [c for c, _ in rb.meshprep_batch(cfgs)]
    torch.cuda.synchronize()
    t_prep = time.perf_counter() - t0
    t0 = time.perf_counter()
    bf = BatchForward(prepped)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    mesh = bf.mesh
    out = {"workload": "%d seeded Kuhn boxes, %d nodes / %d tets in all (%d-%d nodes each), %d columns per mesh "
                       "at most, %d per-mesh columns in all" % (len(cfgs), mesh.nn, mesh.ne,
                                                                min(c["node"].shape[0] for c in cfgs),
                                                                max(c["node"].shape[0] for c in cfgs), bf.C,
                                                                bf.columns),
           "meshprep_batch_s": t_prep, "setup_s": t_setup, "tol": 1e-10}
    L = _abi.lib()
    dets = {}
    if a.autotune:
        out["autotune"] = {}
    for prec in ("fp64", "mixed"):
        if a.autotune:  # the first 64-column group's launch choices (every full group shares them)
            t0 = time.perf_counter()
            tr = bf.tune(prec, 1e-10)
            out["autotune"][prec] = {"lds": tr["lds"], "grid": tr["grid"], "check_every": tr["check_every"],
                                     "precond": tune.PRECOND_NAMES[tr["precond"]],
                                     "tuning_s": time.perf_counter() - t0}
        bf.solve(prec, 1e-10, fields=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = bf.solve(prec, 1e-10, fields=False)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        dets[prec] = res
        L.rb_prof_reset()
        L.rb_prof_enable(1)
        bf.solve(prec, 1e-10, fields=False)
        torch.cuda.synchronize()
        L.rb_prof_enable(0)
        ms, cnt = ctypes.c_double(0), ctypes.c_int64(0)
        _abi.call("rb_prof_read", _abi.RB_PROF_SPMM, ctypes.byref(ms), ctypes.byref(cnt))
        sv = 8 if prec == "fp64" else 4
        # a solve wider than the sweep's descriptor runs as column groups: one SpMM per iteration
        # of each group, each launch priced at its own width
        groups = bf.info.groups or [(bf.C, bf.info.iters)]
        launches = sum(it for _, it in groups)
        total = sum(it * _spmm_bytes(mesh.nn, mesh.nnz, g, sv) for g, it in groups)
        rec = {"rhs_solves_per_s": bf.columns / dt, "s": dt, "iters": bf.info.iters, "relres": bf.info.relres,
               "column_groups": [g for g, _ in groups], "spmm_launches": cnt.value,
               "spmm_avg_us": ms.value / max(cnt.value, 1) * 1e3}
        if launches == cnt.value and ms.value > 0:
            rec["spmm_bytes_avg"] = total / launches
            rec["spmm_frac_hbm"] = total / (ms.value / 1e3) / 1e9 / HBM_PEAK_GBS
        out[prec] = rec
    out["speedup_mixed_vs_fp64"] = out["fp64"]["s"] / out["mixed"]["s"]
    pairs = [(x.detval, y.detval) for x, y in zip(dets["mixed"], dets["fp64"]) if y.detval is not None]
    # element-wise (dominated by the farthest detectors) and relative to each source's largest value
    out["mixed"]["err_detval_maxrel_vs_fp64"] = max(float(np.max(np.abs(x - y) / np.abs(y))) for x, y in pairs)
    out["mixed"]["err_detval_colmax_vs_fp64"] = max(
        float(np.max(np.abs(x - y) / np.max(np.abs(y), axis=0, keepdims=True))) for x, y in pairs)
    return out


def gram_bench(nn, m=4096, reps=3):
    """J J^T at the Gauss-Newton shape: K = nodes, m = 4 wavelengths x 1024 pairs.

    SYRK flops m(m+1)K; bf16x3 does three MFMA products per term.  Peaks:
    bf16 dense MFMA 2.5 PFLOP/s, fp64 78.6 TFLOP/s.
    """
    import torch
    from redbirdpy_amd import _gpu

    g = torch.Generator(device="cuda").manual_seed(0)
    JT = torch.randn((nn, m), generator=g, device="cuda", dtype=torch.float64)  # float64, as in the GN step
    JT *= torch.exp(torch.linspace(-6, 0, m, device="cuda", dtype=torch.float64))[None, :]
    flops = m * (m + 1) * nn
    out = {"K": nn, "m": m, "syrk_flop": flops}
    ref = None
    for mode, peak, terms in (("bf16x3", 2500.0, 3), ("bf16", 2500.0, 1), ("fp64", 78.6, 1)):
        _gpu.gram(JT, None, mode)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            G = _gpu.gram(JT, None, mode)
        e1.record()
        torch.cuda.synchronize()
        dt = e0.elapsed_time(e1) / 1e3 / reps
        tf = terms * flops / dt / 1e12
        rec = {"ms": dt * 1e3, "tflops": tf, "peak_tflops": peak, "frac": tf / peak}
        if mode == "fp64":
            ref = G
        out[mode] = rec
    for mode in ("bf16x3", "bf16"):
        G = _gpu.gram(JT, None, mode)
        d = torch.sqrt(torch.diagonal(ref))
        out[mode]["max_abs_err_normalised"] = (torch.abs(G - ref) / (d[:, None] * d[None, :])).max().item()
    return out


def assembly_bench(a, cfg, reps=20):
    """femlhs assembly at 115k nodes: the FD element-property system (1 complex system + its
    complex64 copy) and the 4-wavelength node-property CW systems of the Gauss-Newton leg (+ float32
    copies), through rb_assemble_ws (property pass + binned slot pass) and the per-slot rb_assemble,
    event-timed per call.  slot_pass_bytes: algorithmic bytes of the slot pass (codes, slot_ptr,
    deldotdel rows, property records, [evol, el, nodal mua], values written)."""
    import copy
    import torch
    import redbirdpy_amd as rb
    from redbirdpy_amd import _gpu, workloads

    def timeit(fn):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3

    out = {}
    cases = [("fd_elem_1sys", cfg, [("", 1)])]
    cfg0, bg, truth, extin, E = workloads.chrome_box(a.h)
    ct, _ = rb.meshprep(copy.deepcopy(cfg0))
    ct["param"] = {k: v.copy() for k, v in truth.items()}
    ct["prop"] = rb.updateprop(ct, extin=extin)
    cases.append(("cw_node_4sys", ct, [(wv, 1) for wv in ct["prop"]]))
    for name, c, systems in cases:
        mesh = _gpu.mesh_for(c)
        ai = _gpu.assemble_inputs(mesh, c, c["deldotdel"], systems)
        ncodes, ne, nnz, ns = mesh.codes.numel(), mesh.ne, mesh.nnz, ai.nsys
        vb = (16 + 8) if ai.is_complex else (8 + 4)
        if ai.kind == 0:  # element records of 32 bytes
            nbytes = 4 * ncodes + 4 * (nnz + 1) + 80 * ne + 32 * ne * ns + vb * nnz * ns
        else:  # node records of 16 bytes, evol, el and the nodal mua
            nbytes = (4 * ncodes + 4 * (nnz + 1) + 80 * ne + 16 * ne * ns + 8 * ne + 16 * ne + 8 * mesh.nn * ns
                      + vb * nnz * ns)
        t_ws = timeit(lambda: _gpu.assemble_call(mesh, ai, want32=True))
        t_old = timeit(lambda: _gpu.assemble_call(mesh, ai, want32=True, entry="rb_assemble"))
        out[name] = {"nodes": mesh.nn, "tets": ne, "nnz": nnz, "systems": ns, "us_per_call": t_ws,
                     "us_per_call_per_slot_kernel": t_old, "slot_pass_bytes": nbytes}
    return out


def jac_bench(a, reps=5):
    """Node Jacobian of 1024 pairs on the 1 mm CW box (115,351 nodes), float64 out.

    Algorithmic bytes nsd*nn*8 (output) + nn*k*8 (fields) + sum_n np_n^2 * 8 (the mesh's
    patch matrices, built once per mesh and read on every call).
    """
    import torch
    import redbirdpy_amd as rb
    from redbirdpy_amd import _gpu, workloads
    from redbirdpy_amd.forward import forward_device, _active_pairs

    cfg, sd = rb.meshprep(workloads.jac_box(a.h))
    res = forward_device(cfg, [1], "fp64", 1e-12)
    mesh = res.mesh
    s, d = _active_pairs(sd)
    phi = res.phi(0)
    JT = torch.empty((mesh.nn, len(s)), dtype=torch.float64, device=mesh.dev)
    _gpu.jac_node(mesh, phi, s, d, torch.float64, out=JT)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        _gpu.jac_node(mesh, phi, s, d, torch.float64, out=JT)
    e1.record()
    torch.cuda.synchronize()
    dt = e0.elapsed_time(e1) / 1e3 / reps
    qptr, _ = _gpu.patch_matrices(mesh)
    nbytes = len(s) * mesh.nn * 8 + mesh.nn * res.k * 8 + int(qptr[-1].item()) * 8
    out = {"workload": "CW box %.1f mm, %d pairs x %d nodes, float64" % (a.h, len(s), mesh.nn), "ms": dt * 1e3,
           "jacobians_per_s": 1.0 / dt, "bytes": nbytes, "achieved_gbs": nbytes / dt / 1e9,
           "frac_hbm": nbytes / dt / 1e9 / HBM_PEAK_GBS}
    if not a.skip_cpu:
        out["cpu_baseline"] = cpu_jac_baseline(_gpu.to_host(mesh.user_rows(phi)), sd, cfg)
    return out


def main():
    a = _args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (--gpus %d refused)" % a.gpus)
        sys.exit(2)
    import torch
    import redbirdpy_amd as rb
    from redbirdpy_amd import _gpu, _abi, workloads
    from redbirdpy_amd.forward import forward_device, locate_optodes

    torch.cuda.set_device(0)
    if a.copy_calibration:
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 512 MiB
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print("copied 512 MiB x3", flush=True)
        return
    cfg = workloads.fd_box(a.h)
    cfg, sd = rb.meshprep(cfg)
    mesh = _gpu.mesh_for(cfg)
    if a.only_assembly:
        print(json.dumps({"assembly": assembly_bench(a, cfg)}), flush=True)
        return
    opt = locate_optodes(cfg)
    k = opt[""].k
    L = _abi.lib()
    tuned = {}
    if a.autotune:
        import tempfile

        # the run's own cache file (a user's cache is left alone)
        os.environ.setdefault("RB_TUNE_CACHE", os.path.join(tempfile.gettempdir(), "rb_bench_tune_%d.json" % os.getpid()))
        tuned["mixed"] = _tuned(cfg, "mixed", a.tol_mixed)
        if not a.skip_fp64:
            tuned["fp64"] = _tuned(cfg, "fp64", a.tol_fp64)

    def step(prec, tol):
        return forward_device(cfg, [1], prec, tol, optodes=opt)

    def timed(prec, tol, steps, warmup, prof=False):
        for _ in range(warmup):
            step(prec, tol)
        torch.cuda.synchronize()
        if prof:
            L.rb_prof_reset()
            L.rb_prof_enable(1)
        t0 = time.perf_counter()
        res = None
        for _ in range(steps):
            res = step(prec, tol)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        L.rb_prof_enable(0)
        return res, dt

    # the timed steps carry no per-kernel events (a start/stop pair on every
    # dispatch costs ~9 % of the step); the kernel durations come from the
    # same number of steps run right after with the events on
    res_m, dt_m = timed("mixed", a.tol_mixed, a.steps, a.warmup)
    _, dt_prof = timed("mixed", a.tol_mixed, a.steps, 0, prof=True)
    import ctypes

    ms = ctypes.c_double(0)
    cnt = ctypes.c_int64(0)
    _abi.call("rb_prof_read", _abi.RB_PROF_SPMM, ctypes.byref(ms), ctypes.byref(cnt))
    spmm_avg_s = ms.value / 1e3 / max(cnt.value, 1)
    per_kernel = {}
    for name, kid in (("spmm", 0), ("update", 1), ("pupdate", 2), ("resid64", 3), ("init", 4), ("poly", 5)):
        m2 = ctypes.c_double(0)
        c2 = ctypes.c_int64(0)
        _abi.call("rb_prof_read", kid, ctypes.byref(m2), ctypes.byref(c2))
        per_kernel[name] = {"launches_per_step": c2.value / a.steps, "ms_per_step": m2.value / a.steps}
    C = k
    nbytes = _spmm_bytes(mesh.nn, mesh.nnz, C, 8)  # complex64 inner solve
    achieved = nbytes / spmm_avg_s / 1e9
    pk = per_kernel["poly"]
    degree2 = False
    if pk["launches_per_step"] > 0:
        # the polynomial preconditioner's product, LDS-staged with its z epilogue: the records of
        # A D^-1, r read once, z written once, plus D^-1.  Degree 2 takes two launches per iteration:
        # w1 = A D^-1 r (the q = A p bytes), then the epilogue's launch on w1, which also reads r
        degree2 = pk["launches_per_step"] > 1.5 * per_kernel["spmm"]["launches_per_step"]
        pbytes = nbytes + (mesh.nn * C * 8 + mesh.nn * 8) / 2 if degree2 else nbytes + mesh.nn * 8
        pk["degree"] = 2 if degree2 else 1
        pavg = pk["ms_per_step"] / 1e3 / pk["launches_per_step"]
        pk.update({"bytes_per_launch": pbytes, "avg_launch_us": pavg * 1e6, "achieved_GBs": pbytes / pavg / 1e9})

    traffic = traffic_plain = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        rec = tj.get(SPMM_KERNEL)
        if rec and a.h == 1.0:
            traffic = rec["hbm_bytes_per_launch"]
            traffic_plain = rec.get("hbm_bytes_median_launch")
        prec_ = tj.get(SPMM_KERNEL.replace(",false,false>", ",false,true>"))
        if prec_ and a.h == 1.0 and "bytes_per_launch" in per_kernel["poly"] and not degree2:
            per_kernel["poly"]["traffic"] = prec_["hbm_bytes_per_launch"]
    out = {
        "metric": "RHS-solves/s + GN iters/s @115k nodes: FP32+refine vs FP64, 1 GPU, <=1e-6",
        "value": k / dt_m,
        "unit": "RHS-solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt_m * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+c128-refine",
        "data": "synthetic",
        "config": {
            "workload": "FD forward, 60x60x30 mm Kuhn box at %.1f mm (%d nodes, %d tets, nnz %d), sphere inclusion, "
                        "100 MHz, 32 src + 32 det = %d complex RHS in one multi-RHS solve" % (
                            a.h, mesh.nn, mesh.ne, mesh.nnz, k),
            "nodes": mesh.nn,
            "rhs": k,
            "tol_mixed": a.tol_mixed,
        },
        "mixed": {"iters": res_m.info.iters, "relres": res_m.info.relres, "flag": res_m.info.flag,
                  "kernels": per_kernel, "ms_per_step_with_kernel_events": dt_prof * 1e3},
        "autotune": tuned,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated, mean over dispatches "
                              "(profiles/spmm_lds_fd115k_c64_pmc.md)",
            "traffic_median_launch": traffic_plain,
            "kernel": SPMM_KERNEL + " (q = A p with the p.q reduction, neighbour rows staged in LDS)",
            "bytes_per_launch": nbytes,
            "avg_launch_us": spmm_avg_s * 1e6,
        },
    }
    if not a.skip_fp64:
        res_d, dt_d = timed("fp64", a.tol_fp64, max(1, a.steps // 2), 1)
        phi_d = res_d.X
        phi_m = res_m.X
        err_phi = (torch.linalg.norm(phi_m - phi_d) / torch.linalg.norm(phi_d)).item()
        dm, dd = res_m.detvals[0], res_d.detvals[0]
        err_det = (torch.abs(dm - dd) / torch.abs(dd)).max().item()
        # relative to each source's largest detector value (the element-wise maximum is set by the
        # farthest detectors, whose values lie ~1e-6 below the largest)
        err_colmax = (torch.abs(dm - dd) / torch.abs(dd).amax(dim=0, keepdim=True)).max().item()
        out["fp64"] = {"rhs_solves_per_s": k / dt_d, "ms_per_step": dt_d * 1e3, "iters": res_d.info.iters,
                       "relres": res_d.info.relres, "tol": a.tol_fp64}
        out["mixed"]["err_phi_relL2_vs_fp64"] = err_phi
        out["mixed"]["err_detval_maxrel_vs_fp64"] = err_det
        out["mixed"]["err_detval_colmax_vs_fp64"] = err_colmax
        out["speedup_mixed_vs_fp64"] = dt_d / dt_m
    if a.gn:
        out["gn"] = gn_bench(a)
    if a.batch:
        out["batch"] = batch_bench(a)
    if not a.no_kernels:
        out["jacobian"] = jac_bench(a)
        out["gram"] = gram_bench(mesh.nn)
        out["assembly"] = assembly_bench(a, cfg)
    if not a.skip_cpu:
        out["cpu_baseline"] = cpu_baseline(a.cpu_h)
        out["config"]["cpu_sample_nodes"] = out["cpu_baseline"]["nodes"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
