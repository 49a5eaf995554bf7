#!/usr/bin/env python
"""Throughput of the batched RCWA solve path on one MI355X.

Workload (one "step"): the Bloch-sweep workload at 81 harmonics -- pw (9,9),
3-layer stack [disc, gap, disc] (128x128 pixmap eps 12 with a centred disc
r = 0.4 of eps 1, depth 0.55; uniform eps 1 gap, depth 1.1), 64 frequencies
linspace(0.45, 0.65) x a 32x32 cell-centred Bloch grid = 65 536 solves, each
(device, lambda, k_par) -> (R, T).  Inputs are built on the device before
the timed region; a step is Crystal.solve() + poynting_flux_end() over all
solves (R/T copied back to the host).

Prints one JSON line (the benchmark contract) with a roofline entry for the
dominant kernel (HIP events on its launch stream) and a CPU baseline (the
oracle on a bounded sample of the same workload, 1 BLAS thread).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s @81 harmonics, plain FP64 vs mixed+batched; grad evals/s; max|ΔT| vs FP64"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def eig_flops(n):
    # Hessenberg + complex Schur + triangular eigenvectors: 110 n^3 real flops
    # per matrix, minus the 4 n^3 back-transform W = Z X that runs as a GEMM.
    return 106.0 * n ** 3


def counted_flops_per_solve(n):
    """Real flops the library's own op sequence spends per disc_stack_bz solve (one unique
    patterned layer, the R/T-only chain), complex multiply-add = 8 flops:
    P Q (two N x N x n GEMMs) 4 n^3; eigensolver 106 n^3 and its
    back-transform W = Z X 8 n^3; V = Q W (two GEMMs) 4 n^3; W^-1 from the
    eigensolver's factors (one upper triangular solve with n right-hand sides
    and one GEMM) 4 + 8 n^3; V^-1 V0 (LU + n right-hand sides) 8/3 + 8 n^3;
    A^-1 (LU + n right-hand sides) and G2 = A^-1 B (one GEMM) 8/3 + 8 + 8 n^3;
    T^-1 R2 (LU + 2n right-hand sides) 8/3 + 16 n^3; four n x n x n GEMMs
    (X B G1, M B, M A, B G2) 32 n^3; the one chain junction (one GEMM, one LU,
    vector solves) 11 n^3."""
    return (4 + 106 + 8 + 4 + (4 + 8) + (8 / 3 + 8) + (8 / 3 + 8 + 8) + (8 / 3 + 16) + 32 + 11) * n ** 3


def cpu_info():
    """Host cores this process may use and the CPU model (/proc/cpuinfo)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": usable, "cpu_count": os.cpu_count(), "cpu_model": model}


def _pool_worker(args):
    """One process of the process-pool CPU baseline: the oracle on its share
    of the sample with one BLAS thread; returns (solves, seconds)."""
    cfg_idx, pos = args
    from threadpoolctl import threadpool_limits
    from khepri_amd import workloads
    from oracle import crystal as OC
    cfg = workloads.disc_stack_bz()
    cl = OC.Crystal(cfg["pw"])
    cl.add_layer_uniform("S1", cfg["eps_gap"], cfg["depth_gap"])
    cl.add_layer_pixmap("Scyl", cfg["pixmap"], cfg["depth_disc"])
    cl.set_device(["Scyl", "S1", "Scyl"], False)
    with threadpool_limits(limits=1), np.errstate(all="ignore"):
        t0 = time.perf_counter()
        for p in pos:
            s = cfg_idx[p]
            cl.set_source(1 / cfg["freqs"][s], 1.0, 0.0, kp=tuple(cfg["kp"][s]))
            cl.solve()
            cl.poynting_flux_end()
        return len(pos), time.perf_counter() - t0


def cpu_baseline_modes(cfg, idx_all, sample, workers):
    """The two CPU layouts the reference's users run: one process whose BLAS
    uses every core, and a pool of single-threaded processes (the pattern of
    khepri's examples/crystal_api/bic.py, multiprocessing.Pool with
    OPENBLAS_NUM_THREADS=1).  Same oracle, evenly spread solves of the
    workload."""
    import multiprocessing as mp
    from threadpoolctl import threadpool_limits
    from oracle import crystal as OC
    info = cpu_info()
    out = dict(info)
    pos = np.linspace(0, len(idx_all) - 1, sample).astype(int)
    cl = OC.Crystal(cfg["pw"])
    cl.add_layer_uniform("S1", cfg["eps_gap"], cfg["depth_gap"])
    cl.add_layer_pixmap("Scyl", cfg["pixmap"], cfg["depth_disc"])
    cl.set_device(["Scyl", "S1", "Scyl"], False)
    with threadpool_limits(limits=info["nproc"]), np.errstate(all="ignore"):
        t0 = time.perf_counter()
        for p in pos:
            s = idx_all[p]
            cl.set_source(1 / cfg["freqs"][s], 1.0, 0.0, kp=tuple(cfg["kp"][s]))
            cl.solve()
            cl.poynting_flux_end()
        el = time.perf_counter() - t0
    out["all_core_blas"] = {"value": round(sample / el, 3), "unit": "solves/s", "cores": info["nproc"],
                            "kind": "port", "sample": f"{sample} solves, one process, BLAS on {info['nproc']} threads"}
    nw = max(1, min(workers, info["nproc"]))
    per = max(1, (sample * 4) // nw)  # every worker gets about four times the one-process sample / workers
    wpos = np.linspace(0, len(idx_all) - 1, per * nw).astype(int)
    chunks = [(idx_all, wpos[w::nw]) for w in range(nw)]
    ctx = mp.get_context("spawn")  # fresh interpreters: the children never touch the GPU
    env_old = os.environ.get("OPENBLAS_NUM_THREADS")
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    try:
        with ctx.Pool(nw) as pool:
            res = pool.map_async(_pool_worker, chunks).get(timeout=600)
    finally:
        if env_old is None:
            os.environ.pop("OPENBLAS_NUM_THREADS", None)
        else:
            os.environ["OPENBLAS_NUM_THREADS"] = env_old
    done = sum(r[0] for r in res)
    slowest = max(r[1] for r in res)
    out["process_pool"] = {"value": round(done / slowest, 3), "unit": "solves/s", "cores": nw, "kind": "port",
                           "sample": f"{done} solves over {nw} single-threaded processes, rate = solves / the "
                                     f"slowest worker's solve loop ({slowest:.1f} s; process start-up excluded)"}
    return out


def grad_grid_rate(steps=2):
    """grad evals/s of the adjoint workload (pw (11,11), four 64 x 64 layers,
    16 frequencies): one grad eval = one frequency's dT/d eps for every pixel,
    forward + backward."""
    import torch
    from khepri_amd import Crystal, workloads
    from khepri_amd.adjoint import flux_gradient
    cfg = workloads.grad_grid()
    cl = Crystal(cfg["pw"])
    names = []
    for j, pix in enumerate(cfg["pixmaps"]):
        cl.add_layer_pixmap(f"L{j}", pix, cfg["depth"])
        names.append(f"L{j}")
    cl.set_device(names, False)
    cl.set_source(1 / cfg["freqs"], 1.0, 0.0)
    for _ in range(2):  # warm-up: GEMM tile autotuning for this workload's shapes, then one plain step
        flux_gradient(cl, wT=1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        flux_gradient(cl, wT=1.0)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    nf = len(cfg["freqs"])
    return {"value": round(nf * steps / el, 3), "unit": "grad evals/s",
            "sample": f"{steps} steps of {nf} frequencies, dT/d eps over 4 x 64 x 64 pixels each, after two warm-up steps"}


def mixed_leg(nsolves, R64, T64, idx_all):
    """The mixed-precision path (KHX_LU_MIXED=1: FP32-MFMA products in the
    layer solves, one FP64 residual refinement step) on an evenly spread
    subset of the disc_stack_bz workload (65 536 solves), timed over one step after a warm-up step, and
    its max |dR|, |dT| against the FP64 line's own values for those solves."""
    import torch
    cl, _, idx = build_crystal(nsolves)
    pos = np.minimum(np.searchsorted(idx_all, idx), len(idx_all) - 1)
    same = idx_all[pos] == idx          # solves that the FP64 line also ran
    os.environ["KHX_LU_MIXED"] = "1"
    try:
        cl.solve()
        cl.poynting_flux_end()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cl.solve()
        R, T = cl.poynting_flux_end()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
    finally:
        del os.environ["KHX_LU_MIXED"]
    dR = dT = 0.0
    for r, t, p, ok in zip(R, T, pos, same):
        if ok:
            dR = parity_max(dR, R64[p], r)
            dT = parity_max(dT, T64[p], t)
    return {"value": round(len(idx) / el, 3), "unit": "solves/s", "solves": int(len(idx)),
            "max_abs_dR_vs_fp64": dR, "max_abs_dT_vs_fp64": dT,
            "precision": "mixed: FP64 LU, FP32-MFMA triangular-solve products, one FP64 residual refinement"}


def graph_leg(cl, steps):
    """The disc_stack_bz step replayed from one captured HIP graph (khepri_amd.graph.SolveGraph:
    all chunks of the solve and the R/T reduction), timed over `steps` replays."""
    import torch
    from khepri_amd.graph import SolveGraph
    g = SolveGraph(cl)
    g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return {"value": round(cl.batch * steps / el, 3), "unit": "solves/s", "steps": steps,
            "ms_per_step": round(el / steps * 1e3, 3)}


def build_crystal(nsolves):
    from khepri_amd import Crystal, workloads
    cfg = workloads.disc_stack_bz()
    idx = np.arange(len(cfg["freqs"]))
    if nsolves < len(idx):
        idx = np.linspace(0, len(idx) - 1, nsolves).astype(int)
    cl = Crystal(cfg["pw"])
    cl.add_layer_uniform("S1", cfg["eps_gap"], cfg["depth_gap"])
    cl.add_layer_pixmap("Scyl", cfg["pixmap"], cfg["depth_disc"])
    cl.set_device(["Scyl", "S1", "Scyl"], False)
    cl.set_source(1 / cfg["freqs"][idx], 1.0, 0.0, kp=cfg["kp"][idx])
    return cl, cfg, idx


def parity_max(acc, ref, got):
    """Running max|ref - got| over the sample.  Both finite: the difference;
    both non-finite with the same NaN / inf pattern: no contribution (the
    reference's evanescent-incidence NaN is reproduced); any mismatch of
    finiteness in either direction: inf."""
    ref, got = complex(ref), complex(got)
    rf = np.isfinite(ref.real) and np.isfinite(ref.imag)
    gf = np.isfinite(got.real) and np.isfinite(got.imag)
    if rf and gf:
        d = abs(ref - got)
    elif not rf and not gf and np.isnan(ref.real) == np.isnan(got.real):
        d = 0.0
    else:
        d = float("inf")
    return d if d > acc else acc


def cpu_baseline(cfg, gpu_R, gpu_T, idx_all, sample):
    """The oracle (numpy restatement, same op sequence as the reference) on a
    bounded, evenly spread sample of the workload; 1 BLAS thread."""
    from threadpoolctl import threadpool_limits
    from oracle import crystal as OC
    pos = np.linspace(0, len(idx_all) - 1, sample).astype(int)
    cl = OC.Crystal(cfg["pw"])
    cl.add_layer_uniform("S1", cfg["eps_gap"], cfg["depth_gap"])
    cl.add_layer_pixmap("Scyl", cfg["pixmap"], cfg["depth_disc"])
    cl.set_device(["Scyl", "S1", "Scyl"], False)
    dR = dT = 0.0
    with threadpool_limits(limits=1), np.errstate(all="ignore"):
        t0 = time.perf_counter()
        for p in pos:
            s = idx_all[p]
            cl.set_source(1 / cfg["freqs"][s], 1.0, 0.0, kp=tuple(cfg["kp"][s]))
            cl.solve()
            r, t = cl.poynting_flux_end()
            dR = parity_max(dR, r, gpu_R[p])
            dT = parity_max(dT, t, gpu_T[p])
        el = time.perf_counter() - t0
    return dict(value=round(sample / el, 3), unit="solves/s", cores=1, kind="port",
                sample=f"{sample} solves evenly spread over the 65536-solve workload, oracle (numpy complex128, "
                       f"reference op sequence incl. reverse star chain), 1 BLAS thread, {el:.1f} s"), dR, dT


def plain_batch1(cfg, idx_all, sample):
    """The plain FP64 path one solve at a time (batch 1, no batching across
    sources): the same kernels and algorithm, timed on an evenly spread
    sample of the workload."""
    import torch
    from khepri_amd import Crystal
    pos = np.linspace(0, len(idx_all) - 1, sample).astype(int)
    cl = Crystal(cfg["pw"])
    cl.add_layer_uniform("S1", cfg["eps_gap"], cfg["depth_gap"])
    cl.add_layer_pixmap("Scyl", cfg["pixmap"], cfg["depth_disc"])
    cl.set_device(["Scyl", "S1", "Scyl"], False)
    s0 = idx_all[pos[0]]
    cl.set_source(1 / cfg["freqs"][s0], 1.0, 0.0, kp=tuple(cfg["kp"][s0]))
    cl.solve()
    cl.poynting_flux_end()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p in pos:
        s = idx_all[p]
        cl.set_source(1 / cfg["freqs"][s], 1.0, 0.0, kp=tuple(cfg["kp"][s]))
        cl.solve()
        cl.poynting_flux_end()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return dict(value=round(sample / el, 3), unit="solves/s", sample=f"{sample} solves one at a time, {el:.2f} s")


def extra_workload(name, args):
    """Secondary measurements (not the main line): returns a JSON dict."""
    import torch
    from khepri_amd import Crystal, Expansion, Layer, workloads
    from khepri_amd.batch import CrystalBatch
    cpu = None  # oracle timing of a bounded sample, for the workloads that have one
    if name == "disc_sweep":
        cfg = workloads.disc_stack_sweep()
        cl = Crystal(cfg["pw"])
        cl.add_layer_uniform("S1", cfg["eps_gap"], cfg["depth_gap"])
        cl.add_layer_pixmap("Scyl", cfg["pixmap"], cfg["depth_disc"])
        cl.set_device(["Scyl", "S1", "Scyl"], False)
        cl.set_source(1 / cfg["freqs"], 1.0, 0.0)

        def step():
            cl.solve(rt_only=True)
            return cl.poynting_flux_end()
        units, unit = len(cfg["freqs"]), "solves/s"
    elif name == "random_devices":
        devs = workloads.random_device_params()[:args.devices]
        cls = []
        for dv in devs:
            c = Crystal(dv["pw"])
            names = []
            for j, ly in enumerate(dv["layers"]):
                c.add_layer_pixmap(f"L{j}", workloads.render_layer(ly), ly["depth"])
                names.append(f"L{j}")
            c.set_device(names, False)
            c.set_source(1 / dv["freq"], 1.0, 1.0, dv["theta"], dv["phi"])
            cls.append(c)
        batch = CrystalBatch(cls)

        def step():
            return batch.solve()
        units, unit = len(cls), "devices/s"
    elif name == "twisted":
        # every twist angle is its own extended expansion (625 harmonics): one Crystal per angle,
        # each solving all frequencies in one batch
        cfg = workloads.twisted_bilayer()
        fr = cfg["freqs"][:args.freqs]
        cls = []
        for ang in cfg["angles_deg"][:args.angles]:
            e1, e2 = Expansion(cfg["pw"]), Expansion(cfg["pw"])
            e2.rotate(np.deg2rad(ang))
            etw = e1 + e2
            cl = Crystal.from_expansion(etw)
            cl.add_layer("L1", Layer.pixmap(e1, cfg["pixmap"], cfg["depth"]), extended=True)
            cl.add_layer("gap", Layer.uniform(etw, 1.0, cfg["gap"]))
            cl.add_layer("L2", Layer.pixmap(e2, cfg["pixmap"], cfg["depth"]), extended=True)
            cl.set_device(["L1", "gap", "L2"], False)
            cl.set_source(1 / fr, 1.0, 0.0)
            cls.append(cl)

        def step():
            out = []
            for cl in cls:
                cl.solve(rt_only=True)
                out.append(cl.poynting_flux_end())
            return out
        units, unit = len(fr) * len(cls), "solves/s"
    elif name == "twisted_fields":
        # the twisted bilayer's field output: for every seeded angle at the middle frequency, solve with
        # the full chain and evaluate E and H on a 64 x 64 x 32 volume (x, y over the cell, z over 0..1)
        cfg = workloads.twisted_bilayer()
        fmid = cfg["freqs"][len(cfg["freqs"]) // 2]
        xs = np.linspace(0, 1, 64)
        X, Y = np.meshgrid(xs, xs)
        Zs = np.linspace(0, 1.0, 32)
        cls = []
        for ang in cfg["angles_deg"][:args.angles]:
            e1, e2 = Expansion(cfg["pw"]), Expansion(cfg["pw"])
            e2.rotate(np.deg2rad(ang))
            etw = e1 + e2
            cl = Crystal.from_expansion(etw)
            cl.add_layer("L1", Layer.pixmap(e1, cfg["pixmap"], cfg["depth"]), extended=True)
            cl.add_layer("gap", Layer.uniform(etw, 1.0, cfg["gap"]))
            cl.add_layer("L2", Layer.pixmap(e2, cfg["pixmap"], cfg["depth"]), extended=True)
            cl.set_device(["L1", "gap", "L2"], True)
            cl.set_source(1 / fmid, 1.0, 0.0)
            cls.append(cl)

        def step():
            out = []
            for cl in cls:
                cl.solve()
                out.append(cl.fields_volume(X, Y, Zs))
            return out
        units, unit = len(cls), "field volumes/s"
    elif name == "grad_grid":
        from khepri_amd.adjoint import flux_gradient
        cfg = workloads.grad_grid()
        cl = Crystal(cfg["pw"])
        names = []
        for j, pix in enumerate(cfg["pixmaps"]):
            cl.add_layer_pixmap(f"L{j}", pix, cfg["depth"])
            names.append(f"L{j}")
        cl.set_device(names, False)
        cl.set_source(1 / cfg["freqs"], 1.0, 0.0)

        def step():
            # forward (R, T) + backward: one dT/d eps map per frequency for all 4 x 64 x 64 pixels
            return flux_gradient(cl, wT=1.0)
        units, unit = len(cfg["freqs"]), "grad evals/s"
    elif name == "bz_fields":
        # the reference's Brillouin-zone field integration (examples/bzi/twisted_phc.py:126-147) at pw (9,9)
        # on the disc stack: a tilted Gaussian beam on 8 x 8 cells of 16 x 16 points, its Bloch amplitudes at
        # 8 x 8 Bloch vectors, one batched solve with the field chain, one batched fields_volume on a
        # 32 x 32 x 8 grid per Bloch vector (E and H copied to the host); unit = Bloch vectors
        from functools import partial
        from khepri_amd import beams as B
        cfg = workloads.disc_stack_bz()
        ns, bzs, wl = 16, (8, 8), 1 / 0.55
        xs = np.arange(ns * bzs[0]) / ns
        X, Y = np.meshgrid(xs, xs)
        Z = np.zeros_like(X)
        geo = (4.0, 4.0, -0.5, 0.2, 0.3, 0.1)
        kps = workloads.bz_grid(bzs).reshape(2, -1).T
        xg, yg = np.meshgrid(np.linspace(0, 1, 32), np.linspace(0, 1, 32))
        zs = np.linspace(-0.2, 2.4, 8)
        f = B.shifted_rotated_fields(partial(B._paraxial_gaussian_field_fn, beam_waist=2.0), X, Y, Z, wl, *geo)
        fr = torch.as_tensor(np.swapaxes(np.swapaxes(f[..., 0], 0, 2), 1, 3), device="cuda")
        Xd, Yd = torch.as_tensor(X, device="cuda"), torch.as_tensor(Y, device="cuda")

        def build(cls_):
            c = cls_(cfg["pw"])
            c.add_layer_uniform("S1", cfg["eps_gap"], cfg["depth_gap"])
            c.add_layer_pixmap("Scyl", cfg["pixmap"], cfg["depth_disc"])
            c.set_device(["Scyl", "S1", "Scyl"], [True] * 3)
            return c
        cl = build(Crystal)

        def step():
            cl.set_source(np.full(len(kps), wl), np.nan, np.nan, kp=kps)
            cl.solve()
            F = np.stack([B.amplitudes_from_fields(fr, cl.expansion, wl, tuple(k), Xd, Yd, bzs).reshape(-1)
                          for k in kps])
            E, H = cl.fields_volume(xg, yg, zs, incident_fields=F)
            return E.sum(axis=0), H.sum(axis=0)
        units, unit = len(kps), "Bloch vectors/s"

        def cpu(nsample):
            from oracle import beams as OB, crystal as OC
            oc = build(OC.Crystal)
            fo = OB.shifted_rotated_fields(partial(OB.paraxial_gaussian_field_fn, beam_waist=2.0), X, Y, Z, wl, *geo)
            fro = np.swapaxes(np.swapaxes(fo[..., 0], 0, 2), 1, 3)
            t0 = time.perf_counter()
            for k in kps[:nsample]:
                oc.set_source(wl, np.nan, np.nan, kp=tuple(k))
                oc.solve()
                Fo = OB.amplitudes_from_fields(fro, oc.expansion.g, cfg["pw"], tuple(k), X, Y, bzs).reshape(-1)
                oc.fields_volume(xg, yg, zs, incident_fields=Fo)
            return nsample / (time.perf_counter() - t0), f"{nsample} Bloch vectors of the same workload, the reference's loop"
    elif name == "islands":
        # analytical island layer (disc + rectangle, closed-form coefficients) in the disc stack, pw (9, 9),
        # 64 frequencies x 8 x 8 Bloch vectors
        isl = [{"type": "disc", "params": [0.5, 0.5, 0.3], "epsilon": 1.0},
               {"type": "rectangle", "params": [0.05, 0.05, 0.25, 0.9], "epsilon": 4.0}]
        fr = np.repeat(np.linspace(0.45, 0.65, 64), 64)
        kp = np.tile(workloads.bz_grid((8, 8)).reshape(2, -1).T, (64, 1))

        def build(cls_, pw=(9, 9)):
            c = cls_(pw)
            c.add_layer_uniform("S1", 1.0, 1.1)
            c.add_layer_analytical("Sisl", isl, 12.0, 0.55)
            c.set_device(["Sisl", "S1", "Sisl"], False)
            return c
        cl = build(Crystal)
        cl.set_source(1 / fr, 1.0, 0.0, kp=kp)

        def step():
            cl.solve()
            return cl.poynting_flux_end()
        units, unit = len(fr), "solves/s"

        def cpu(nsample):
            from oracle import crystal as OC
            oc = build(OC.Crystal)
            pos = np.linspace(0, len(fr) - 1, nsample).astype(int)
            t0 = time.perf_counter()
            for p_ in pos:
                oc.set_source(1 / fr[p_], 1.0, 0.0, kp=tuple(kp[p_]))
                oc.solve()
                oc.poynting_flux_end()
            return nsample / (time.perf_counter() - t0), f"{nsample} solves of the same workload"
    elif name == "bands":
        # Bloch eigenpairs of 2 048 period S-matrices (the disc-stack period, pw (5, 5): 4ng = 100),
        # 32 frequencies x 8 x 8 Bloch vectors, solved in one batch
        from khepri_amd import eigentricks as E
        cfg = workloads.disc_stack_bz()
        fr = np.repeat(np.linspace(0.45, 0.65, 32), 64)
        kp = np.tile(workloads.bz_grid((8, 8)).reshape(2, -1).T, (32, 1))
        cl = Crystal((5, 5))
        cl.add_layer_uniform("S1", cfg["eps_gap"], cfg["depth_gap"])
        cl.add_layer_pixmap("Scyl", cfg["pixmap"], cfg["depth_disc"])
        cl.set_device(["Scyl", "S1"], False)
        cl.set_source(1 / fr, 1.0, 0.0, kp=kp)
        cl.solve(rt_only=False)
        S = E.scattering_matrix(cl)
        del cl

        def step():
            return E.scattering_eigenvalues(S)
        units, unit = S.shape[0], "S-matrices/s"

        def cpu(nsample):
            from oracle import eigentricks as OE
            Sn = S[np.linspace(0, S.shape[0] - 1, nsample).astype(int)].cpu().numpy()
            t0 = time.perf_counter()
            for b in range(nsample):
                OE.scattering_eigenvalues(Sn[b])
            return nsample / (time.perf_counter() - t0), f"{nsample} S-matrices of the same batch, scipy QZ eig(Sl, Sr)"
    elif name == "beams":
        # Gaussian beam (tilted, shifted) sampled on 8 x 8 unit cells of 32 x 32 points, then its Bloch
        # amplitudes (pw (9, 9)) at 16 Bloch vectors: one step = fields + 16 amplitude sets
        from functools import partial
        from khepri_amd import beams as B
        ns, bzs = 32, (8, 8)
        xs = np.arange(ns * bzs[0]) / ns
        X, Y = np.meshgrid(xs, xs)
        Z = np.zeros_like(X)
        geo = (4.0, 4.0, -0.5, 0.2, 0.3, 0.1)
        kps = workloads.bz_grid((4, 4)).reshape(2, -1).T
        e = Expansion((9, 9))

        def step():
            f = B.shifted_rotated_fields(partial(B._paraxial_gaussian_field_fn, beam_waist=2.0), X, Y, Z, 0.63, *geo)
            fr_ = np.swapaxes(np.swapaxes(f[..., 0], 0, 2), 1, 3)
            return [B.amplitudes_from_fields(fr_, e, 0.63, tuple(k), X, Y, bzs) for k in kps]
        units, unit = len(kps), "k-points/s"

        def cpu(nsample):
            from khepri_amd.expansion import Expansion as PE
            from oracle import beams as OB
            gv = PE((9, 9)).g_vectors
            t0 = time.perf_counter()
            f = OB.shifted_rotated_fields(partial(OB.paraxial_gaussian_field_fn, beam_waist=2.0), X, Y, Z, 0.63, *geo)
            fr_ = np.swapaxes(np.swapaxes(f[..., 0], 0, 2), 1, 3)
            for k in kps[:nsample]:
                OB.amplitudes_from_fields(fr_, gv, (9, 9), tuple(k), X, Y, bzs)
            return nsample / (time.perf_counter() - t0), f"fields + {nsample} amplitude sets of the same workload"
    else:
        raise SystemExit(f"unknown workload {name}")
    tuned = run_autotune(step, args)
    graph = None
    if args.graph:
        # the step captured once as a HIP graph (khepri_amd.graph) and replayed
        from khepri_amd.graph import GradientGraph, SolveGraph
        if name == "disc_sweep":
            graph = SolveGraph(cl, rt_only=True)
        elif name == "grad_grid":
            graph = GradientGraph(cl, wT=1.0)
        else:
            raise SystemExit("--graph applies to the disc_sweep and grad_grid workloads")
        step = graph.replay
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    out = {"metric": f"{name} throughput", "value": round(units * args.steps / el, 3), "unit": unit,
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 3),
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": name, "units_per_step": units, "hip_graph": graph is not None},
           "gemm_autotune": tuned}
    if cpu is not None and args.cpu_sample > 0:
        from threadpoolctl import threadpool_limits
        with threadpool_limits(limits=1), np.errstate(all="ignore"):
            v, what = cpu(min(args.cpu_sample, {"islands": 30, "bands": 200, "beams": 4, "bz_fields": 2}[name]))
        out["cpu_baseline"] = {"value": round(v, 3), "unit": unit, "cores": 1, "kind": "port",
                               "sample": f"oracle, 1 BLAS thread: {what}"}
    return out


def run_autotune(step, args):
    """Record the GEMMs of one (untimed) step, time the tile shapes on each and
    install the fastest; returns a summary for the JSON line (None when off)."""
    if not args.autotune:
        return None
    from khepri_amd import autotune
    t0 = time.perf_counter()
    table = autotune.autotune(step)
    changed = sum(1 for e in table.values() if e["tile"] != autotune.default_tile(*e["shape"][:2]))
    mfma = sum(1 for e in table.values() if e["tile"] >= 4)
    print(f"[bench] gemm autotune: {len(table)} classes, {changed} changed, {mfma} on MFMA", file=sys.stderr,
          flush=True)
    return {"classes": len(table), "changed": changed, "mfma": mfma, "seconds": round(time.perf_counter() - t0, 2)}


def load_traffic(kernel, workload, matrices_per_launch, avg_launch_ms):
    """HBM bytes per launch from the committed FETCH_SIZE/WRITE_SIZE profile of
    the same workload and launch size (None when the profile does not match,
    or when the kernel's launch time moved by more than 5 % since, i.e. the
    profile is of a different build of the kernel)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d.get("workload") != workload:
        return None
    if abs(d.get("matrices_per_launch", 0) - matrices_per_launch) > 0.01 * matrices_per_launch:
        return None
    ref_ms = d.get("avg_launch_ms")
    if ref_ms is None or abs(ref_ms - avg_launch_ms) > 0.05 * ref_ms:
        return None
    return d.get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--solves", type=int, default=65536, help="subset of the workload (default: all)")
    ap.add_argument("--cpu-sample", type=int, default=100, help="oracle solves timed for cpu_baseline (0: skip; 100 = about 10 s of CPU work)")
    ap.add_argument("--b1-sample", type=int, default=24, help="solves timed one at a time for the plain batch-1 line")
    ap.add_argument("--grad-steps", type=int, default=4,
                    help="steps of the adjoint workload (16 grad evals each) timed after the main line (0: skip)")
    ap.add_argument("--mixed-solves", type=int, default=8192,
                    help="solves of the mixed-precision leg (KHX_LU_MIXED=1) after the main line (0: skip)")
    ap.add_argument("--graph-steps", type=int, default=1,
                    help="replays of the whole step captured as one HIP graph, after the main line (0: skip)")
    ap.add_argument("--cpu-workers", type=int, default=16,
                    help="processes of the process-pool CPU baseline (capped at the usable cores; 0: skip both "
                         "multi-core modes)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="run one plain device copy of this many bytes before the workload (byte-counter calibration)")
    ap.add_argument("--workload", default="disc_stack_bz",
                    help="disc_stack_bz (default, the reported line) | disc_sweep | random_devices | twisted | grad_grid | "
                         "twisted_fields | islands | bands | beams | bz_fields")
    ap.add_argument("--devices", type=int, default=4096, help="random_devices: number of devices")
    ap.add_argument("--freqs", type=int, default=32, help="twisted: number of frequencies")
    ap.add_argument("--angles", type=int, default=8, help="twisted: number of twist angles")
    ap.add_argument("--graph", action="store_true",
                    help="disc_sweep / grad_grid: replay the step from a captured HIP graph")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1 (default): pick GEMM tile shapes with khepri_amd.autotune on one untimed step first; 0: "
                         "built-in tile rule")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for {max(args.gpus, world)}): every solve is independent, "
              "so N GPUs would each run the same per-GPU workload; multi-GPU runs are not provided.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from khepri_amd import _lib

    lib = _lib.load()
    if args.copy_bytes > 0:
        nb = (args.copy_bytes // 16) * 16
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        _lib.check(lib.khx_copy(src.data_ptr(), dst.data_ptr(), nb, torch.cuda.current_stream().cuda_stream), "copy")
        torch.cuda.synchronize()
        del src, dst
    if args.workload != "disc_stack_bz":
        print(json.dumps(extra_workload(args.workload, args)), flush=True)
        return
    cl, cfg, idx = build_crystal(args.solves)
    nsol = len(idx)
    n = 2 * 81
    print(f"[bench] {nsol} solves per step, warmup {args.warmup}, steps {args.steps}", file=sys.stderr, flush=True)

    def solve_step():
        cl.solve()
        return cl.poynting_flux_end()
    tuned = run_autotune(solve_step, args)
    for w in range(args.warmup):
        cl.solve()
        R, T = cl.poynting_flux_end()
        print(f"[bench] warmup {w} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    lib.khx_prof_enable(1)
    t0 = time.perf_counter()
    for s in range(args.steps):
        cl.solve()
        R, T = cl.poynting_flux_end()
        print(f"[bench] step {s} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kern = {}
    import ctypes
    for name in ("zeig", "zgesv_kernel", "zgemm_kernel"):
        tot, cnt = ctypes.c_double(), ctypes.c_longlong()
        _lib.check(lib.khx_prof_collect(name.encode(), ctypes.byref(tot), ctypes.byref(cnt)), "prof")
        kern[name] = (tot.value, cnt.value)
    lib.khx_prof_enable(0)
    value = nsol * args.steps / el

    # dominant kernel roofline: the eigen-solver (FP64, one matrix per workgroup)
    eig_ms, eig_n = kern["zeig"]
    avg_ms = eig_ms / max(eig_n, 1)
    matrices_per_launch = nsol * args.steps / max(eig_n, 1)
    flops_per_launch = eig_flops(n) * matrices_per_launch
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    workload = "disc_stack_bz pw(9,9) 65536 solves"
    roofline = dict(bound="fp64", kernel="zeig (zeig_hess_kernel + zeig_qr_kernel + zeig_vec_kernel)", achieved=round(achieved, 4), peak=PEAK_FP64_TFLOPS,
                    unit="TFLOP/s", frac=round(achieved / PEAK_FP64_TFLOPS, 6),
                    traffic=load_traffic("zeig", workload, matrices_per_launch, avg_ms),
                    avg_launch_ms=round(avg_ms, 4), matrices_per_launch=matrices_per_launch,
                    flops_per_matrix=eig_flops(n))
    share = {k: round(v[0] / (el * 1e3 * 1.0), 4) for k, v in kern.items()}

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload if nsol == 65536 else f"disc_stack_bz pw(9,9) {nsol}-solve subset",
                   "pw": [9, 9], "harmonics": 81, "solves_per_step": nsol, "layers": ["Scyl", "S1", "Scyl"],
                   "precision": "fp64 (plain)", "chunk": cl._chunk(nsol)},
        "roofline": roofline,
        "kernel_time_share": share,
        "gemm_autotune": tuned,
    }
    if args.b1_sample > 0:
        pb = plain_batch1(cfg, idx, args.b1_sample)
        out["plain_fp64_batch1"] = pb
        out["speedup_vs_plain_batch1"] = round(out["value"] / pb["value"], 2)
    cf = counted_flops_per_solve(n)
    out["counted_flops"] = {"per_solve": cf, "tflops": round(value * cf / 1e12, 3),
                            "frac_fp64_peak": round(value * cf / 1e12 / PEAK_FP64_TFLOPS, 4),
                            "note": "the library's own op sequence (cheaper than the reference's, whose count is "
                                    "2.578e9 per solve)"}
    if args.grad_steps > 0:
        out["grad_grid"] = grad_grid_rate(args.grad_steps)
        print("[bench] grad_grid done", file=sys.stderr, flush=True)
    if args.mixed_solves > 0:
        out["mixed_batched"] = mixed_leg(min(args.mixed_solves, nsol), np.asarray(R), np.asarray(T), idx)
        print("[bench] mixed leg done", file=sys.stderr, flush=True)
    if args.graph_steps > 0:
        out["graph_replay"] = graph_leg(cl, args.graph_steps)
        print("[bench] graph replay done", file=sys.stderr, flush=True)
    if args.cpu_sample > 0:
        Rn, Tn = np.asarray(R), np.asarray(T)
        cb, dR, dT = cpu_baseline(cfg, Rn, Tn, idx, args.cpu_sample)
        out["cpu_baseline"] = cb
        out["max_abs_dR"] = dR
        out["max_abs_dT"] = dT
        if args.cpu_workers > 0:
            out["cpu_baseline_modes"] = cpu_baseline_modes(cfg, idx, args.cpu_sample, args.cpu_workers)
            print("[bench] cpu baseline modes done", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
