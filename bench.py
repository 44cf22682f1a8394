#!/usr/bin/env python3
"""Benchmark: OC iterations/s of the GPU density-method loop.

Workload (default): structured hex8 120x60x60 (432,000 elements, 1,350,723
DOF), x=0 face clamped, 8 unit point loads at seeded random nodes of the
x=120 face solved as one 8-column block, SIMP p=3, E0=1, Emin=1e-9, nu=0.3,
volume fraction 0.2 (initial density 0.2), Gaussian neighbour filter r=2.0,
Heaviside beta=1 (eta=0.5), OC move limit 0.2, eta=0.5.  The state solve is
the batched CG with the geometric-multigrid preconditioner in mixed
precision (float32 iterations, float64 reliable updates, relative residual
1e-8); the same loop in float64 runs beside it for the mixed-vs-fp64 ratio
and the compliance error.

A step is one OC iteration (filter, projection, 8-load solve, energies,
sensitivities, filter adjoint, device-side lambda bisection).  W warm-up
iterations run untimed, then K iterations are timed one by one, each
between device synchronisations; ``value`` = 1 / median per-iteration time
(``ms_per_step`` is that median; the mean is reported beside it).  The
timed iterations carry no per-kernel instrumentation: the kernel times
behind ``roofline`` come from two further, untimed iterations with HIP
events on the solve stream around every fp32 operator apply of the CG and
every fine-level launch of the V-cycle; ``roofline`` is the fine-level
modal kernel with the largest total time (``fine_level_kernels`` lists all
five).  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "scikit-topt_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "hex8-120x60x60-8load": dict(grid=(120, 60, 60), n_loads=8, vol=0.2, r=2.0),
    "hex8-40x20x20-8load": dict(grid=(40, 20, 20), n_loads=8, vol=0.2, r=2.0),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_optimizer(wl, precision, rtol, iters, dst):
    import numpy as np  # noqa: F401
    from sktopt_amd import core
    from sktopt_amd.mesh import toy_problem
    from sktopt_amd.tools import SchedulerConfig as SC
    nx, ny, nz = wl["grid"]
    task = toy_problem.multiload(nx, ny, nz, n_loads=wl["n_loads"], seed=0)
    cfg = core.OC_Config(
        dst_path=dst, max_iters=iters, p=SC.constant(target_value=3.0),
        vol_frac=SC.constant(target_value=wl["vol"]), beta=SC.constant(target_value=1.0),
        filter_type="spacial", filter_radius=SC.constant(target_value=wl["r"]),
        move_limit=SC.constant(target_value=0.2), E_min_coeff=1e-9, precision=precision,
        solver_rtol=rtol)
    opt = core.OC_Optimizer(cfg, task)
    opt.parameterize()
    return opt


APPLY_KERNEL = "skt::k_hex_apply_modal<float, 3, 2, 0, 0, true>"  # plain apply + fused p.Kp
# the V-cycle's fine-level variants in the order skt_solver_get_vcycle_timing
# reports them: fused first two Chebyshev steps, residual, Chebyshev step,
# last Chebyshev step with the fused r . z
VCYCLE_KERNELS = ["skt::k_hex_apply_modal<float, 3, 2, 3, 0, false>",
                  "skt::k_hex_apply_modal<float, 3, 2, 1, 0, false>",
                  "skt::k_hex_apply_modal<float, 3, 2, 2, 0, false>",
                  "skt::k_hex_apply_modal<float, 3, 2, 2, 0, true>"]
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hex8_modal_apply_f32_120x60x60_8col_pmc.txt")
COPY_DOUBLES = 1 << 25            # 256 MiB read + 256 MiB written per copy


def hbm_copy(reps=5):
    """Plain device copy of a known size through the library's element-wise
    kernel: the HBM counter calibration for the pmc runs and a sanity
    bandwidth figure.  Runs outside the timed region."""
    import torch
    from sktopt_amd import _lib
    x = torch.ones(COPY_DOUBLES, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    s = torch.cuda.current_stream().cuda_stream
    for _ in range(2):
        _lib.call("skt_ewise", 0, x.numel(), 1.0, x.data_ptr(), 0.0, None, y.data_ptr(), s)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        _lib.call("skt_ewise", 0, x.numel(), 1.0, x.data_ptr(), 0.0, None, y.data_ptr(), s)
    e1.record()
    torch.cuda.synchronize()
    us = 1e3 * e0.elapsed_time(e1) / reps
    return dict(bytes=2 * 8 * COPY_DOUBLES, us=us, gbs=2 * 8 * COPY_DOUBLES / us / 1e3)


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` from the committed pmc summary
    (FETCH_SIZE and WRITE_SIZE passes, scaled by the copy calibration), if
    present."""
    key = f"traffic_bytes_per_launch {kernel}:"
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith(key):
                    return float(line[len(key):])
    except OSError:
        pass
    return None


def jacobi_iterations(opt):
    """Iterations the reference's own iterative option (Jacobi-preconditioned
    CG at rtol 1e-8, fea/solver_elastic.py:84-92) needs on the current state,
    counted with one float64 Jacobi-PCG solve on the GPU (untimed).  The CPU
    baseline charges that many Jacobi-CG iterations per load."""
    import dataclasses
    import torch
    from sktopt_amd import _lib
    fem = opt.fem
    cfg = fem.solver_config
    fem.solver_config = dataclasses.replace(cfg, precision="fp64", warm_start=False)
    _lib.call("skt_solver_configure", fem._solver, _lib.PC_JACOBI, cfg.mg_smooth_degree,
              cfg.mg_max_coarse)
    U = torch.zeros((fem.ndof, fem.n_tasks), dtype=torch.float64, device="cuda")
    try:
        fem.solve(U)
        its = fem.last_solve["iters"]
    finally:
        fem.solver_config = cfg
        _lib.call("skt_solver_configure", fem._solver, _lib.PC_AUTO, cfg.mg_smooth_degree,
                  cfg.mg_max_coarse)
        fem.solve_history.pop()
    return its


def timed_run(opt, warmup, steps):
    """W untimed iterations, then K iterations each bracketed by device
    synchronisations; returns the per-iteration wall times (s)."""
    import torch
    if warmup:
        opt.optimize_steps(warmup)
    torch.cuda.synchronize()
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        opt.optimize_steps(1)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    return times


def apply_timing_pass(opt, steps=2):
    """Untimed OC iterations with HIP events (on the solve stream) around
    every operator apply of the CG loop and every fine-level launch of the
    multigrid V-cycle: totals and counts per fp32 / fp64 apply and per
    V-cycle variant."""
    import ctypes as C
    import torch
    from sktopt_amd import _lib
    s = opt.fem._solver
    _lib.call("skt_solver_set_timing", s, 1)
    ms = (C.c_double * 2)()
    n = (C.c_int64 * 2)()
    vms = (C.c_double * 4)()
    vn = (C.c_int64 * 4)()
    _lib.call("skt_solver_get_timing", s, ms, n, 1)     # reset
    _lib.call("skt_solver_get_vcycle_timing", s, vms, vn, 1)
    opt.optimize_steps(steps)
    torch.cuda.synchronize()
    _lib.call("skt_solver_get_timing", s, ms, n, 1)
    _lib.call("skt_solver_get_vcycle_timing", s, vms, vn, 1)
    _lib.call("skt_solver_set_timing", s, 0)
    return dict(f64_ms=ms[0], f64_n=n[0], f32_ms=ms[1], f32_n=n[1],
                vcycle_ms=list(vms), vcycle_n=list(vn))


def time_launches(fn, reps=20, warm=3):
    """Average duration (us) of fn()'s launches, HIP events on the current
    stream around `reps` back-to-back calls."""
    import torch
    for _ in range(warm):
        fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return 1e3 * e0.elapsed_time(e1) / reps


def tet_oc_iterations(n=8, precision="mixed"):
    """OC iterations on the tet configuration (100x50x50 cells x 6 jittered
    tets, 1.5M tets, r = 1.5 filter, Emin = 1e-9, vol 0.3, SIMP p = 3,
    multigrid solve to rtol 1e-8) in the given precision: median of n
    iterations after one warm-up, each between device synchronisations."""
    from sktopt_amd import core
    from sktopt_amd.mesh import toy_problem
    from sktopt_amd.tools import SchedulerConfig as SC
    task = toy_problem.cantilever(100, 50, 50, kind="tet4", jitter=0.2, seed=0)
    cfg = core.OC_Config(dst_path="/tmp/skt_bench_tet", max_iters=n + 3,
                         p=SC.constant(target_value=3.0), vol_frac=SC.constant(target_value=0.3),
                         beta=SC.constant(target_value=1.0), filter_type="spacial",
                         filter_radius=SC.constant(target_value=1.5),
                         move_limit=SC.constant(target_value=0.2), E_min_coeff=1e-9,
                         precision=precision, solver_rtol=1e-8)
    opt = core.OC_Optimizer(cfg, task)
    opt.parameterize()
    return _loop_record("tet4-100x50x50x6-1load (c1)", "tet OC", opt, n, precision)


def heat_logmoc_iterations(n=6, precision="mixed"):
    """Log-space modified OC on the heat configuration (160^3 hex8 scalar
    conduction, 4.2M DOF, k0 = 1, kmin = 1e-3, source 1e-2, T = 0 on a 16x16
    patch of z = 0, vol 0.3, r = 2.0 filter, SIMP p = 3, multigrid solve to
    rtol 1e-8) in the given precision: median of n iterations after one
    warm-up."""
    from sktopt_amd import core
    from sktopt_amd.mesh import toy_problem
    from sktopt_amd.tools import SchedulerConfig as SC
    task = toy_problem.heat_patch(160, patch=16, source=1e-2)
    cfg = core.LogMOC_Config(
        dst_path="/tmp/skt_bench_heat", max_iters=n + 3, p=SC.constant(target_value=3.0),
        vol_frac=SC.constant(target_value=0.3), beta=SC.constant(target_value=1.0),
        filter_type="spacial", filter_radius=SC.constant(target_value=2.0),
        move_limit=SC.constant(target_value=0.2), eta=SC.constant(target_value=0.6),
        E_min_coeff=1e-3, precision=precision, solver_rtol=1e-8)
    opt = core.LogMOC_Optimizer(cfg, task)
    opt.parameterize()
    return _loop_record("heat-hex8-160^3 (c4)", "heat LogMOC", opt, n, precision)


def gpu_c0_iterations(n=10):
    """The CPU-runnable case on the GPU, for comparison with the restated
    reference path timed on the host: 40x20x20 hex8 cantilever, vol 0.3,
    r = 1.5, move limit 0.2, SIMP p = 3, Emin = 1e-9, FP64, solve to rtol
    1e-8.  Wall time of the first OC iteration (filter build, multigrid
    set-up and graph captures included, as the host's first iteration
    builds its filter), then the median of n more."""
    import torch
    from sktopt_amd import core
    from sktopt_amd.mesh import toy_problem
    from sktopt_amd.tools import SchedulerConfig as SC
    task = toy_problem.cantilever(40, 20, 20)
    cfg = core.OC_Config(dst_path="/tmp/skt_bench_c0", max_iters=n + 3,
                         p=SC.constant(target_value=3.0), vol_frac=SC.constant(target_value=0.3),
                         beta=SC.constant(target_value=1.0), filter_type="spacial",
                         filter_radius=SC.constant(target_value=1.5),
                         move_limit=SC.constant(target_value=0.2), E_min_coeff=1e-9,
                         precision="fp64", solver_rtol=1e-8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    opt = core.OC_Optimizer(cfg, task)
    opt.parameterize()
    opt.optimize_steps(1)
    torch.cuda.synchronize()
    first = time.perf_counter() - t0
    times = timed_run(opt, 0, n)
    t = median(times)
    log(f"gpu c0 (fp64): first iteration {first:.3f} s, then {1.0 / t:.1f} OC it/s "
        f"(median {1e3 * t:.2f} ms)")
    return dict(workload="hex8-40x20x20-1load (54,243 DOF)", precision="fp64",
                first_iteration_s=first, median_ms=1e3 * t, value=1.0 / t,
                unit="OC iterations/s", steps=n)


def _loop_record(config, label, opt, n, precision):
    times = timed_run(opt, 1, n)
    solves = [s["iters"] for s in opt.fem.solve_history[-n:]]
    t = median(times)
    comp = [float(c) for c in opt.recorder.as_object().compliance]
    log(f"{label} ({precision}): {1.0 / t:.2f} it/s (median {1e3 * t:.1f} ms), solves {solves}")
    return dict(config=config, kernel=f"{label} iteration ({precision}, multigrid)",
                value=1.0 / t, unit="iterations/s", ms_per_step=1e3 * t,
                mg_pcg_iterations=solves, steps=n, warmup=1, compliance=comp)


def _precision_pair(rec_mixed, rec_fp64):
    """Mixed and float64 records of one loop, with the rate ratio and the
    largest relative compliance difference over the iterations both ran."""
    cm, cf = rec_mixed["compliance"], rec_fp64["compliance"]
    m = min(len(cm), len(cf))
    err = max(abs(a - b) / abs(b) for a, b in zip(cm[:m], cf[:m]))
    log(f"{rec_mixed['config']}: mixed/fp64 {rec_mixed['value'] / rec_fp64['value']:.2f}x, "
        f"compliance rel err {err:.1e}")
    return dict(config=rec_mixed["config"], kernel=rec_mixed["kernel"].split(" (")[0],
                mixed=rec_mixed, fp64=rec_fp64, mixed_over_fp64=rec_mixed["value"] /
                rec_fp64["value"], compliance_rel_err_vs_fp64=err)


def apply_suite():
    """Rooflines of the operator apply on the other configurations and of
    the CSR filter gather on the tet mesh, each at uniform density 0.3, fp32
    operator (float64 filter), HIP events around back-to-back launches.
    Algorithmic bytes: apply k N (4 + 4) + n_el s_E [+ tet4 connectivity
    16 + gradients 12 * 4 + volume 4 per element]; CSR gather nnz (4 + 8)
    (column index, float64 weight) + n_design 8 (row pointer) + n_el 16
    (the source read once, the output written once)."""
    import torch
    from sktopt_amd import fea, filters
    from sktopt_amd.mesh import toy_problem
    out = []
    cases = [
        ("tet4-100x50x50x6-1load (c1)", lambda: toy_problem.cantilever(100, 50, 50, kind="tet4",
                                                                    jitter=0.2, seed=0)),
        ("hex8-200x100x100-1load (c3)", lambda: toy_problem.cantilever(200, 100, 100)),
        ("heat-hex8-160^3 (c4)", lambda: toy_problem.heat_patch(160, patch=16, source=1e-2)),
    ]
    for name, mk in cases:
        task = mk()
        cls = fea.FEM_SimpLinearHeatConduction if name.startswith("heat") else \
            fea.FEM_SimpLinearElasticity
        fem = cls(task, 1e-3 if name.startswith("heat") else 1e-9,
                  solver_config=fea.LinearSolverConfig(precision="mixed"))
        rho = torch.full((fem.n_el,), 0.3, dtype=torch.float64, device="cuda")
        fem._moduli(rho, 3.0)
        X = torch.randn((fem.ndof, fem.n_tasks), dtype=torch.float32, device="cuda")
        us = time_launches(lambda: fem.apply(X, dtype=torch.float32))
        b = fem.n_tasks * fem.ndof * 8 + fem.n_el * 4
        if task.mesh.kind == "tet4":
            b += fem.n_el * (16 + 12 * 4 + 4)
        gbs = b / (us * 1e-6) / 1e9
        out.append(dict(config=name, kernel="operator apply (fp32)", ndof=fem.ndof,
                        n_elements=fem.n_el, bytes=b, avg_us=us, achieved_gbs=gbs,
                        frac=gbs / HBM_PEAK_GBS, dof_per_s=fem.n_tasks * fem.ndof / (us * 1e-6)))
        log(f"apply {name}: {us:.1f} us, {gbs:.0f} GB/s")
        if task.mesh.kind == "tet4":
            f = filters.SpacialFilter.from_defaults(task.mesh, task.elements_volume, 1.5,
                                                    design_mask=task.design_mask)
            x = torch.rand(fem.n_el, dtype=torch.float64, device="cuda")
            y = torch.empty_like(x)
            us = time_launches(lambda: f.forward_into(x, y))
            nd, nnz = f.n_design, f.nnz
            b = nnz * (4 + 8) + nd * 8 + fem.n_el * 16
            gbs = b / (us * 1e-6) / 1e9
            out.append(dict(config=name, kernel="CSR filter gather (fp64, r=1.5)", nnz=nnz,
                            bytes=b, avg_us=us, achieved_gbs=gbs, frac=gbs / HBM_PEAK_GBS))
            log(f"filter {name}: nnz={nnz} {us:.1f} us, {gbs:.0f} GB/s")
            del f
        del fem
        torch.cuda.empty_cache()
        if task.mesh.kind == "tet4":
            out.append(_precision_pair(tet_oc_iterations(precision="mixed"),
                                       tet_oc_iterations(precision="fp64")))
        elif name.startswith("heat"):
            out.append(_precision_pair(heat_logmoc_iterations(precision="mixed"),
                                       heat_logmoc_iterations(precision="fp64")))
        torch.cuda.empty_cache()
    return out


def median(xs):
    ys = sorted(xs)
    m = len(ys) // 2
    return ys[m] if len(ys) % 2 else 0.5 * (ys[m - 1] + ys[m])


KERNEL_TRACE_FILE = os.path.join(ROOT, "profiles",
                                 "oc_loop_hex8_120x60x60_8load_mixed_mg_kernel_trace.txt")
# Algorithmic bytes per launch of the fine-level modal variants at c2 (fp32,
# 8 columns, n = ndof, e = elements): every variant reads x (and writes y)
# once per column; mode 1 also reads b; mode 2 reads b, the previous iterate
# and D^-1 (one value per dof, shared by the columns); mode 3 reads b and
# D^-1 and writes the second iterate.
MODAL_BYTES = {
    "skt::k_hex_apply_modal<float, 3, 2, 0, 0, true>": lambda n, e, k: k * n * 8 + e * 4,
    "skt::k_hex_apply_modal<float, 3, 2, 1, 0, false>": lambda n, e, k: k * n * 12 + e * 4,
    "skt::k_hex_apply_modal<float, 3, 2, 2, 0, true>": lambda n, e, k: k * n * 16 + n * 4 + e * 4,
    "skt::k_hex_apply_modal<float, 3, 2, 2, 0, false>": lambda n, e, k: k * n * 16 + n * 4 + e * 4,
    "skt::k_hex_apply_modal<float, 3, 2, 3, 0, false>": lambda n, e, k: k * n * 8 + n * 4 + e * 4,
}


def modal_variants_from_profile(ndof, n_el, k):
    """Average launch time of the fine-level modal variants from the
    committed kernel-trace summary (rocprofv3 --kernel-trace --stats of this
    benchmark), with their algorithmic bytes and HBM fraction."""
    out = []
    try:
        with open(KERNEL_TRACE_FILE) as f:
            lines = f.readlines()
    except OSError:
        return out
    seen = set()
    for line in lines:
        for name, fn in MODAL_BYTES.items():
            if line.startswith(name + " "):
                parts = line[len(name):].split()
                # rows of the stats table: calls avg min max per-OC-it pct
                if name in seen or len(parts) != 6:
                    continue
                seen.add(name)
                calls, avg_us = int(parts[0]), float(parts[1])
                b = fn(ndof, n_el, k)
                gbs = b / (avg_us * 1e-6) / 1e9
                out.append(dict(kernel=name, calls=calls, avg_us=avg_us, bytes=b,
                                achieved_gbs=gbs, frac=gbs / HBM_PEAK_GBS,
                                per_iter_ms=float(parts[4])))
    out.sort(key=lambda d: -d["per_iter_ms"])
    return out


def timing_trace_from_profile():
    """Trace durations (average per launch, microseconds) of the fine-level
    modal launches that the committed kernel-trace summary lists for the
    per-launch timing iterations (rows: kernel, launches, trace avg, events
    avg), the launches the HIP events of this benchmark time."""
    out = {}
    try:
        with open(KERNEL_TRACE_FILE) as f:
            lines = f.readlines()
    except OSError:
        return out
    for line in lines:
        for name in MODAL_BYTES:
            if line.startswith(name + " "):
                parts = line[len(name):].split()
                if len(parts) == 3 and name not in out:
                    out[name] = float(parts[1])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="hex8-120x60x60-8load", choices=sorted(WORKLOADS))
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="timed float64 iterations for the ratio (default: --steps)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-c0", action="store_true",
                    help="skip the fully timed CPU iteration of the 40x20x20 case (~100 s)")
    ap.add_argument("--no-c0", action="store_true",
                    help="skip the 40x20x20 GPU iterations (kernel traces of the c2 loop alone)")
    ap.add_argument("--apply-suite", action="store_true",
                    help="also time the operator apply on the tet, 200x100x100 and heat "
                         "configurations and the tet mesh's CSR filter gather")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU split is a design note only).", file=sys.stderr)
        sys.exit(2)
    import torch
    import sktopt_amd  # noqa: F401
    from sktopt_amd import _lib
    _lib.load()
    if not torch.cuda.is_available():
        print("no GPU", file=sys.stderr)
        sys.exit(1)
    wl = WORKLOADS[args.workload]
    K, W = args.steps, args.warmup
    copy = hbm_copy()
    log(f"hbm copy: {copy['gbs']:.0f} GB/s ({copy['us']:.1f} us for {copy['bytes']} B)")
    iters = W + K + 4
    t_setup = time.perf_counter()
    opt = make_optimizer(wl, "mixed", 1e-8, iters, "/tmp/skt_bench_mixed")
    log(f"setup {time.perf_counter() - t_setup:.1f}s  ndof={opt.fem.ndof} "
        f"n_el={opt.fem.n_el} loads={opt.fem.n_tasks}")
    times = timed_run(opt, W, K)
    t_med = median(times)
    t_mean = sum(times) / len(times)
    value = 1.0 / t_med
    comp_mixed = opt.recorder.as_object().compliance
    solves = opt.fem.solve_history[-K:]
    evals = opt.recorder.as_object().bisection_evals[-K:]
    log(f"mixed: {value:.3f} OC it/s (median {1e3 * t_med:.2f} ms, mean {1e3 * t_mean:.2f} ms)"
        f"  solves={[s['iters'] for s in solves]}")
    # the fine-level modal kernels (the CG's apply and the V-cycle's four
    # variants) from an untimed pass with HIP events around each launch; the
    # roofline is the one with the largest total time
    apply = apply_timing_pass(opt, 2)
    ndof, n_el, k = opt.fem.ndof, opt.fem.n_el, opt.fem.n_tasks
    bytes_apply = k * ndof * (4 + 4) + n_el * 4
    avg_ms = apply["f32_ms"] / max(apply["f32_n"], 1)
    achieved = bytes_apply / (avg_ms * 1e-3) / 1e9
    log(f"fp32 apply: {1e3 * avg_ms:.2f} us avg over {apply['f32_n']} launches, "
        f"{achieved:.0f} GB/s")
    trace_us = timing_trace_from_profile()
    fine = [dict(kernel=APPLY_KERNEL, total_ms=apply["f32_ms"], launches=apply["f32_n"])]
    for name, tms, tn in zip(VCYCLE_KERNELS, apply["vcycle_ms"], apply["vcycle_n"]):
        fine.append(dict(kernel=name, total_ms=tms, launches=tn))
    for d in fine:
        d["avg_us"] = 1e3 * d["total_ms"] / max(d["launches"], 1)
        d["bytes"] = MODAL_BYTES[d["kernel"]](ndof, n_el, k)
        d["achieved_gbs"] = d["bytes"] / (d["avg_us"] * 1e-6) / 1e9 if d["launches"] else None
        d["frac"] = d["achieved_gbs"] / HBM_PEAK_GBS if d["launches"] else None
        d["trace_avg_us"] = trace_us.get(d["kernel"])
        log(f"  {d['kernel']}: {d['avg_us']:.2f} us x {d['launches']}, "
            f"{d['total_ms']:.2f} ms in all")
    dom = max(fine, key=lambda d: d["total_ms"])
    jac_its = None if args.no_cpu else jacobi_iterations(opt)
    fp64 = None
    err = None
    if not args.no_fp64:
        del opt
        torch.cuda.empty_cache()
        K64 = args.fp64_steps or K
        opt64 = make_optimizer(wl, "fp64", 1e-8, W + K64 + 4, "/tmp/skt_bench_fp64")
        times64 = timed_run(opt64, W, K64)
        comp64 = opt64.recorder.as_object().compliance
        n = min(len(comp64), len(comp_mixed))
        err = max(abs(a - b) / abs(b) for a, b in zip(comp_mixed[:n], comp64[:n]))
        apply64 = apply_timing_pass(opt64, 2)
        t64 = median(times64)
        fp64 = dict(value=1.0 / t64, ms_per_step=1e3 * t64,
                    ms_per_step_mean=1e3 * sum(times64) / len(times64),
                    apply_avg_us=1e3 * apply64["f64_ms"] / max(apply64["f64_n"], 1),
                    mg_pcg_iterations=[s["iters"] for s in opt64.fem.solve_history[-K64 - 2:-2]])
        log(f"fp64: {fp64}  max compliance rel err mixed vs fp64 = {err:.2e}")
        del opt64
        torch.cuda.empty_cache()
    c0_gpu = None if args.no_c0 else gpu_c0_iterations()
    torch.cuda.empty_cache()
    cpu = None
    if not args.no_cpu:
        from oracle import cpu_baseline
        n_bis = sum(evals) / len(evals)
        cpu = cpu_baseline.measure(wl, jac_its, n_bis)
        log(f"cpu (extrapolated c2): {cpu['value']:.4f} OC it/s")
        if not args.no_cpu_c0 and c0_gpu is not None:
            cpu["c0_full_iteration"] = cpu_baseline.measure_c0_iteration(log)
            cs = cpu["c0_full_iteration"]["seconds"]
            c0_gpu["vs_cpu_first_iteration"] = cs / c0_gpu["first_iteration_s"]
            c0_gpu["vs_cpu_steady"] = cs / (1e-3 * c0_gpu["median_ms"])
            log(f"cpu c0 iteration: {cs:.1f} s (GPU first iteration "
                f"{c0_gpu['vs_cpu_first_iteration']:.0f}x faster, steady state "
                f"{c0_gpu['vs_cpu_steady']:.0f}x)")
    variants = modal_variants_from_profile(ndof, n_el, k)
    suite = apply_suite() if args.apply_suite else None
    line = {
        "metric": "OC iters/s, 1.36M-DOF hex8, GPU FP32-mixed vs FP64 at compliance err≤1e-5",
        "value": value, "unit": "OC iterations/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": 1e3 * t_med, "ms_per_step_mean": 1e3 * t_mean,
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-operator/f64-reductions",
        "data": "synthetic (seeded random point loads)",
        "config": {"workload": args.workload, "ndof": ndof, "n_elements": n_el,
                   "load_cases": k,
                   "solver": ("batched CG, geometric-multigrid preconditioner (Galerkin, "
                              "Chebyshev-Jacobi smoothing), fp32 with fp64 reliable updates"),
                   "rtol": 1e-8, "mg_pcg_iterations": [s["iters"] for s in solves],
                   "bisection_evals": evals, "timing": "median of per-iteration wall times"},
        "roofline": {"bound": "hbm", "kernel": dom["kernel"],
                     "achieved": dom["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["frac"], "traffic": traffic_from_profile(dom["kernel"]),
                     "bytes_per_launch": dom["bytes"], "avg_launch_us": dom["avg_us"],
                     "launches": dom["launches"], "hbm_copy_gbs": copy["gbs"],
                     "dominant": "largest total time among the fine-level modal kernels",
                     "measured": "HIP events per launch on the solve stream, 2 untimed OC "
                                 "iterations after the timed ones"},
        "operator_apply": {"kernel": APPLY_KERNEL, "avg_launch_us": 1e3 * avg_ms,
                           "achieved": achieved, "frac": achieved / HBM_PEAK_GBS,
                           "bytes_per_launch": bytes_apply,
                           "traffic": traffic_from_profile(APPLY_KERNEL)},
        "operator_apply_dof_per_s": k * ndof / (avg_ms * 1e-3),
        "fine_level_kernels": fine,
        "modal_variants": {"source": os.path.relpath(KERNEL_TRACE_FILE, ROOT),
                           "kernels": variants},
        "other_configs": suite,
        "fp64": fp64, "mixed_over_fp64": (value / fp64["value"]) if fp64 else None,
        "compliance_rel_err_vs_fp64": err,
        "cpu_baseline": cpu,
        "c0_gpu": c0_gpu,
        "host": platform.node(),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
