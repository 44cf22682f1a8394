"""Timed CPU baseline for the benchmark (oracle = the reference algorithm
restated in numpy/scipy).  TEST/BENCH INFRASTRUCTURE ONLY.

The reference's own path cannot run at the benchmark size (its multi-load
solve is a 1.35M-DOF SuperLU factorisation, ``fea/solver_elastic.py:384-390``,
which does not fit in host memory), so the per-iteration cost is measured on
a bounded sample and scaled:

* sample = the same x/y extent and loads but only ``nz_sample`` element
  layers (times scale linearly with the element count);
* assembly of K (``fea/composer.py:53-101`` restated), one Jacobi-CG
  iteration of the ``cg_jacobi`` substitution (``solver_elastic.py:84-92``,
  SpMV + vector updates, one load column), one filter-matrix build as the
  reference does it on every ``forward`` (cKDTree + Python loop,
  ``filters/spacial.py:51-103,143-149``) and one cached filter SpMV;
* per OC iteration = assembly + n_loads * cg_iters * t_cg_iter
  + (1 + n_bisect) * t_filter [+ n_loads adjoint SpMVs], with cg_iters the
  Jacobi-CG iteration count of the same system (counted by the caller with a
  Jacobi-preconditioned solve of the GPU state) and n_bisect the OC
  bisection's evaluation count.
"""
from __future__ import annotations

import os
import time

import numpy as np
import scipy.sparse.linalg as spla
from threadpoolctl import threadpool_limits

from . import fem as ofem
from . import filter as ofilt
from . import problems as oprob
from . import transforms as otr


def measure(wl: dict, gpu_solves: list, n_bisect: float = 38.0, nz_sample: int = 6) -> dict:
    """gpu_solves: [{"iters": Jacobi-CG iterations per solve}, ...]."""
    with threadpool_limits(1):
        return _measure(wl, gpu_solves, n_bisect, nz_sample)


def _measure(wl, gpu_solves, n_bisect, nz_sample):
    nx, ny, nz = wl["grid"]
    k = wl["n_loads"]
    frac = nz / nz_sample
    task = oprob.multiload(nx, ny, nz_sample, n_loads=k, seed=0)
    rho = np.full(task.n_elements, wl["vol"])
    op = task.element_operator()
    t = time.perf_counter()
    E = otr.simp_interpolation(rho, 1.0, 1e-9, 3.0)
    K = op.assemble(E)
    Ke, _ = ofem.enforce(K, np.zeros(task.ndof), task.dirichlet_dofs)
    t_asm = time.perf_counter() - t
    f = task.forces[0].copy()
    f[task.dirichlet_dofs] = 0.0
    m_inv = 1.0 / Ke.diagonal()
    M = spla.LinearOperator(Ke.shape, matvec=lambda x: m_inv * x)
    n_cg = 40
    t = time.perf_counter()
    spla.cg(Ke, f, M=M, rtol=1e-30, maxiter=n_cg)
    t_cg = (time.perf_counter() - t) / n_cg
    mask = task.design_mask
    t = time.perf_counter()
    W = ofilt.build_filter_matrix(task.centers, wl["r"], mask)
    t_build = time.perf_counter() - t
    x = np.random.default_rng(0).uniform(0, 1, int(mask.sum()))
    t = time.perf_counter()
    for _ in range(10):
        W @ x
    t_spmv = (time.perf_counter() - t) / 10
    iters = [s.get("iters", 0) for s in gpu_solves] or [1000]
    cg_iters = float(np.mean(iters))
    per_iter_written = frac * (t_asm + k * cg_iters * t_cg + (1 + n_bisect) * t_build
                               + k * t_spmv)
    per_iter_cached = frac * (t_asm + k * cg_iters * t_cg + (1 + n_bisect + k) * t_spmv)
    return {
        "value": 1.0 / per_iter_written, "unit": "OC iterations/s", "cores": 1,
        "kind": "port",
        "value_filter_cached": 1.0 / per_iter_cached,
        "sample": (f"{nx}x{ny}x{nz_sample} slab of the {nx}x{ny}x{nz} workload, scaled x{frac:.2f}:"
                   f" assembly {t_asm:.2f}s, Jacobi-CG iteration {1e3 * t_cg:.1f}ms x {k} loads x"
                   f" {cg_iters:.0f} iterations, filter build (as in the reference, per forward)"
                   f" {t_build:.2f}s x {1 + n_bisect:.0f}, cached filter SpMV {1e3 * t_spmv:.2f}ms;"
                   f" 1 thread (threadpoolctl) on a host with {os.cpu_count()} cpus"),
    }
