"""Write the committed golden trajectory of the oracle (FP64, SuperLU).

c0: 3D cantilever 40x20x20 unit hexes (54,243 DOF), x=0 clamped, unit -z
load split over the 21 nodes of the bottom edge x=40, z=0; E0=1, Emin=1e-9,
nu=0.3, SIMP p=3, vol_frac 0.3 (initial 0.3), Gaussian filter r=1.5,
beta=1 (eta 0.5), OC move limit 0.2, eta 0.5, 50 iterations; state solve
``scipy.sparse.linalg.spsolve`` (SuperLU), the reference's default
``solver_option`` (``core/optimizers/common_density.py:253-255``,
``fea/solver_elastic.py:106-108``).

Stored per iteration: compliance, lmid, vol_error, n_evals (bisection
evaluations), kkt residual, rho_change_max, running_scale; the design
densities after iterations 10, 25 and 50.  One SuperLU factorisation of
this 3D system takes ~1-2 minutes on one core, so the run takes about an
hour.

Run: python tests/golden/make_golden.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import driver, problems  # noqa: E402

GOLDEN_NAME = "c0_oc_50iter.npz"
RHO_AT = (10, 25, 50)


def c0_config(iters=50, solver="spsolve", rtol=1e-12):
    S = driver.Sched
    return driver.Config(max_iters=iters, p=S("Constant", 3.0, 3.0),
                         vol_frac=S("Constant", 0.3, 0.3), beta=S("Constant", 1.0, 1.0),
                         filter_radius=S("Constant", 1.5, 1.5),
                         move_limit=S("Constant", 0.2, 0.2), E_min_coeff=1e-9,
                         solver=solver, solver_rtol=rtol)


def main(n_iter=50):
    task = problems.cantilever(40, 20, 20)
    loop = driver.DensityLoop(c0_config(n_iter), task)
    recs = []
    rho_at = {}
    for it in range(1, n_iter + 1):
        t = time.perf_counter()
        recs.append(loop.step(it))
        if it in RHO_AT:
            rho_at[f"rho_{it}"] = loop.rho[task.design_elements].copy()
        print(f"iter {it}: compliance {recs[-1]['compliance']:.15e} "
              f"evals {recs[-1]['n_evals']} ({time.perf_counter() - t:.1f}s)", flush=True)
    out = dict(
        compliance=np.array([r["compliance"] for r in recs]),
        lmid=np.array([r["lmid"] for r in recs]),
        vol_error=np.array([r["vol_error"] for r in recs]),
        n_evals=np.array([r["n_evals"] for r in recs]),
        kkt=np.array([r["kkt_residual"] for r in recs]),
        rho_change_max=np.array([r["rho_change_max"] for r in recs]),
        running_scale=np.array([r["running_scale"] for r in recs]),
        design_elements=task.design_elements.astype(np.int32),
        **rho_at,
    )
    np.savez_compressed(os.path.join(HERE, GOLDEN_NAME), **out)
    print({k: v for k, v in out.items() if v.size <= 50})


if __name__ == "__main__":
    main()
