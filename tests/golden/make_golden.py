"""Write the committed golden trajectories of the oracle (FP64).

c0: 3D cantilever 40x20x20 unit hexes (54,243 DOF), x=0 clamped, unit -z
load split over the 21 nodes of the bottom edge x=40, z=0; E0=1, Emin=1e-9,
nu=0.3, SIMP p=3, vol_frac 0.3 (initial 0.3), Gaussian filter r=1.5,
beta=1 (eta 0.5), OC move limit 0.2, eta 0.5; state solve FP64 Jacobi-CG to
rtol 1e-12 (agrees with SuperLU to ~1e-11 relative compliance, see
tests/test_oracle_fem.py).

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import driver, problems  # noqa: E402


def c0_config(iters=50):
    S = driver.Sched
    return driver.Config(max_iters=iters, p=S("Constant", 3.0, 3.0),
                         vol_frac=S("Constant", 0.3, 0.3), beta=S("Constant", 1.0, 1.0),
                         filter_radius=S("Constant", 1.5, 1.5),
                         move_limit=S("Constant", 0.2, 0.2), E_min_coeff=1e-9,
                         solver="cg_jacobi", solver_rtol=1e-12)


def main(n_iter=3):
    task = problems.cantilever(40, 20, 20)
    loop = driver.DensityLoop(c0_config(), task)
    recs = [loop.step(it) for it in range(1, n_iter + 1)]
    out = dict(
        compliance=np.array([r["compliance"] for r in recs]),
        lmid=np.array([r["lmid"] for r in recs]),
        vol_error=np.array([r["vol_error"] for r in recs]),
        n_evals=np.array([r["n_evals"] for r in recs]),
        kkt=np.array([r["kkt_residual"] for r in recs]),
        rho_change_max=np.array([r["rho_change_max"] for r in recs]),
        rho_final=loop.rho[task.design_elements],
        design_elements=task.design_elements.astype(np.int32),
    )
    np.savez_compressed(os.path.join(HERE, "c0_oc_3iter.npz"), **out)
    print({k: v for k, v in out.items() if v.size < 10})


if __name__ == "__main__":
    main()
