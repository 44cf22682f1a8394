"""FP64 CPU oracle for the density-method loop -- TEST INFRASTRUCTURE ONLY.

This package is a from-scratch numpy/scipy restatement of the per-iteration
density-method loop of scikit-topt v0.3.9 (filter -> Heaviside projection ->
SIMP/RAMP interpolation -> state solve -> element energy -> compliance
sensitivity -> filter adjoint -> OC bisection / LogMOC update).  Each function
cites the reference file:line it restates (paths relative to the reference
repository root, ``scikit-topt/sktopt/...``).

It exists to CHECK the HIP implementation in ``scikit-topt_amd``.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU
baseline.  The product path never imports, links or falls back to it.

Pinning (what fixes each function independently of the reference, which
cannot be imported here: scikit-fem, pyamg and numba are absent):

* ``elements``   exact symbolic integration (sympy), rigid-body null space,
                 constant-strain patch test, closed-form bending-mode energy.
* ``fem``        dense brute-force solves on tiny meshes, the uniaxial-bar
                 closed form, ``C = f.u`` and ``sum_e U_e = C/2`` identities.
* ``filter``     brute-force O(n^2) neighbour search, row sums, adjoint
                 identity <Wx, y> = <x, W^T y>.
* ``transforms`` closed forms incl. the ``+1e-12`` and clamp quirks,
                 central finite differences.
* ``oc``         brute-force pure-Python loop restatement on small inputs,
                 exit-condition and bound invariants, monotone volume in lambda.
* ``logmoc``     hand-evaluated single steps on tiny vectors.
* ``scheduler``  the reference's own test properties
                 (``tests/test_scheduler.py:16-114``) plus closed forms.
* ``driver``     full-vs-stepwise bit-identical runs
                 (``tests/test_optimize_steps_equivalence.py:49-52``).

No reference test holds golden numbers for this path (its tests only assert
finiteness, ``tests/test_global_flow.py:145,151,157``), so no function is
pinned by a reference golden vector; the checks above stand in for them.
"""
