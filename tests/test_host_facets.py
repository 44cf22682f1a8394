"""Facet tables, facet selection and surface loads (host only):
``mesh/task_elastic.py:15-81``, ``mesh/task_common.py:273-347`` and the
reference's toy problems (``mesh/toy_problem.py:10-145``)."""
import numpy as np
import pytest

from oracle import facets as of
from sktopt_amd.mesh import jittered_tet, structured_hex, toy_problem as tp
from sktopt_amd.mesh.task import assemble_surface_forces


@pytest.mark.parametrize("grid", [(3, 2, 2), (5, 4, 3)])
def test_hex_facet_table_equals_brute_force(grid):
    nx, ny, nz = grid
    m = structured_hex(nx, ny, nz)
    got = {tuple(sorted(c)) for c in m.facets.T.tolist()}
    assert got == of.box_grid_faces(nx, ny, nz)
    assert m.facets.shape[1] == 3 * nx * ny * nz + nx * ny + ny * nz + nz * nx
    assert len(m.boundary_facets()) == 2 * (nx * ny + ny * nz + nz * nx)
    # cyclic order: every facet is a planar unit square of area 1
    np.testing.assert_allclose(m.facet_areas(), 1.0, rtol=0, atol=1e-14)


def test_tet_facets_and_areas():
    m = jittered_tet(2, 2, 2, jitter=0.15, seed=2)
    n_el = m.nelements
    f = m.facets
    nb = len(m.boundary_facets())
    # each tet has 4 faces; interior faces are shared by two tets
    assert 4 * n_el == 2 * f.shape[1] - nb
    # boundary triangles tile the surface of the 2x2x2 cube (boundary nodes
    # are not jittered): total area 6 * 4
    np.testing.assert_allclose(m.facet_areas(m.boundary_facets()).sum(), 24.0, rtol=1e-13)


def test_facet_selection_uses_midpoints():
    m = structured_hex(4, 3, 2)
    sel = m.facets_satisfying(lambda x: x[0] <= 0.03)
    assert len(sel) == 3 * 2                                  # the x = 0 face only
    assert np.all(m.p[0, m.facets[:, sel]] == 0.0)
    # a box that catches interior facets too, unless boundaries_only
    box = tp.get_points_in_range((2.5, 4.0), (0.0, 3.0), (0.0, 2.0))
    allf = m.facets_satisfying(box)
    bnd = m.facets_satisfying(box, boundaries_only=True)
    mid = m.p[:, m.facets].mean(axis=1)
    assert set(allf.tolist()) == set(np.nonzero(box(mid))[0].tolist())
    assert set(bnd.tolist()) < set(allf.tolist())
    el = m.elements_satisfying(lambda x: x[2] < 1.0)
    assert el.tolist() == list(range(12))


@pytest.mark.parametrize("kind", ["hex", "tet"])
def test_surface_forces_match_quadrature_oracle(kind):
    if kind == "hex":
        m = structured_hex(4, 3, 3, 0.5, 1.5, 2.0)
    else:
        m = jittered_tet(3, 3, 2, jitter=0.2, seed=4)
    rng = np.random.default_rng(1)
    regions = [rng.choice(m.facets.shape[1], size=9, replace=False) for _ in range(2)]
    F = assemble_surface_forces(m, regions, ["u^3", "u^1"], [-100.0, 7.5])
    for f, ids, comp, val in zip(F, regions, (2, 0), (-100.0, 7.5)):
        ref = of.surface_load(m.p, m.facets[:, ids], comp, val)
        np.testing.assert_allclose(f, ref, rtol=1e-13, atol=1e-13 * abs(val))
        np.testing.assert_allclose(f.sum(), val, rtol=1e-13)
        assert np.all(f.reshape(-1, 3)[:, [c for c in range(3) if c != comp]] == 0.0)


def test_bad_direction_raises():
    m = structured_hex(2, 2, 2)
    with pytest.raises(ValueError):
        assemble_surface_forces(m, np.arange(3), "x", 1.0)
    with pytest.raises(ValueError):
        assemble_surface_forces(m, [np.arange(3)], ["u^1", "u^2"], [1.0])


def test_toy_problems():
    t = tp.toy_test()
    m = t.mesh
    assert m.grid == (8, 6, 4)
    assert t.n_tasks == 1
    np.testing.assert_allclose(t.force[0].sum(), -100.0, rtol=1e-13)
    # clamped: every node on x = 0, all three components
    x0 = np.nonzero(m.p[0] == 0.0)[0]
    np.testing.assert_array_equal(t.dirichlet_nodes, x0)
    np.testing.assert_array_equal(t.dirichlet_dofs, np.sort(np.concatenate([3 * x0, 3 * x0 + 1,
                                                                            3 * x0 + 2])))
    # load facets: midpoints in [6.8, 8.1] x [2.4, 3.6] x [2.8, 4.0]
    mid = m.p[:, m.facets].mean(axis=1)
    box = tp.get_points_in_range((6.8, 8.1), (2.4, 3.6), (2.8, 4.0))
    ref = of.surface_load(m.p, m.facets[:, np.nonzero(box(mid))[0]], 2, -100.0)
    np.testing.assert_allclose(t.force[0], ref, rtol=1e-13, atol=1e-12)
    t2 = tp.toy2()
    assert t2.n_tasks == 2
    np.testing.assert_allclose([f.sum() for f in t2.neumann_linear], [-1.0, 1.0], rtol=1e-13)
    assert np.all(t2.mesh.p[0, t2.neumann_nodes] == 8.0)
