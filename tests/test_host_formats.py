"""On-disk formats of the density loop (host only): ``histories.npz``
(tools/history.py:125-206,536-700), ``cfg.json``
(core/optimizers/common_density.py:312-324) and VTU snapshots
(core/visualization.py:22-84, mesh/visualization.py:11-37)."""
import xml.etree.ElementTree as ET

import numpy as np

from sktopt_amd.core import visualization as vis
from sktopt_amd.mesh import structured_hex, jittered_tet
from sktopt_amd.tools import SchedulerConfig
from sktopt_amd.tools.history import HistoryCollection


def test_history_semantics_and_npz_round_trip(tmp_path):
    h = HistoryCollection(str(tmp_path))
    h.add("rho_projected", plot_type="min-max-mean-std")
    h.add("u_max", plot_type="min-max-mean")
    h.add("compliance", ylog=True)
    h.add("never_fed")
    rng = np.random.default_rng(0)
    xs = [rng.uniform(size=50) for _ in range(3)]
    for i, x in enumerate(xs):
        h.feed_data("rho_projected", x)
        h.feed_data("u_max", x[:5])
        h.feed_data("compliance", np.float64(10.0 + i))      # 0-d array -> float
    # stats entries in the reference's order [min, mean, max, std]
    np.testing.assert_array_equal(h.latest("rho_projected"),
                                  [xs[2].min(), xs[2].mean(), xs[2].max(), xs[2].std()])
    np.testing.assert_array_equal(h.latest("u_max"),
                                  [xs[2][:5].min(), xs[2][:5].mean(), xs[2][:5].max()])
    h.export_histories()
    with np.load(tmp_path / "histories.npz", allow_pickle=False) as d:
        assert set(d.files) == {"rho_projected", "rho_projected_header", "u_max",
                                "u_max_header", "compliance", "compliance_header"}
        assert d["rho_projected"].shape == (4, 3)        # (n_stats, n_iters)
        assert d["u_max"].shape == (3, 3)
        np.testing.assert_array_equal(d["compliance"], [10.0, 11.0, 12.0])
        assert d["rho_projected_header"].tolist() == ["rho_projected", "min-max-mean-std"]
        np.testing.assert_array_equal(d["rho_projected"][:, 1],
                                      [xs[1].min(), xs[1].mean(), xs[1].max(), xs[1].std()])
    g = HistoryCollection(str(tmp_path))
    g.import_histories()
    assert g.histories["rho_projected"].plot_type == "min-max-mean-std"
    np.testing.assert_array_equal(g.latest("rho_projected"), h.latest("rho_projected"))
    assert g.as_object().compliance == [10.0, 11.0, 12.0]


def test_device_stats_entry_equals_array_entry(tmp_path):
    h = HistoryCollection(str(tmp_path))
    h.add("a", plot_type="min-max-mean-std")
    h.add("b", plot_type="min-max-mean-std")
    x = np.random.default_rng(1).standard_normal(40)
    h.feed_data("a", x)
    h.feed_stats("b", [x.min(), x.mean(), x.max(), x.std()])
    assert h.histories["a"].data == h.histories["b"].data


def test_cfg_json_round_trip(tmp_path):
    from sktopt_amd import core
    cfg = core.OC_Config(dst_path=str(tmp_path), max_iters=7,
                         vol_frac=SchedulerConfig.constant(target_value=0.25),
                         filter_radius=SchedulerConfig.constant(target_value=1.75),
                         precision="fp64")
    cfg.export(str(tmp_path))
    back = core.OC_Config.import_from(str(tmp_path))
    assert back.max_iters == 7 and back.precision == "fp64"
    assert isinstance(back.vol_frac, SchedulerConfig)
    assert back.vol_frac.target_value == 0.25
    assert back.move_limit == cfg.move_limit
    assert back.p == cfg.p


def _vtk_hex_volume(p, c):
    """Volume of a VTK_HEXAHEDRON from its 8 points (VTK corner order), by
    the divergence-free 6-tet split around the 0-6 diagonal."""
    tets = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
    v = 0.0
    for a, b, cc, d in tets:
        v += np.linalg.det(np.stack([p[:, c[b]] - p[:, c[a]], p[:, c[cc]] - p[:, c[a]],
                                     p[:, c[d]] - p[:, c[a]]])) / 6.0
    return v


def test_vtu_write_read_hex_and_tet(tmp_path):
    m = structured_hex(4, 3, 2, 0.5, 1.0, 2.0)
    rho = np.linspace(0, 1, m.nelements)
    nd = np.arange(m.nvertices, dtype=np.float64)
    f = str(tmp_path / "mesh_rho" / "info_mesh-00000001.vtu")
    vis.export_mesh_with_info(m, point_data_values=[nd], point_data_names=["id"],
                              cell_data_values=[rho], cell_data_names=["rho_projected"],
                              filepath=f)
    root = ET.parse(f).getroot()
    assert root.get("type") == "UnstructuredGrid"
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfCells")) == m.nelements
    assert int(piece.get("NumberOfPoints")) == m.nvertices
    p, c, types, pd, cd = vis.read_vtu(f)
    np.testing.assert_array_equal(p, m.p)
    np.testing.assert_array_equal(c, m.t)
    assert set(types.tolist()) == {12}
    np.testing.assert_array_equal(cd["rho_projected"], rho)
    np.testing.assert_array_equal(pd["id"], nd)
    # the corner order is VTK's: every cell has positive volume hx*hy*hz
    for e in range(m.nelements):
        assert abs(_vtk_hex_volume(p, c[:, e]) - 1.0) < 1e-12
    t = jittered_tet(2, 2, 2, jitter=0.1, seed=1)
    g = str(tmp_path / "tet.vtu")
    vis.export_mesh_with_info(t, filepath=g)
    p2, c2, types2, _, _ = vis.read_vtu(g)
    np.testing.assert_array_equal(c2, t.t)
    assert set(types2.tolist()) == {10}


def test_submesh_matches_reference_rule(tmp_path):
    from sktopt_amd.mesh import toy_problem
    tsk = toy_problem.cantilever(6, 3, 3)
    rng = np.random.default_rng(3)
    rho = rng.uniform(size=tsk.mesh.nelements)
    p, t = vis.export_submesh(tsk, rho, 0.5, str(tmp_path / "cubic_top.vtu"))
    # the reference's rule written out with a dict map (mesh/visualization.py:17-33)
    design = np.asarray(tsk.design_elements)
    remove = design[rho[design] <= 0.5]
    keep = np.array([e for e in range(tsk.mesh.nelements) if e not in set(remove.tolist())])
    kept_t = tsk.mesh.t[:, keep]
    uniq = np.unique(kept_t)
    index_map = {old: new for new, old in enumerate(uniq)}
    ref_t = np.vectorize(index_map.get)(kept_t)
    np.testing.assert_array_equal(t, ref_t)
    np.testing.assert_array_equal(p, tsk.mesh.p[:, uniq])
    p2, c2, _, _, _ = vis.read_vtu(str(tmp_path / "cubic_top.vtu"))
    np.testing.assert_array_equal(c2, ref_t)
