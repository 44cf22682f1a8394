"""The benchmark's readers of the committed profile summaries: every
fine-level modal kernel it times has a kernel-trace row, a trace row of the
timed launches and a traffic line (no GPU needed)."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_profile_summaries_parse():
    ndof, n_el, k = 1350723, 432000, 8
    variants = bench.modal_variants_from_profile(ndof, n_el, k)
    assert sorted(d["kernel"] for d in variants) == sorted(bench.MODAL_BYTES)
    for d in variants:
        assert d["calls"] > 0 and d["avg_us"] > 0 and d["per_iter_ms"] > 0
        assert 0 < d["frac"] < 1
    trace = bench.timing_trace_from_profile()
    assert sorted(trace) == sorted(bench.MODAL_BYTES)
    assert all(v > 0 for v in trace.values())
    for name in bench.MODAL_BYTES:
        t = bench.traffic_from_profile(name)
        assert t is not None and t > bench.MODAL_BYTES[name](ndof, n_el, k) * 0.5, name
    assert bench.APPLY_KERNEL in bench.MODAL_BYTES
    assert set(bench.VCYCLE_KERNELS) <= set(bench.MODAL_BYTES)


def test_mode3_bytes():
    # 8 columns x 1,350,723 dofs x (4 B read of b + 4 B write) + 1,350,723 x 4 B
    # of D^-1 + 432,000 x 4 B of moduli
    assert bench.MODAL_BYTES[bench.VCYCLE_KERNELS[0]](1350723, 432000, 8) == 93577164
