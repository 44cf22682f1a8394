"""The allocation-free percentile (skt_percentile_abs_noalloc) against the
allocating entry point it mirrors (skt_percentile_abs) and numpy: two calls
in a row, the second on a smaller input, so the static scratch's leftovers
from the first call must not leak into the second."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n) * 10.0 ** rng.integers(-5, 5, n)
    x[: n // 3] = np.round(x[: n // 3], 1)        # ties
    return x


@pytest.mark.parametrize("q", [0.0, 3.3, 50.0, 95.0, 100.0])
def test_noalloc_percentile_matches_allocating_entry_point(gpu, q):
    import torch
    from sktopt_amd import _lib
    s = torch.cuda.current_stream().cuda_stream
    for n, seed in ((1_500_001, 1), (4097, 2), (1, 3)):
        x = _data(n, seed) if n > 1 else np.array([-2.5])
        xd = torch.as_tensor(x, device="cuda")
        a = torch.zeros(1, dtype=torch.float64, device="cuda")
        b = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
        _lib.call("skt_percentile_abs", n, xd.data_ptr(), q, a.data_ptr(), s)
        _lib.call("skt_percentile_abs_noalloc", n, xd.data_ptr(), q, b.data_ptr(), s)
        torch.cuda.synchronize()
        assert b.item() == a.item(), (n, q)
        assert b.item() == np.percentile(np.abs(x), q), (n, q)


def test_noalloc_percentile_rejects_bad_sizes(gpu):
    """Host-side checks before any launch: empty input, n beyond 2^53, q
    outside [0, 100]."""
    import torch
    from sktopt_amd import _lib
    lib = _lib.load()
    s = torch.cuda.current_stream().cuda_stream
    x = torch.ones(4, dtype=torch.float64, device="cuda")
    o = torch.zeros(1, dtype=torch.float64, device="cuda")
    assert lib.skt_percentile_abs_noalloc(0, x.data_ptr(), 50.0, o.data_ptr(), s) == \
        _lib.SKT_EINVAL
    assert lib.skt_percentile_abs_noalloc((1 << 53) + 1, x.data_ptr(), 50.0, o.data_ptr(),
                                          s) == _lib.SKT_EINVAL
    assert "2^53" in _lib.last_error()
    assert lib.skt_percentile_abs_noalloc(4, x.data_ptr(), 101.0, o.data_ptr(), s) == \
        _lib.SKT_EINVAL
