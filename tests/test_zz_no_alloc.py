"""The no-allocation FP64 direct sum (augur_direct_sum_f64_noalloc: its
partial sums in the library's static device scratch) against the workspace
entry point augur_direct_sum_f64 on the same inputs: the same bits, on two
calls in a row (the second smaller, so stale partials of the first would
show), in 3D and 2D; sizes past the scratch are refused on the host."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return torch.device("cuda:0")


def _noalloc():
    from augur_amd import _lib
    return _lib.get().augur_direct_sum_f64_noalloc     # ctypes signature from augur_amd._lib


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def test_noalloc_direct_sum_equals_workspace_entry(dev):
    from augur_amd import ops
    f = _noalloc()
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    for n, t, D, seed in ((60000, 4096, 3, 1), (9000, 700, 3, 2), (5000, 333, 2, 3)):
        print(f"noalloc direct sum: n={n} t={t} D={D}", flush=True)
        rng = np.random.default_rng(seed)
        pos = torch.from_numpy(rng.random((n, D), dtype=np.float32)).to(dev)
        mass = torch.from_numpy(rng.uniform(0.5, 1.5, n)).to(dev)
        tg = torch.from_numpy(np.sort(rng.choice(n, t, replace=False)).astype(np.int64)).to(dev)
        ref = ops.direct_sum_f64(pos, 0.01, mass, tg)
        out = torch.full((t, D), float("nan"), dtype=torch.float64, device=dev)
        rc = f(None, _p(pos), _p(mass), n, D, _p(tg), t, 0.01, _p(out), stream)
        assert rc == 0
        torch.cuda.synchronize()
        assert torch.equal(out, ref)


def test_noalloc_refuses_sizes_past_its_scratch(dev):
    from augur_amd import _lib
    f = _noalloc()
    pos = torch.zeros(16, 3, device=dev)
    out = torch.zeros(1, dtype=torch.float64, device=dev)
    rc = f(None, _p(pos), None, 16, 3, None, 700000, 0.01, _p(out), None)
    assert rc != 0 and b"static scratch" in _lib.get().augur_last_error()
