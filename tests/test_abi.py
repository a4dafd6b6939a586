"""CPU checks of the C ABI: the library loads, exports every symbol the header
declares, validates arguments on the host, and the Python layer refuses CPU
tensors instead of falling back."""
import ctypes

import pytest
import torch

from augur_amd import _lib, ops


def test_exports_every_header_symbol(lib):
    names = _lib.header_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)
    assert lib.augur_abi_version() == 1


def test_planned_deposit_drift_word_is_in_its_workspace(lib):
    n = _lib.size("augur_cic_deposit_ragged_planned_workspace_bytes", 4, 64, 10000)
    assert n >= _lib.size("augur_cic_deposit_ragged_tiled_workspace_bytes", 4, 64, 10000)
    base = 1 << 20
    w = lib.augur_ragged_plan_drift_word(base, n)
    assert w is not None and base <= w < base + 4096
    assert lib.augur_ragged_plan_drift_word(base, 8) is None


def test_workspace_queries_are_host_only(lib):
    assert _lib.size("augur_sort_workspace_bytes", 1 << 20, 1 << 18) >= 8 << 20
    assert _lib.size("augur_cic_deposit_workspace_bytes", 2, 3, 64, 4) >= 2 * 4 * 64 ** 3 * 8
    assert _lib.size("augur_pm_workspace_bytes", 1 << 16, 64) >= 64 ** 3 * 8
    assert _lib.size("augur_direct_sum_workspace_bytes", 4096, 512, 3) > 0
    assert _lib.size("augur_ewald_workspace_bytes", 4096, 64, 8) > 0


@pytest.mark.parametrize("name,args,needle", [
    ("augur_cic_deposit", (None, None, None, 1.0, None, 1, 10, 4, 8, 0.0, 1.0, 1, 1.0, None, None, 0, None), "D must"),
    ("augur_cic_deposit", (None, None, None, 1.0, None, 1, 10, 3, 12, 0.0, 1.0, 1, 1.0, None, None, 0, None),
     "power-of-two"),
    ("augur_cell_keys", (None, None, 1, 4, 5, 0.0, 1.0, 4, None, None, None), "D must"),
    ("augur_pm_kdk", (None, None, None, None, None, 1.0, 100, 48, 1.0, 1.0, 0.1, 1, None, 0, None), "power of two"),
    ("augur_pm_kdk", (None, None, None, None, None, 1.0, 100, 16, 1.0, 1.0, 0.1, 1, None, 0, None), "32, 512"),
    ("augur_conv3_bf16_fwd", (None, 8, None, None, 1, 64, 0, 3, None, None, None, None, None), "bad arguments"),
    ("augur_conv3_bf16_fwd", (None, 16, None, None, 1, 48, 0, 3, None, None, None, None, None), "power of two"),
    ("augur_poisson_green_grad", (None, 1, 1024, 1.0, 1.0, 1.0, None, None, None), "power-of-two"),
    ("augur_direct_sum_f64", (None, None, None, 10, 3, None, 4, 0.1, None, None, 0, None), "exactly one"),
    ("augur_persistent_step", (1, 1, 1, 1, 0, 10, 1, 32, 1.0, 0.1, 0, None), "B must be"),
    ("augur_persistent_step_tracked", (1, 1, 1, 1, 2, 10, 1, 32, 1.0, 0.1, 3, None, None), "null drift word"),
    ("augur_cic_deposit_ragged_planned", (1, None, 1.0, 1, 2, 10, 48, 0.1, 1.0, 0.1, 1, 1, None, 0, None),
     "power of two"),
    ("augur_cic_deposit_ragged_planned", (1, None, 1.0, 1, 2, 10, 32, 0.1, 1.0, -1.0, 1, 1, None, 0, None),
     "drift_dt"),
    ("augur_persistent_step", (1, 1, 1, 1, 2, 10, 1, 48, 1.0, 0.1, 0, None), "power of two"),
    ("augur_persistent_step", (1, 1, 1, 1, 2, 10, 1, 32, 1.0, 0.1, 7, None), "unknown phase"),
    ("augur_persistent_step", (1, None, 1, 1, 2, 10, 1, 32, 1.0, 0.1, 3, None), "vel is required"),
    ("augur_persistent_step", (1, 1, 1, 1, 2, 10, None, 32, 1.0, 0.1, 2, None), "force is required"),
    ("augur_conv3_bf16_fwd_tile", (1, 32, 1, 1, 1, 64, 1, 32, None, 1, None, None, 1, 24, None), "zc must"),
    ("augur_conv3_bf16_fwd_tile", (1, 16, 1, 1, 1, 64, 1, 32, None, 1, None, None, 2, 16, None), "tile variant"),
])
def test_host_validation_errors(lib, name, args, needle):
    with pytest.raises(_lib.AugurError, match=needle):
        _lib.call(name, *args)


def test_fft_rejects_non_power_of_two(lib):
    shape = (ctypes.c_int64 * 2)(12, 16)
    with pytest.raises(_lib.AugurError, match="powers of two"):
        _lib.call("augur_fft_r2c", None, None, None, 2, shape, 1, None, 0, None)


def test_ops_refuse_cpu_tensors():
    pos = torch.rand(1, 8, 3)
    with pytest.raises(RuntimeError, match="ROCm GPU only"):
        ops.cic_deposit(pos, None, 8, 0.125, 0.0, True)
    with pytest.raises(RuntimeError, match="ROCm GPU only"):
        ops.direct_sum_f64(pos[0], 0.1)


def test_conv_tile_candidates_host_only(lib):
    """The autotuner's candidate tiles come from the library's own rule (no GPU needed)."""
    from augur_amd import autotune as A
    assert A.variants(16, 1, 32) == [A.TILE_ZM]
    assert A.variants(32, 2, 32) == [A.TILE_ZM, A.TILE_SPLIT]
    assert A.variants(32, 0, 3) == [A.TILE_ZM, A.TILE_NARROW]
    assert A.variants(32, 3, 32) == [A.TILE_ZM, A.TILE_SPLIT]
    assert A.variants(32, 3, 16) == [A.TILE_ZM, A.TILE_NARROW]
    assert A.variants(16, 2, 32) == [] and A.variants(8, 1, 32) == [] and A.variants(32, 1, 33) == []
    assert A.variants(32, 4, 32) == [A.TILE_SPLIT] and A.variants(32, 4, 16) == [] and A.variants(16, 4, 32) == []
    assert A.variants(4, 1, 32) == [A.TILE_IN4] and A.variants(4, 2, 32) == [] and A.variants(4, 0, 3) == []
    assert A.candidates(32, 2, 32, 32) == [(1, 8), (1, 16), (1, 32), (2, 8), (2, 16), (2, 32)]
    assert len(A.candidates(32, 0, 3, 256)) == 10
    t = A.ConvTuner()
    assert t.fwd_tile(32, 2, 32, 1, 256) is None and t.wgrad_nwg(32, 1, 256) is None
