"""ctypes binding of libaugur_hip.so (the C ABI declared in include/augur.h).

The library is the only compute path: if it is missing, importing the ops
raises instead of falling back to anything else.
"""
import ctypes
import os
import re
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaugur_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "augur.h")

c_int, c_int64, c_float, c_double, c_size_t, c_void_p, c_char_p = (
    ctypes.c_int, ctypes.c_int64, ctypes.c_float, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_char_p)
P = c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "augur_last_error": (c_char_p, []),
    "augur_abi_version": (c_int, []),
    "augur_cell_keys": (c_int, [P, P, c_int64, c_int64, c_int, c_float, c_float, c_int64, P, P, P]),
    "augur_cic_base": (c_int, [P, c_int64, c_int, c_float, c_float, P, P, P]),
    "augur_sort_workspace_bytes": (c_size_t, [c_int64, c_int64]),
    "augur_sort_by_cell": (c_int, [P, c_int64, c_int64, P, P, P, c_size_t, P]),
    "augur_cic_deposit_workspace_bytes": (c_size_t, [c_int64, c_int, c_int64, c_int]),
    "augur_cic_deposit": (c_int, [P, P, P, c_float, P, c_int64, c_int64, c_int, c_int64, c_float, c_float, c_int,
                                  c_float, P, P, c_size_t, P]),
    "augur_cic_deposit_ragged": (c_int, [P, P, P, c_float, P, c_int64, c_int64, c_int, c_int64, c_float, c_float,
                                         c_int, c_float, P, P, c_size_t, P]),
    "augur_tsc_deposit": (c_int, [P, P, P, c_float, P, c_int64, c_int64, c_int, c_int64, c_float, c_float, c_int,
                                  c_float, P, P, c_size_t, P]),
    "augur_tsc_gather": (c_int, [P, P, c_int64, c_int64, c_int, c_int64, c_int, c_float, c_float, c_int, P, P]),
    "augur_cic_gather": (c_int, [P, P, c_int64, c_int64, c_int, c_int64, c_int, c_float, c_float, c_int, P, P]),
    "augur_fft_workspace_bytes": (c_size_t, [c_int, P, c_int64]),
    "augur_fft_r2c": (c_int, [P, P, P, c_int, P, c_int64, P, c_size_t, P]),
    "augur_fft_c2r": (c_int, [P, P, P, c_int, P, c_int64, c_float, P, c_size_t, P]),
    "augur_complex_mul": (c_int, [P, P, P, c_int64, c_int64, c_float, P]),
    "augur_poisson_green_grad": (c_int, [P, c_int64, c_int64, c_float, c_float, c_float, P, P, P]),
    "augur_open_kernel_2d": (c_int, [c_int64, c_float, c_float, P, P]),
    "augur_neg_grad_fd": (c_int, [P, c_int64, c_int64, c_float, P, P]),
    "augur_direct_sum_workspace_bytes": (c_size_t, [c_int64, c_int64, c_int]),
    "augur_direct_sum_f64": (c_int, [P, P, P, c_int64, c_int, P, c_int64, c_double, P, P, c_size_t, P]),
    "augur_direct_sum_f64_noalloc": (c_int, [P, P, P, c_int64, c_int, P, c_int64, c_double, P, P]),
    "augur_direct_potential_f64": (c_int, [P, P, P, c_int64, c_int, P, c_int64, c_double, P, P, c_size_t, P]),
    "augur_ewald_workspace_bytes": (c_size_t, [c_int64, c_int64, c_int]),
    "augur_ewald_f64": (c_int, [P, P, c_double, c_int64, P, c_int64, c_double, c_double, c_double, c_int, P, P,
                                c_size_t, P]),
    "augur_reduce_workspace_bytes": (c_size_t, [c_int64]),
    "augur_reduce_energy_momentum_f64": (c_int, [P, P, c_float, P, c_int64, P, P, c_size_t, P]),
    "augur_pm_workspace_bytes": (c_size_t, [c_int64, c_int64]),
    "augur_pm_prepare": (c_int, [P, P, P, P, P, c_int64, c_int64, c_float, P, c_size_t, P]),
    "augur_pm_accel": (c_int, [P, P, P, c_float, c_int64, c_int64, c_float, c_float, P, c_size_t, P]),
    "augur_pm_density": (c_int, [P, P, c_float, c_int64, c_int64, c_float, P, P, c_size_t, P]),
    "augur_pm_fields": (c_int, [P, P, c_float, c_int64, c_int64, c_float, c_float, P, P, P, c_size_t, P]),
    "augur_pm_kdk": (c_int, [P, P, P, P, P, c_float, c_int64, c_int64, c_float, c_float, c_float, c_int, P,
                             c_size_t, P]),
    "augur_pm_kdk_diag": (c_int, [P, P, P, P, P, c_float, c_int64, c_int64, c_float, c_float, c_float, c_int, P, P,
                                  c_size_t, P]),
    "augur_pm_step_begin": (c_int, [P, P, P, P, P, c_float, c_int64, c_int64, c_float, c_float, c_float, c_int, P,
                                    c_size_t, P]),
    "augur_pm_step_end": (c_int, [P, P, P, c_int64, c_int64, c_float, c_float, c_int, P, c_size_t, P]),
    "augur_pm_grid_offsets": (c_int, [c_int64, c_int64, P, P]),
    "augur_sur_gate_scratch_bytes": (c_size_t, [c_int64]),
    "augur_sur_gate": (c_int, [c_int, P, P, P, c_int64, c_int64, c_float, P, P, P, P, c_int, P]),
    "augur_pm_force_replay": (c_int, [c_int64, c_int64, c_float, P, P, c_size_t, P]),
    "augur_conv3_bf16_fwd_into": (c_int, [P, c_int, P, P, c_int64, c_int64, c_int, c_float, c_int, P, P, c_int, P]),
    "augur_persistent_step": (c_int, [P, P, P, P, c_int64, c_int64, P, c_int64, c_float, c_float, c_int, P]),
    "augur_persistent_step_tracked": (c_int, [P, P, P, P, c_int64, c_int64, P, c_int64, c_float, c_float, c_int, P,
                                              P]),
    "augur_conv3_bf16_fwd": (c_int, [P, c_int, P, P, c_int64, c_int64, c_int, c_int, P, P, P, P, P]),
    "augur_conv3_tile_variants": (c_int, [c_int, c_int, c_int]),
    "augur_conv3_bf16_fwd_tile": (c_int, [P, c_int, P, P, c_int64, c_int64, c_int, c_int, P, P, P, P, c_int, c_int,
                                          P]),
    "augur_conv3_wgrad_workspace_bytes": (c_size_t, [c_int, c_int64, c_int64, c_int]),
    "augur_conv3_bf16_wgrad": (c_int, [P, c_int, P, c_int64, c_int64, P, P, c_int, P, c_size_t, P]),
    "augur_conv3_flip_weights": (c_int, [P, c_int, c_int, c_int, P, P]),
    "augur_gelu_bwd": (c_int, [P, P, c_int64, P, P]),
    "augur_pack_cl_bf16": (c_int, [P, c_int, c_int64, c_int64, c_int, c_float, P, P]),
    "augur_pack_features_bf16": (c_int, [P, P, c_int64, c_int64, c_float, c_float, c_float, c_int, P, P]),
    "augur_adam_update": (c_int, [P, P, P, P, c_int64, c_float, c_float, c_float, c_float, c_int, c_float, P, P]),
    "augur_conv2d_f32": (c_int, [P, c_int, c_int, P, c_int, P, P, c_int64, c_int64, c_int64, c_int, c_int, P, P, P]),
    "augur_sf_kernel_grid": (c_int, [P, c_int64, c_float, c_float, P, P]),
    "augur_sf_neg_grad": (c_int, [P, c_int, c_int64, c_int64, c_float, P, P, P]),
    "augur_pair_knn_f32": (c_int, [P, P, P, c_int64, c_int64, P, P, c_int, c_float, c_float, c_float, c_int, P, P,
                                   P]),
    "augur_sf_momfix": (c_int, [P, P, P, c_int64, c_int64, c_int, c_int, P]),
    "augur_sf_update": (c_int, [P, P, P, P, P, c_int64, c_float, c_int, P]),
    "augur_grad_sumsq": (c_int, [P, c_int64, P, P]),
    "augur_grad_clip_decide": (c_int, [P, P, c_float, P, P, P]),
    "augur_adam_update_gated": (c_int, [P, P, P, P, c_int64, c_float, c_float, c_float, c_float, P, P, P, P]),
    "augur_lp_encode": (c_int, [P, c_int, P, c_int, c_int64, c_float, P, P, P]),
    "augur_lp_decode": (c_int, [P, c_int, P, c_int, c_int64, c_float, P, P, P]),
    "augur_lp_drift_deposit": (c_int, [P, c_int, P, c_int, c_int64, c_int64, c_int64, c_float, c_float, c_int64,
                                       c_float, P, P]),
    "augur_lp_finish": (c_int, [P, c_int64, c_int64, c_float, c_float, P, P]),
    "augur_pm_solve_grid": (c_int, [P, P, c_int64, c_float, c_float, P, c_size_t, P]),
    "augur_cic_deposit_ragged_tiled_workspace_bytes": (c_size_t, [c_int64, c_int64, c_int64]),
    "augur_cic_deposit_ragged_tiled": (c_int, [P, P, c_float, P, c_int64, c_int64, c_int64, c_float, c_float, P, P,
                                               c_size_t, P]),
    "augur_cic_deposit_ragged_planned_workspace_bytes": (c_size_t, [c_int64, c_int64, c_int64]),
    "augur_cic_deposit_ragged_planned": (c_int, [P, P, c_float, P, c_int64, c_int64, c_int64, c_float, c_float,
                                                 c_float, c_int, P, P, c_size_t, P]),
    "augur_ragged_plan_drift_word": (c_void_p, [P, c_size_t]),
    "augur_field_add": (c_int, [P, P, c_float, c_int64, P, P]),
    "augur_pm_solve_grids_workspace_bytes": (c_size_t, [c_int64, c_int64]),
    "augur_pm_solve_grids": (c_int, [P, P, c_int64, c_int64, c_float, c_float, P, c_size_t, P]),
    "augur_grid_mean_workspace_bytes": (c_size_t, []),
    "augur_grid_mean_acc_f64": (c_int, [P, P, c_int64, P, P, c_size_t, P]),
    "augur_lp_kick": (c_int, [P, c_int, P, c_int, c_int64, c_int64, c_float, P, P, c_int64, c_float, P]),
    "augur_timing_enable": (c_int, [c_int]),
    "augur_timing_filter": (c_int, [ctypes.c_char_p]),
    "augur_timing_collect": (c_int, [P, c_size_t, P, P, c_int, P]),
}


def timing_enable(on=True):
    call("augur_timing_enable", int(bool(on)))


def timing_filter(name=None):
    """Record only launches named `name` (None: all) while timing is enabled."""
    call("augur_timing_filter", None if name is None else name.encode())


def timing_collect():
    """{kernel name: (total ms, launches)} since timing_enable(); waits on the events."""
    cap, mx = 1 << 16, 256
    names = ctypes.create_string_buffer(cap)
    ms = (ctypes.c_double * mx)()
    cnt = (ctypes.c_int64 * mx)()
    n = ctypes.c_int(0)
    call("augur_timing_collect", ctypes.cast(names, c_void_p), cap, ctypes.cast(ms, c_void_p),
         ctypes.cast(cnt, c_void_p), mx, ctypes.cast(ctypes.pointer(n), c_void_p))
    keys = names.value.decode().split("\n")[: n.value]
    return {k: (ms[i], cnt[i]) for i, k in enumerate(keys)}

_lock = threading.Lock()
_lib = None


class AugurError(RuntimeError):
    pass


def header_symbols(path=HEADER):
    """Function names declared in include/augur.h."""
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(augur_\w+)\s*\(", txt)))


def get():
    """The loaded library (loads without a GPU)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(f"{LIB_PATH} is missing: build it with `python -m augur_amd.build` "
                                  "(the HIP library is the only compute path)")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype, fn.argtypes = res, args
            _lib = lib
    return _lib


def call(name, *args):
    rc = getattr(get(), name)(*args)
    if rc != 0:
        msg = get().augur_last_error()
        raise AugurError(f"{name} failed ({rc}): {msg.decode() if msg else ''}")


def size(name, *args):
    return int(getattr(get(), name)(*args))
