"""Typed wrappers over the C ABI: allocate outputs and workspaces with torch on
the tensors' device and launch on the current stream.  Every function needs
its tensors on a ROCm GPU and raises otherwise (there is no CPU path)."""
import ctypes
import math

import torch

from . import _lib

F32, F64, I64, I32 = torch.float32, torch.float64, torch.int64, torch.int32


def _dev(*ts):
    dev = None
    for t in ts:
        if t is None:
            continue
        if not t.is_cuda:
            raise RuntimeError("augur_amd runs on a ROCm GPU only: got a tensor on " + str(t.device))
        dev = t.device if dev is None else dev
    if dev is None:
        raise RuntimeError("augur_amd: no device tensor given")
    return dev


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _c(t, dtype):
    return None if t is None else t.contiguous().to(dtype)


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _ws(nbytes, dev):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=dev)


def _shape_arr(shape):
    return (ctypes.c_int64 * len(shape))(*[int(s) for s in shape])


# ------------------------------------------------------------------ keys / sort
def cell_keys(pos, mask, off, cell, C, counts=False):
    """pos (B,N,D) -> int64 keys (B*N,), optional counts (B*C^D+1,)."""
    pos = _c(pos, F32)
    dev = _dev(pos)
    B, N, D = pos.shape
    mask = _c(mask, F32) if mask is not None else None
    keys = torch.empty(B * N, dtype=I64, device=dev)
    cnt = torch.empty(B * C ** D + 1, dtype=I64, device=dev) if counts else None
    _lib.call("augur_cell_keys", _p(pos), _p(mask), B, N, D, float(off), float(cell), int(C), _p(keys), _p(cnt),
              _stream(dev))
    return (keys, cnt) if counts else keys


def cic_base(pos, off, h):
    """pos (..., D) -> (i0 int32 (..., D), w float32 (..., D))."""
    pos = _c(pos, F32)
    dev = _dev(pos)
    D = pos.shape[-1]
    n = pos.numel() // D
    i0 = torch.empty(pos.shape, dtype=I32, device=dev)
    w = torch.empty(pos.shape, dtype=F32, device=dev)
    _lib.call("augur_cic_base", _p(pos), n, D, float(off), float(h), _p(i0), _p(w), _stream(dev))
    return i0, w


def sort_by_cell(keys, nkeys, starts=False):
    """Stable argsort by key -> int64 perm (and starts (nkeys+1,))."""
    keys = _c(keys, I64)
    dev = _dev(keys)
    n = keys.numel()
    perm = torch.empty(n, dtype=I64, device=dev)
    st = torch.empty(nkeys + 1, dtype=I64, device=dev) if starts else None
    wsb = _lib.size("augur_sort_workspace_bytes", n, nkeys)
    ws = _ws(wsb, dev)
    _lib.call("augur_sort_by_cell", _p(keys), n, int(nkeys), _p(perm), _p(st), _p(ws), wsb, _stream(dev))
    return (perm, st) if starts else perm


# ------------------------------------------------------------- deposit / gather
def cic_deposit(pos, mass, G, h, off, periodic, vel=None, mask=None, scale=None, mass_value=1.0):
    """pos (B,N,D), mass (B,N) or None -> (B, C, G, .., G) float32, C = 1 or 1+D."""
    pos = _c(pos, F32)
    dev = _dev(pos)
    B, N, D = pos.shape
    mass = _c(mass, F32)
    vel = _c(vel, F32)
    mask = _c(mask, F32)
    C = 1 if vel is None else 1 + D
    if scale is None:
        scale = 1.0 / h ** D
    out = torch.empty((B, C) + (G,) * D, dtype=F32, device=dev)
    wsb = _lib.size("augur_cic_deposit_workspace_bytes", B, D, G, C)
    ws = _ws(wsb, dev)
    _lib.call("augur_cic_deposit", _p(pos), _p(vel), _p(mass), float(mass_value), _p(mask), B, N, D, int(G),
              float(off), float(h), int(bool(periodic)), float(scale), _p(out), _p(ws), wsb, _stream(dev))
    return out


def tsc_deposit(pos, mass, G, h, off, periodic, vel=None, mask=None, scale=None, mass_value=1.0):
    """The TSC ("quad") deposit (augur_tsc_deposit); arguments as cic_deposit."""
    pos = _c(pos, F32)
    dev = _dev(pos)
    B, N, D = pos.shape
    mass, vel, mask = _c(mass, F32), _c(vel, F32), _c(mask, F32)
    C = 1 if vel is None else 1 + D
    if scale is None:
        scale = 1.0 / h ** D
    out = torch.empty((B, C) + (G,) * D, dtype=F32, device=dev)
    wsb = _lib.size("augur_cic_deposit_workspace_bytes", B, D, G, C)
    ws = _ws(wsb, dev)
    _lib.call("augur_tsc_deposit", _p(pos), _p(vel), _p(mass), float(mass_value), _p(mask), B, N, D, int(G),
              float(off), float(h), int(bool(periodic)), float(scale), _p(out), _p(ws), wsb, _stream(dev))
    return out


def tsc_gather(field, pos, h, off, periodic):
    """The TSC gather (augur_tsc_gather): field (B, C, G..G), pos (B, N, D) -> (B, N, C)."""
    field, pos = _c(field, F32), _c(pos, F32)
    dev = _dev(field, pos)
    B, N, D = pos.shape
    C, G = field.shape[1], field.shape[-1]
    out = torch.empty(B, N, C, dtype=F32, device=dev)
    _lib.call("augur_tsc_gather", _p(field), _p(pos), B, N, D, int(G), int(C), float(off), float(h),
              int(bool(periodic)), _p(out), _stream(dev))
    return out


def cic_deposit_ragged(pos, starts, G, h, off, periodic, mass=None, vel=None, scale=None, mass_value=1.0,
                       tiled=None):
    """Ragged batch: pos (total, D) with scene b = rows [starts[b], starts[b+1]),
    starts an int64 tensor of B+1 offsets on the device; mass (total,) or
    None (mass_value each) -> (B, C, G, .., G) float32, every scene at once.
    A 3D periodic mass deposit (off 0, no velocity channels, G a power of two
    in [32, 512]) runs on the tiled LDS deposit (augur_cic_deposit_ragged_tiled,
    the same bits); tiled=False forces the global-atomic deposit."""
    pos = _c(pos, F32)
    dev = _dev(pos)
    total, D = pos.shape
    starts = _c(starts, torch.int64)
    B = starts.shape[0] - 1
    mass = _c(mass, F32)
    vel = _c(vel, F32)
    C = 1 if vel is None else 1 + D
    if scale is None:
        scale = 1.0 / h ** D
    ok = (D == 3 and bool(periodic) and vel is None and float(off) == 0.0 and G & (G - 1) == 0 and 32 <= G <= 512
          and B * G ** 3 < 2 ** 32 and B <= 65535 and total < 2 ** 31)
    if tiled is None:
        tiled = ok
    if tiled:
        assert ok, "the tiled ragged deposit needs D = 3, periodic, off 0, no vel, G a power of two in [32, 512]"
        out = torch.empty((B, 1, G, G, G), dtype=F32, device=dev)
        ws = _ws(_lib.size("augur_cic_deposit_ragged_tiled_workspace_bytes", B, G, total), dev)
        _lib.call("augur_cic_deposit_ragged_tiled", _p(pos), _p(mass), float(mass_value), _p(starts), B, total, int(G),
                  float(h), float(scale), _p(out), _p(ws), ws.numel(), _stream(dev))
        return out
    out = torch.empty((B, C) + (G,) * D, dtype=F32, device=dev)
    wsb = _lib.size("augur_cic_deposit_workspace_bytes", B, D, G, C)
    ws = _ws(wsb, dev)
    _lib.call("augur_cic_deposit_ragged", _p(pos), _p(vel), _p(mass), float(mass_value), _p(starts), B, total, D,
              int(G), float(off), float(h), int(bool(periodic)), float(scale), _p(out), _p(ws), wsb, _stream(dev))
    return out


class RaggedPlan:
    """The persistent workspace of a planned ragged deposit
    (augur_cic_deposit_ragged_planned): the batch's sort and tile plans kept
    between calls and redone on the device only when the drift recorded by
    augur_persistent_step_tracked (through ``drift_word``) may have moved a
    particle a cell since the last sort."""

    def __init__(self, B, G, total, device):
        self.key = (int(B), int(G), int(total))
        self.ws = _ws(_lib.size("augur_cic_deposit_ragged_planned_workspace_bytes", B, G, total), device)
        self.drift_word = _lib.get().augur_ragged_plan_drift_word(_p(self.ws), self.ws.numel())
        assert self.drift_word, "planned ragged deposit: workspace too small"


def cic_deposit_ragged_planned(pos, starts, G, h, plan, drift_dt, force_sort, mass=None, mass_value=1.0,
                               scale=None):
    """As cic_deposit_ragged on the tiled deposit (3D periodic, off 0, mass
    only; the same bits), with the sort and plans kept in `plan` (a
    RaggedPlan of this batch's sizes): force_sort re-sorts unconditionally,
    otherwise the device re-sorts when the tracked drift (drift_dt times the
    largest |v| of the drifts since the previous call) would pass a cell."""
    pos = _c(pos, F32)
    dev = _dev(pos)
    total = pos.shape[0]
    starts = _c(starts, torch.int64)
    B = starts.shape[0] - 1
    mass = _c(mass, F32)
    assert plan.key == (B, int(G), total), (plan.key, (B, G, total))
    if scale is None:
        scale = 1.0 / h ** 3
    out = torch.empty((B, 1, G, G, G), dtype=F32, device=dev)
    _lib.call("augur_cic_deposit_ragged_planned", _p(pos), _p(mass), float(mass_value), _p(starts), B, total, int(G),
              float(h), float(scale), float(drift_dt), int(bool(force_sort)), _p(out), _p(plan.ws), plan.ws.numel(),
              _stream(dev))
    return out


def cic_gather(field, pos, h, off, periodic):
    """field (B, C, G..G), pos (B, N, D) -> (B, N, C)."""
    field = _c(field, F32)
    pos = _c(pos, F32)
    dev = _dev(field, pos)
    B, N, D = pos.shape
    C, G = field.shape[1], field.shape[-1]
    out = torch.empty(B, N, C, dtype=F32, device=dev)
    _lib.call("augur_cic_gather", _p(field), _p(pos), B, N, D, int(G), int(C), float(off), float(h),
              int(bool(periodic)), _p(out), _stream(dev))
    return out


# ------------------------------------------------------------------------ FFT
def rfftn(x, s=None):
    """Real FFT over the last 2 or 3 axes of x (batch dims in front), zero
    extended to s (the reference's rfft2(x, s=...)).  complex64 output."""
    x = _c(x, F32)
    dev = _dev(x)
    nd = 3 if (s is not None and len(s) == 3) else (2 if s is not None else None)
    if nd is None:
        raise ValueError("rfftn: pass s (the FFT shape over the last 2 or 3 axes)")
    in_shape = x.shape[-nd:]
    batch = x.numel() // max(1, int(torch.tensor(in_shape).prod()))
    out = torch.empty(x.shape[:-nd] + tuple(s[:-1]) + (s[-1] // 2 + 1,), dtype=torch.complex64, device=dev)
    _lib.call("augur_fft_r2c", _p(x), _shape_arr(in_shape), _p(out), nd, _shape_arr(s), batch, None, 0, _stream(dev))
    return out


def irfftn(X, s, out_shape=None, scale=None):
    """Inverse of rfftn (torch 'backward' normalisation unless scale given),
    cropped to out_shape.  X is copied (the kernels consume their input)."""
    Xc = X.contiguous().to(torch.complex64).clone()
    dev = _dev(Xc)
    nd = len(s)
    batch = Xc.numel() // (int(torch.tensor(s[:-1]).prod()) * (s[-1] // 2 + 1))
    os_ = tuple(out_shape) if out_shape is not None else tuple(s)
    if scale is None:
        scale = 1.0 / float(torch.tensor(s).prod())
    out = torch.empty(X.shape[:-nd] + os_, dtype=F32, device=dev)
    _lib.call("augur_fft_c2r", _p(Xc), _p(out), _shape_arr(os_), nd, _shape_arr(s), batch, float(scale), None, 0,
              _stream(dev))
    return out


def complex_mul(a, b, scale=1.0):
    a = a.contiguous().to(torch.complex64)
    b = b.contiguous().to(torch.complex64)
    dev = _dev(a, b)
    n = b.numel()
    out = torch.empty_like(a)
    _lib.call("augur_complex_mul", _p(a), _p(b), _p(out), n, a.numel() // n, float(scale), _stream(dev))
    return out


def poisson_green_grad(rho_hat, G, L, Gconst=1.0, scale=1.0, want_phi=False):
    rho_hat = rho_hat.contiguous().to(torch.complex64)
    dev = _dev(rho_hat)
    B = rho_hat.numel() // (G * G * (G // 2 + 1))
    acc = torch.empty((B, 3, G, G, G // 2 + 1), dtype=torch.complex64, device=dev)
    phi = torch.empty((B, G, G, G // 2 + 1), dtype=torch.complex64, device=dev) if want_phi else None
    _lib.call("augur_poisson_green_grad", _p(rho_hat), B, G, float(L), float(Gconst), float(scale), _p(acc), _p(phi),
              _stream(dev))
    return (acc, phi) if want_phi else acc


def pm_solve_grids(rho, L=1.0, Gconst=1.0, out=None):
    """Periodic PM force planes of B density grids: rho (B, G, G, G) FP32 ->
    (B, 3, G, G, G) FP32 [a_x, a_y, a_z], the single-scene stepper's fused
    FFT passes with the batch as a grid dimension (each grid gets the bits it
    gets alone)."""
    rho = _c(rho, F32)
    dev = _dev(rho)
    B, G = rho.shape[0], rho.shape[-1]
    assert tuple(rho.shape[1:]) == (G, G, G), rho.shape
    if out is None:
        out = torch.empty(B, 3, G, G, G, dtype=F32, device=dev)
    assert out.dtype == F32 and out.is_contiguous() and tuple(out.shape) == (B, 3, G, G, G)
    ws = _ws(_lib.size("augur_pm_solve_grids_workspace_bytes", B, G), dev)
    _lib.call("augur_pm_solve_grids", _p(rho), _p(out), B, G, float(L), float(Gconst), _p(ws), ws.numel(),
              _stream(dev))
    return out


def open_kernel_2d(G, h, eps, device):
    out = torch.empty(2 * G, 2 * G, dtype=F32, device=device)
    _lib.call("augur_open_kernel_2d", int(G), float(h), float(eps), _p(out), _stream(torch.device(device)))
    return out


def neg_grad_fd(phi, h):
    """phi (B, 1, G, G) -> (B, 2, G, G), zero-padded central differences."""
    phi = _c(phi, F32)
    dev = _dev(phi)
    B, G = phi.shape[0], phi.shape[-1]
    out = torch.empty(B, 2, G, G, dtype=F32, device=dev)
    _lib.call("augur_neg_grad_fd", _p(phi), B, G, float(h), _p(out), _stream(dev))
    return out


# ------------------------------------------------------------------- oracles
def _check_targets(tg, n):
    """Target indices index the source arrays inside the kernels: validate on the host."""
    if tg is not None and tg.numel() and (int(tg.min()) < 0 or int(tg.max()) >= n):
        raise ValueError(f"target indices must lie in [0, {n})")


def direct_potential_f64(pos, eps, mass=None, targets=None):
    """FP64 softened potential sums phi_t = sum_{j != t} m_j / sqrt(r^2 + eps^2)."""
    return direct_sum_f64(pos, eps, mass, targets, potential=True)


def direct_sum_f64(pos, eps, mass=None, targets=None, potential=False):
    """FP64 softened direct sum.  pos (n, D) float32 or float64."""
    dev = _dev(pos)
    n, D = pos.shape
    p64 = _c(pos, F64) if pos.dtype == F64 else None
    p32 = _c(pos, F32) if pos.dtype != F64 else None
    m = _c(mass, F64) if mass is not None else None
    tg = _c(targets, I64) if targets is not None else None
    _check_targets(tg, n)
    t = n if tg is None else tg.numel()
    out = torch.empty((t,) if potential else (t, D), dtype=F64, device=dev)
    wsb = _lib.size("augur_direct_sum_workspace_bytes", n, t, D)
    ws = _ws(wsb, dev)
    name = "augur_direct_potential_f64" if potential else "augur_direct_sum_f64"
    _lib.call(name, _p(p64), _p(p32), _p(m), n, D, _p(tg), t, float(eps), _p(out), _p(ws), wsb, _stream(dev))
    return out


def ewald_f64(pos, mass, targets, L=1.0, eps=0.0, alpha_l=8.0, kmax=None):
    """FP64 Ewald (periodic) accelerations at the targets.  pos (n, 3) float32."""
    import math
    pos = _c(pos, F32)
    dev = _dev(pos)
    n = pos.shape[0]
    if kmax is None:
        kmax = int(math.ceil(2 * alpha_l * math.sqrt(18.5) / (2 * math.pi))) + 1
    m = None
    mval = 1.0
    if torch.is_tensor(mass):
        m = _c(mass, F64)
    elif mass is not None:
        mval = float(mass)
    tg = _c(targets, I64)
    _check_targets(tg, n)
    t = tg.numel()
    out = torch.empty(t, 3, dtype=F64, device=dev)
    wsb = _lib.size("augur_ewald_workspace_bytes", n, t, kmax)
    ws = _ws(wsb, dev)
    _lib.call("augur_ewald_f64", _p(pos), _p(m), mval, n, _p(tg), t, float(L), float(eps), float(alpha_l), int(kmax),
              _p(out), _p(ws), wsb, _stream(dev))
    return out


def energy_momentum_f64(vel, mass=None, mass_value=1.0, phi=None):
    """-> float64 tensor [px, py, pz, kinetic, 0.5*sum m phi, sum m]."""
    vel = _c(vel, F32)
    dev = _dev(vel)
    n = vel.shape[0]
    m = _c(mass, F32)
    ph = _c(phi, F32)
    out = torch.empty(6, dtype=F64, device=dev)
    wsb = _lib.size("augur_reduce_workspace_bytes", n)
    ws = _ws(wsb, dev)
    _lib.call("augur_reduce_energy_momentum_f64", _p(vel), _p(m), float(mass_value), _p(ph), n, _p(out), _p(ws), wsb,
              _stream(dev))
    return out


# ------------------------------------------------- 2D scatter-field step
def conv2d(a, weight, bias=None, act=0, amode=0, b=None, res=None):
    """FP32 3x3 zero-padded conv (augur_conv2d_f32).  a (B, ca, Ha, Wa) read
    through amode (0 same size, 1 2x2 average, 2 nearest x2), b (B, cb, H, W)
    concatenated after it; weight (cout, ca + cb, 3, 3); act 0 none, 1 GELU,
    2 res + GELU.  -> (B, cout, H, W)."""
    a = _c(a, F32)
    dev = _dev(a)
    B, ca, Ha, Wa = a.shape
    H, W = (2 * Ha, 2 * Wa) if amode == 2 else ((Ha // 2, Wa // 2) if amode == 1 else (Ha, Wa))
    b = _c(b, F32)
    cb = 0 if b is None else b.shape[1]
    if b is not None:
        assert b.shape == (B, cb, H, W), (b.shape, (B, cb, H, W))
    w = _c(weight, F32)
    cout = w.shape[0]
    assert w.shape == (cout, ca + cb, 3, 3)
    res = _c(res, F32)
    out = torch.empty(B, cout, H, W, dtype=F32, device=dev)
    _lib.call("augur_conv2d_f32", _p(a), ca, int(amode), _p(b), cb, _p(w), _p(_c(bias, F32)), B, H, W, cout,
              int(act), _p(res), _p(out), _stream(dev))
    return out


def mlp_params(seq):
    """Flat FP32 parameters of nn.Sequential(Linear(2, 64), GELU, Linear(64, 64),
    GELU, Linear(64, 1)) in module order (the layout the kernels read)."""
    ps = [seq[0].weight, seq[0].bias, seq[2].weight, seq[2].bias, seq[4].weight, seq[4].bias]
    assert tuple(seq[0].weight.shape) == (64, 2) and tuple(seq[4].weight.shape) == (1, 64), \
        "the kernels evaluate 2 -> 64 -> 64 -> 1 MLPs"
    return torch.cat([p.detach().reshape(-1).float() for p in ps]).contiguous()


def sf_kernel_grid(mlp_flat, G, h, pp=0.0):
    """Learned kernel on the (2G)^2 min-image grid (augur_sf_kernel_grid)."""
    dev = _dev(mlp_flat)
    K = torch.empty(2 * G, 2 * G, dtype=F32, device=dev)
    _lib.call("augur_sf_kernel_grid", _p(mlp_flat), int(G), float(h), float(pp), _p(K), _stream(dev))
    return K


def sf_neg_grad(phi, G, h, ext=False, addend=None):
    """phi (B, 1, G, G) (ext: (B, 1, 2G, 2G)) -> -grad (B, 2, G, G) (+ addend)."""
    phi = _c(phi, F32)
    dev = _dev(phi)
    B = phi.shape[0]
    out = torch.empty(B, 2, G, G, dtype=F32, device=dev)
    _lib.call("augur_sf_neg_grad", _p(phi), int(bool(ext)), B, int(G), float(h), _p(_c(addend, F32)), _p(out),
              _stream(dev))
    return out


def pair_knn(pos, mass, mask, extent, pp, knn, mlp_flat, acc):
    """Adds the cell-list kNN pair term to acc (B, N, 2) in place."""
    pos = _c(pos, F32)
    dev = _dev(pos, acc)
    B, N, _ = pos.shape
    C = int(math.ceil(extent / pp))
    keys = cell_keys(pos, mask, extent / 2, pp, C)
    perm, starts = sort_by_cell(keys, B * C * C + 1, starts=True)
    _lib.call("augur_pair_knn_f32", _p(pos), _p(_c(mass, F32)), _p(_c(mask, F32)), B, N, _p(perm), _p(starts), C,
              float(extent / 2), float(pp), float(pp), int(knn), _p(mlp_flat), _p(acc), _stream(dev))
    return acc


def sf_momfix(acc, mass, mask, fix=True):
    """acc (B, N, D) in place: minus the mass-weighted mean (fix), times mask."""
    dev = _dev(acc)
    B, N, D = acc.shape
    _lib.call("augur_sf_momfix", _p(acc), _p(_c(mass, F32)), _p(_c(mask, F32)), B, N, D, int(bool(fix)),
              _stream(dev))
    return acc


def sf_update(pos, vel, a0, a1, aux, dt, phase):
    """In-place step updates (augur_sf_update), all tensors FP32 contiguous of one shape."""
    dev = _dev(pos, vel, a0)
    for t in (pos, vel, a0, a1, aux):
        assert t is None or (t.dtype == F32 and t.is_contiguous() and t.numel() == pos.numel())
    _lib.call("augur_sf_update", _p(pos), _p(vel), _p(a0), _p(a1), _p(aux), pos.numel(), float(dt), int(phase),
              _stream(dev))
