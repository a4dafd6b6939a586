"""numpy restatement of the reference particle-mesh path (test infrastructure).

Conventions shared with the HIP path
-----------------------------------
* Positions ``pos`` have shape (B, N, D), D in {2, 3}, component 0 = x.
* Grids are (B, C, G, G) for D=2 and (B, C, G, G, G) for D=3, x fastest
  (flat cell index ``((b*G + iz)*G + iy)*G + ix``), as in the reference's
  ``out.view(B, G, G, C).permute(0, 3, 1, 2)`` (scripts/scatter_field.py:250).
* ``off`` is the shift added to positions before dividing by the cell size:
  the reference ScatterField uses ``off = extent/2`` (grid covering
  [-extent/2, extent/2)); the 3D periodic unit box uses ``off = 0``.
* Keys and CIC base corners are computed in FP32 with the reference's op
  order (rounded add, rounded divide, rounded subtract, floor), so they are
  bit-exact against it; everything downstream is FP64.
"""
import numpy as np

F32 = np.float32


# --------------------------------------------------------------------- a1 / a3
def cell_keys(pos, mask, extent, cell, C):
    """Pair-term cell keys, restating ScatterField._cell_keys
    (scripts/scatter_field.py:302-311) for D in {2, 3}.

    ``c = floor((p + extent/2) / cell).clamp(0, C-1)``; key ``b*C^D + lin(c)``
    with x fastest; masked particles get ``nc = B*C^D`` (sorted last).
    Returns (keys int64 (B*N,), counts int64 (nc+1,)).
    """
    pos = np.asarray(pos, dtype=F32)
    B, N, D = pos.shape
    pf = pos.reshape(B * N, D)
    c = np.floor((pf + F32(extent / 2)) / F32(cell))
    c = np.clip(c, 0, C - 1).astype(np.int64)
    b = np.repeat(np.arange(B, dtype=np.int64), N)
    nc = B * C ** D
    lin = b * C ** D
    for d in range(D):
        lin = lin + c[:, d] * C ** d
    key = np.where(np.asarray(mask).reshape(-1) > 0, lin, nc)
    counts = np.bincount(key, minlength=nc + 1).astype(np.int64)
    return key, counts


def stable_sort_by_cell(keys):
    """``torch.argsort(key, stable=True)`` plus counts/starts
    (scripts/scatter_field.py:331-333): permutation ordering by (key, index)."""
    keys = np.asarray(keys)
    perm = np.argsort(keys, kind="stable").astype(np.int64)
    return perm


def cic_base(pos, off, h):
    """CIC base corner and weights, restating ScatterField._corners
    (scripts/scatter_field.py:179-182) and cic_density
    (model/tiered/energy.py:16-19): ``u = (pos + off)/h - 0.5``,
    ``i0 = floor(u)``, ``w = u - i0``, all FP32.  Returns (i0 int64, w float32)."""
    pos = np.asarray(pos, dtype=F32)
    u = (pos + F32(off)) / F32(h) - F32(0.5)
    i0 = np.floor(u)
    return i0.astype(np.int64), (u - i0).astype(F32)


# --------------------------------------------------------------------- a2 / a6
def _corner_list(D):
    return [tuple((k >> d) & 1 for d in range(D)) for k in range(1 << D)]


def cic_deposit(pos, mass, mask, G, h, off, periodic, vel=None, in_scale=1.0, dtype=np.float64):
    """CIC deposit, restating ScatterField.scatter (scripts/scatter_field.py:237-250)
    and cic_density (model/tiered/energy.py:9-34), generalised to D=3 and to a
    periodic wrap.

    Channels: [m] or, with ``vel``, [m, m*v_0, ..., m*v_{D-1}], each scaled by
    ``in_scale / h**D``.  Open boundary drops corners outside [0, G) (the
    reference's ``ok`` mask, :206-209); periodic wraps them mod G.
    """
    pos = np.asarray(pos, dtype=F32)
    B, N, D = pos.shape
    mm = (np.asarray(mass, dtype=np.float64) * np.asarray(mask, dtype=np.float64))
    chans = [mm]
    if vel is not None:
        vel = np.asarray(vel, dtype=np.float64)
        chans += [mm * vel[..., d] for d in range(D)]
    val = np.stack(chans, -1) * (in_scale / h ** D)          # (B, N, C)
    C = val.shape[-1]
    i0, w = cic_base(pos, off, h)
    w = w.astype(np.float64)
    out = np.zeros((B, G ** D, C), dtype=dtype)
    bidx = np.repeat(np.arange(B), N).reshape(B, N)
    for o in _corner_list(D):
        wt = np.ones((B, N))
        ok = np.ones((B, N), dtype=bool)
        flat = np.zeros((B, N), dtype=np.int64)
        for d in range(D):
            ix = i0[..., d] + o[d]
            wt = wt * (w[..., d] if o[d] else 1 - w[..., d])
            if periodic:
                ix = np.mod(ix, G)
            else:
                ok &= (ix >= 0) & (ix < G)
                ix = np.clip(ix, 0, G - 1)
            flat = flat + ix * G ** d
        wt = wt * ok
        contrib = val * wt[..., None]
        np.add.at(out, (bidx, flat), contrib)
    return np.moveaxis(out, -1, 1).reshape((B, C) + (G,) * D)


def cic_deposit_bruteforce(pos, mass, G, h, off, periodic):
    """Independent form of the same deposit: every cell c gets
    sum_p m_p * prod_d max(0, 1 - |u_pd - c_d|) (the CIC tent function),
    evaluated over all (particle, cell) pairs.  O(N G^D): tiny inputs only."""
    pos = np.asarray(pos, dtype=np.float64)
    B, N, D = pos.shape
    u = (pos + off) / h - 0.5
    grids = np.meshgrid(*[np.arange(G)] * D, indexing="ij")   # (z,y,x) order reversed below
    cells = np.stack([g.reshape(-1) for g in grids[::-1]], -1)  # (G^D, D), column d = coordinate d
    out = np.zeros((B, G ** D))
    for b in range(B):
        dist = u[b][:, None, :] - cells[None, :, :]
        if periodic:
            dist = (dist + G / 2) % G - G / 2
        wt = np.prod(np.clip(1 - np.abs(dist), 0, None), -1)
        out[b] = (np.asarray(mass, dtype=np.float64)[b][:, None] * wt).sum(0)
    return (out / h ** D).reshape((B, 1) + (G,) * D)


def cic_gather(field, pos, h, off, periodic):
    """CIC gather, restating ScatterField.gather (scripts/scatter_field.py:252-261),
    generalised to D=3 and periodic wrap.  field (B, C, G..G) -> (B, N, C)."""
    field = np.asarray(field, dtype=np.float64)
    pos = np.asarray(pos, dtype=F32)
    B, N, D = pos.shape
    C, G = field.shape[1], field.shape[-1]
    fl = np.moveaxis(field.reshape(B, C, G ** D), 1, -1)       # (B, G^D, C)
    i0, w = cic_base(pos, off, h)
    w = w.astype(np.float64)
    out = np.zeros((B, N, C))
    bidx = np.repeat(np.arange(B), N).reshape(B, N)
    for o in _corner_list(D):
        wt = np.ones((B, N))
        ok = np.ones((B, N), dtype=bool)
        flat = np.zeros((B, N), dtype=np.int64)
        for d in range(D):
            ix = i0[..., d] + o[d]
            wt = wt * (w[..., d] if o[d] else 1 - w[..., d])
            if periodic:
                ix = np.mod(ix, G)
            else:
                ok &= (ix >= 0) & (ix < G)
                ix = np.clip(ix, 0, G - 1)
            flat = flat + ix * G ** d
        out += fl[bidx, flat] * (wt * ok)[..., None]
    return out


# ------------------------------------------------------------------ a4 / a5 (2D open)
PM_EPS = 0.5


def neg_grad_fd(phi, h):
    """Zero-padded central difference -grad, restating ScatterField.neg_grad
    (scripts/scatter_field.py:551-555) for (B, 1, G, G) -> (B, 2, G, G)."""
    p = np.pad(phi, [(0, 0), (0, 0), (1, 1), (1, 1)])
    gx = (p[:, :, 1:-1, 2:] - p[:, :, 1:-1, :-2]) / (2 * h)
    gy = (p[:, :, 2:, 1:-1] - p[:, :, :-2, 1:-1]) / (2 * h)
    return -np.concatenate([gx, gy], 1)


def open_kernel_2d(G, h, eps=PM_EPS):
    """Softened kernel sampled at min-image distances on the (2G)^2 domain,
    restating pm_acc (scripts/pm_baseline_force_error.py:71-73)."""
    idx = np.arange(2 * G, dtype=np.float64)
    d = np.minimum(idx, 2 * G - idx) * h
    return -1.0 / np.sqrt(d[:, None] ** 2 + d[None, :] ** 2 + eps ** 2)


def open_potential_2d(rho, h, eps=PM_EPS):
    """Zero-padded FFT convolution, restating pm_acc
    (scripts/pm_baseline_force_error.py:74): irfft2(rfft2(rho, s=(2G,2G)) *
    rfft2(K)) * h^2, cropped to [:G, :G]."""
    G = rho.shape[-1]
    K = open_kernel_2d(G, h, eps)
    full = np.fft.irfft2(np.fft.rfft2(rho, s=(2 * G, 2 * G)) * np.fft.rfft2(K), s=(2 * G, 2 * G)) * h ** 2
    return full[..., :G, :G]


def open_potential_2d_bruteforce(rho, h, eps=PM_EPS):
    """Independent form: phi(i) = h^2 sum_j rho(j) * -1/sqrt(|i-j|^2 h^2 + eps^2)
    over the G^2 cells (open boundary, no images).  O(G^4)."""
    B, C, G, _ = rho.shape
    iy, ix = np.meshgrid(np.arange(G), np.arange(G), indexing="ij")
    cy, cx = iy.reshape(-1), ix.reshape(-1)
    r2 = ((cy[:, None] - cy[None, :]) ** 2 + (cx[:, None] - cx[None, :]) ** 2) * h * h
    K = -1.0 / np.sqrt(r2 + eps ** 2)
    flat = rho.reshape(B, C, G * G)
    return (h ** 2 * flat @ K.T).reshape(B, C, G, G)


def pm_acc_2d_open(pos, mass, mask, G, extent, eps=PM_EPS, in_scale=1.0):
    """Plain PM force, restating pm_acc (scripts/pm_baseline_force_error.py:68-76):
    CIC deposit -> zero-padded FFT convolution with -1/sqrt(r^2+eps^2) -> FD
    gradient -> CIC gather, times mask.  FP64."""
    h = extent / G
    rho = cic_deposit(pos, mass, mask, G, h, extent / 2, False, in_scale=in_scale)
    phi = open_potential_2d(rho, h, eps)
    acc = neg_grad_fd(phi, h)
    return cic_gather(acc, pos, h, extent / 2, False) * np.asarray(mask, dtype=np.float64)[..., None]


# ------------------------------------------------------------- a4 / a5 (3D periodic)
def wavenumbers(G, L):
    """Signed FFT frequencies times 2*pi/L for one axis (numpy fftfreq order)."""
    return 2 * np.pi * np.fft.fftfreq(G, d=L / G)


def poisson_periodic_accel_grids(rho, L, Gconst=1.0, zero_nyquist=True):
    """3D periodic Poisson solve with spectral gradient (no reference exists:
    the reference is 2D open-boundary).  phi_hat = -4 pi Gconst rho_hat / |k|^2,
    phi_hat(0) = 0, a_hat_d = -i k_d phi_hat with the Nyquist plane of the
    differentiated axis zeroed.  rho (G,G,G) indexed [z, y, x]; returns
    (3, G, G, G) accelerations [a_x, a_y, a_z] and phi."""
    G = rho.shape[-1]
    rk = np.fft.rfftn(rho)                       # (G, G, G//2+1) over axes (z, y, x)
    kz = wavenumbers(G, L)[:, None, None]
    ky = wavenumbers(G, L)[None, :, None]
    kx = (2 * np.pi * np.arange(G // 2 + 1) / L)[None, None, :]
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k2[0, 0, 0] = 1.0
    phik = -4 * np.pi * Gconst * rk / k2
    phik[0, 0, 0] = 0.0
    out = []
    for d, kd in enumerate((kx, ky, kz)):
        kd = np.broadcast_to(kd, phik.shape).copy()
        if zero_nyquist:
            if d == 0:
                kd[..., G // 2] = 0.0
            elif d == 1:
                kd[:, G // 2, :] = 0.0
            else:
                kd[G // 2, :, :] = 0.0
        out.append(np.fft.irfftn(-1j * kd * phik, s=(G, G, G), axes=(0, 1, 2)))
    phi = np.fft.irfftn(phik, s=(G, G, G), axes=(0, 1, 2))
    return np.stack(out), phi


def pm_acc_3d_periodic(pos, mass, G, L=1.0, Gconst=1.0):
    """Plain periodic PM acceleration for one scene: CIC deposit (periodic) ->
    Green's function -4 pi G/k^2 -> spectral gradient -> CIC gather.  pos (N,3)."""
    pos = np.asarray(pos, dtype=F32)[None]
    N = pos.shape[1]
    h = L / G
    m = np.broadcast_to(np.asarray(mass, dtype=np.float64), (N,))[None]
    rho = cic_deposit(pos, m, np.ones((1, N)), G, h, 0.0, True)[0, 0]
    acc, _ = poisson_periodic_accel_grids(rho, L, Gconst)
    return cic_gather(acc[None], pos, h, 0.0, True)[0]


def kdk_pm_rollout(pos, vel, mass, G, dt, steps, L=1.0, Gconst=1.0, force=None):
    """KDK leapfrog with the periodic PM force (restating the integrator of
    scripts/gravity_sim.py:31-44 with substeps=1 and the force replaced by PM).
    FP64 state; positions wrapped into [0, L).  ``force`` (pos -> acc)
    replaces the PM force when given (the integrator's own pins use it)."""
    if force is None:
        def force(p):
            return pm_acc_3d_periodic(p, mass, G, L, Gconst)
    pos = np.asarray(pos, dtype=np.float64).copy()
    vel = np.asarray(vel, dtype=np.float64).copy()
    a = force(pos)
    for _ in range(steps):
        vel += 0.5 * dt * a
        pos = np.mod(pos + dt * vel, L)
        a = force(pos)
        vel += 0.5 * dt * a
    return pos, vel


# ------------------------------------------------------------------------ a13
def errs(a, at, mask):
    """Per-particle force errors, restating errs (scripts/pm_baseline_force_error.py:88-92):
    |a - a_true| / rms(|a_true|) of the scene, and |a - a_true| / |a_true|.
    a, at (B, N, D); mask (B, N).  Returns the masked values, flattened."""
    a, at, mask = (np.asarray(x, dtype=np.float64) for x in (a, at, mask))
    rms = np.sqrt((np.linalg.norm(at, axis=-1) ** 2 * mask).sum(1) / mask.sum(1))[:, None]
    d = np.linalg.norm(a - at, axis=-1)
    e = (d / rms)[mask > 0]
    rel = (d / np.maximum(np.linalg.norm(at, axis=-1), 1e-12))[mask > 0]
    return e, rel


def stats(x):
    """median / p90 / mean, restating stats (scripts/pm_baseline_force_error.py:95-96)."""
    x = np.asarray(x, dtype=np.float64)
    return {"median": float(np.median(x)), "p90": float(np.quantile(x, 0.9)), "mean": float(x.mean())}


def rel_rms(a, at):
    """Force rel-RMS: sqrt(mean |a - a_t|^2) / sqrt(mean |a_t|^2) over targets."""
    a, at = np.asarray(a, dtype=np.float64), np.asarray(at, dtype=np.float64)
    return float(np.sqrt(((a - at) ** 2).sum(-1).mean() / (at ** 2).sum(-1).mean()))
