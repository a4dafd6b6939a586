"""Seeded 3D scenes for the PM workloads (host numpy, FP64 then FP32; the
N = 2^30 scene is drawn on the GPU with torch, since its FP64 host arrays
would take tens of GB).

plummer:  r = a / sqrt(u^(-2/3) - 1), isotropic direction, radii >= rmax redrawn.
clumpy:   a fraction uniform in the unit box, the rest split equally over
          Gaussian clumps with uniform-random centres, wrapped periodically.
The reference's samplers are 2D (scripts/pm_baseline_force_error.py:28-65).
"""
import numpy as np


def plummer(n, a, rng, center=(0.5, 0.5, 0.5), rmax=0.45):
    out = np.zeros((0, 3))
    while len(out) < n:
        m = 2 * (n - len(out)) + 16
        u = rng.uniform(0.0, 1.0, m)
        u = u[u > 0]
        r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        r = r[r < rmax]
        k = len(r)
        ct = rng.uniform(-1.0, 1.0, k)
        ph = rng.uniform(0.0, 2 * np.pi, k)
        st = np.sqrt(1 - ct * ct)
        out = np.concatenate([out, r[:, None] * np.stack([st * np.cos(ph), st * np.sin(ph), ct], 1)])
    return out[:n] + np.asarray(center)


def clumpy(n, rng, n_clumps=64, sigma=0.02, frac_uniform=0.5):
    nu = int(round(n * frac_uniform))
    nc = n - nu
    uni = rng.uniform(0.0, 1.0, (nu, 3))
    cen = rng.uniform(0.0, 1.0, (n_clumps, 3))
    which = np.arange(nc) % n_clumps
    cl = cen[which] + rng.normal(0.0, sigma, (nc, 3))
    return np.mod(np.concatenate([uni, cl]), 1.0)


def to_box_f32(p, L=1.0):
    """FP32 positions wrapped into [0, L) after rounding."""
    p = np.asarray(p, dtype=np.float32)
    p = np.where(p >= L, p - np.float32(L), p)
    p = np.where(p < 0, p + np.float32(L), p)
    return np.where(p >= L, np.float32(0.0), p).astype(np.float32)


def cfg1(seed=4738, n=65536):
    """Plummer cold collapse: N=65,536, a=0.05, centred, v=0, m=1/N."""
    rng = np.random.default_rng(seed)
    return to_box_f32(plummer(n, 0.05, rng)), np.zeros((n, 3), np.float32)


def cfg2(seed=4738, n=1 << 24, sigma_v=0.01):
    """50% uniform + 50% in 64 Gaussian clumps (sigma 0.02), Maxwellian velocities."""
    rng = np.random.default_rng(seed)
    pos = to_box_f32(clumpy(n, rng))
    vel = rng.normal(0.0, sigma_v, (n, 3)).astype(np.float32)
    return pos, vel


def targets(n, t, seed=4738):
    """Seeded target subset (sorted indices) for the FP64 oracle."""
    rng = np.random.default_rng(seed + 1)
    return np.sort(rng.choice(n, size=min(t, n), replace=False)).astype(np.int64)


def clumpy_device(n, device, seed=4738, n_clumps=4096, sigma=0.02, frac_uniform=0.5, sigma_v=0.01, chunk=1 << 26):
    """The clumpy mixture drawn on the GPU (torch generator, FP32, written in
    chunks): a fraction uniform in the unit box, the rest split equally over
    Gaussian clumps with uniform-random centres, wrapped periodically into
    [0, 1); Maxwellian velocities.  Returns (pos, vel), each (n, 3) FP32 on
    device.  Same distribution as clumpy(), not the same numbers."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    pos = torch.empty(n, 3, dtype=torch.float32, device=device)
    vel = torch.empty(n, 3, dtype=torch.float32, device=device)
    nu = int(round(n * frac_uniform))
    cen = torch.rand(n_clumps, 3, generator=g, device=device, dtype=torch.float64)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        m = b - a
        p = torch.rand(m, 3, generator=g, device=device, dtype=torch.float64)
        lo = max(a, nu)
        if lo < b:   # the clump part of this chunk
            idx = torch.arange(lo - nu, b - nu, device=device) % n_clumps
            p[lo - a:] = cen[idx] + sigma * torch.randn(b - lo, 3, generator=g, device=device, dtype=torch.float64)
        p = torch.remainder(p, 1.0).to(torch.float32)
        p = torch.where(p >= 1.0, torch.zeros_like(p), p)
        pos[a:b] = p
        vel[a:b] = sigma_v * torch.randn(m, 3, generator=g, device=device, dtype=torch.float32)
    return pos, vel
