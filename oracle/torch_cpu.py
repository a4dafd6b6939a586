"""PyTorch-CPU FP32 port of the plain periodic PM step (the CPU baseline that
bench.py times beside the HIP path) and FP64 references for the surrogate
network (test infrastructure).

The PM port restates the reference's own PyTorch formulation
(index_add_ CIC deposit, scripts/scatter_field.py:237-250; FFT Poisson,
scripts/pm_baseline_force_error.py:68-76; CIC gather, scatter_field.py:252-261;
KDK, scripts/gravity_sim.py:31-44) generalised to 3D periodic with the
Green's function -4 pi G / k^2 and a spectral gradient.
"""
import math

import torch
import torch.nn.functional as F


def _corners(pos, G):
    u = pos * G - 0.5
    i0 = torch.floor(u)
    return i0.long(), u - i0


def deposit(pos, mass, G):
    """pos (N,3) in [0,1), mass (N,) -> rho (G,G,G) [z,y,x], mass per cell volume."""
    i0, w = _corners(pos, G)
    rho = torch.zeros(G * G * G, dtype=pos.dtype)
    val = mass * G ** 3
    for oz in (0, 1):
        for oy in (0, 1):
            for ox in (0, 1):
                ix = (i0[:, 0] + ox) % G
                iy = (i0[:, 1] + oy) % G
                iz = (i0[:, 2] + oz) % G
                wt = (w[:, 0] if ox else 1 - w[:, 0]) * (w[:, 1] if oy else 1 - w[:, 1]) * (w[:, 2] if oz else 1 - w[:, 2])
                rho.index_add_(0, (iz * G + iy) * G + ix, val * wt)
    return rho.view(G, G, G)


def accel_grids(rho, L=1.0, Gc=1.0):
    G = rho.shape[-1]
    rk = torch.fft.rfftn(rho)
    f = torch.fft.fftfreq(G, d=1.0 / G, dtype=rho.dtype) * (2 * math.pi / L)
    kx = (torch.arange(G // 2 + 1, dtype=rho.dtype) * (2 * math.pi / L))
    kz, ky = f[:, None, None], f[None, :, None]
    kx = kx[None, None, :]
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k2[0, 0, 0] = 1.0
    phik = -4 * math.pi * Gc * rk / k2
    phik[0, 0, 0] = 0
    kxn, kyn, kzn = kx.clone(), ky.clone(), kz.clone()
    kxn[..., G // 2] = 0
    kyn[:, G // 2] = 0
    kzn[G // 2] = 0
    return torch.stack([torch.fft.irfftn(-1j * k * phik, s=(G, G, G)) for k in (kxn, kyn, kzn)])


def gather(acc, pos, G):
    i0, w = _corners(pos, G)
    fl = acc.reshape(3, -1)
    out = torch.zeros_like(pos)
    for oz in (0, 1):
        for oy in (0, 1):
            for ox in (0, 1):
                ix = (i0[:, 0] + ox) % G
                iy = (i0[:, 1] + oy) % G
                iz = (i0[:, 2] + oz) % G
                wt = (w[:, 0] if ox else 1 - w[:, 0]) * (w[:, 1] if oy else 1 - w[:, 1]) * (w[:, 2] if oz else 1 - w[:, 2])
                out += fl[:, (iz * G + iy) * G + ix].T * wt[:, None]
    return out


def pm_acc(pos, mass, G):
    return gather(accel_grids(deposit(pos, mass, G)), pos, G)


def kdk_step(pos, vel, acc, mass, G, dt):
    """One KDK step in the merged form used by the HIP stepper:
    v += dt/2 a; x += dt v (wrapped); a = PM(x); v += dt/2 a."""
    vel = vel + 0.5 * dt * acc
    pos = torch.remainder(pos + dt * vel, 1.0)
    acc = pm_acc(pos, mass, G)
    vel = vel + 0.5 * dt * acc
    return pos, vel, acc


def conv3d_periodic(x, w, b=None):
    """Periodic ('circular') 3^3 convolution in FP64: x (B,Cin,G,G,G), w (Cout,Cin,3,3,3)."""
    return F.conv3d(F.pad(x, (1, 1, 1, 1, 1, 1), mode="circular"), w, b)


def localnet3d(x, weights, biases):
    """LocalNet topology (scripts/scatter_field.py:15-26) in 3D with periodic padding:
    h = gelu(conv_in(x)); h = h + gelu(conv(h)) for each middle layer; out = conv_out(h)."""
    h = F.gelu(conv3d_periodic(x, weights[0], biases[0]), approximate='tanh')
    for w, b in zip(weights[1:-1], biases[1:-1]):
        h = h + F.gelu(conv3d_periodic(h, w, b), approximate='tanh')
    return conv3d_periodic(h, weights[-1], biases[-1])


def pm_acc_2d(pos, mass, G, extent, eps=0.5):
    """The reference's 2D plain PM force (scripts/pm_baseline_force_error.py:68-76)
    on the CPU: CIC deposit with off-grid corners dropped (scripts/scatter_field.py:
    190-211,237-250), zero-padded rfft2 convolution with -1/sqrt(r^2 + eps^2) on
    the (2G)^2 domain times h^2, crop, zero-padded central differences
    (:551-555), CIC gather (:252-261).  pos (N, 2) FP32 on a grid covering
    [-extent/2, extent/2)^2, mass (N,) -> (N, 2)."""
    h = extent / G
    u = (pos + extent / 2) / h - 0.5
    i0 = torch.floor(u)
    w = u - i0
    i0 = i0.long()
    corners = []
    for ox in (0, 1):
        for oy in (0, 1):
            ix, iy = i0[:, 0] + ox, i0[:, 1] + oy
            wt = (w[:, 0] if ox else 1 - w[:, 0]) * (w[:, 1] if oy else 1 - w[:, 1])
            ok = ((ix >= 0) & (ix < G) & (iy >= 0) & (iy < G)).to(pos.dtype)
            corners.append((iy.clamp(0, G - 1) * G + ix.clamp(0, G - 1), wt * ok))
    rho = torch.zeros(G * G, dtype=pos.dtype)
    val = mass / h ** 2
    for flat, wt in corners:
        rho.index_add_(0, flat, val * wt)
    rho = rho.view(G, G)
    idx = torch.arange(2 * G, dtype=pos.dtype)
    d = torch.minimum(idx, 2 * G - idx) * h
    K = -1.0 / torch.sqrt(d[:, None] ** 2 + d[None, :] ** 2 + eps ** 2)
    full = torch.fft.irfft2(torch.fft.rfft2(rho, s=(2 * G, 2 * G)) * torch.fft.rfft2(K), s=(2 * G, 2 * G)) * h ** 2
    p = F.pad(full[:G, :G][None, None], (1, 1, 1, 1))[0, 0]
    gx = -(p[1:-1, 2:] - p[1:-1, :-2]) / (2 * h)
    gy = -(p[2:, 1:-1] - p[:-2, 1:-1]) / (2 * h)
    acc = torch.stack([gx.reshape(-1), gy.reshape(-1)], -1)
    out = torch.zeros_like(pos)
    for flat, wt in corners:
        out += acc[flat] * wt[:, None]
    return out
