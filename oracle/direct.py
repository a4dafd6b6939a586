"""FP64 direct-sum and Ewald oracles, energies and momenta (test infrastructure)."""
import math

import numpy as np
from scipy.special import erf, erfc


def direct_acc(pos, eps, mass=None, targets=None, chunk=2048):
    """Softened all-pairs acceleration, G=1, restating exact_acc
    (scripts/scatter_scaling_ops.py:56-71) and gravity_sim.accel
    (scripts/gravity_sim.py:6-11): a_i = sum_{j != i} m_j (x_j - x_i)
    (|x_j - x_i|^2 + eps^2)^-1.5.  ``targets`` (indices into pos) restricts the
    rows; the self pair is excluded by index.  pos (N, D) -> (T, D), FP64."""
    pos = np.asarray(pos, dtype=np.float64)
    N = pos.shape[0]
    m = np.ones(N) if mass is None else np.broadcast_to(np.asarray(mass, dtype=np.float64), (N,))
    tgt = np.arange(N) if targets is None else np.asarray(targets)
    out = np.empty((len(tgt), pos.shape[1]))
    for s in range(0, len(tgt), chunk):
        t = tgt[s:s + chunk]
        d = pos[None, :, :] - pos[t, None, :]
        with np.errstate(divide="ignore"):
            inv = ((d ** 2).sum(-1) + eps ** 2) ** -1.5
        inv[np.arange(len(t)), t] = 0.0
        out[s:s + chunk] = (d * (inv * m[None, :])[..., None]).sum(1)
    return out


def kdk_rollout(pos, vel, steps, dt=0.1, substeps=4, eps=0.5, g=1.0):
    """KDK leapfrog with substeps, restating gravity_sim.rollout
    (scripts/gravity_sim.py:31-44).  Returns (steps+1, N, D) positions, velocities."""
    h = dt / substeps
    ps, vs = [pos.copy()], [vel.copy()]
    pos, vel = pos.copy(), vel.copy()
    a = g * direct_acc(pos, eps)
    for _ in range(steps):
        for _ in range(substeps):
            vel += 0.5 * h * a
            pos += h * vel
            a = g * direct_acc(pos, eps)
            vel += 0.5 * h * a
        ps.append(pos.copy())
        vs.append(vel.copy())
    return np.stack(ps), np.stack(vs)


def energy(pos, vel, eps, mass=None):
    """Softened total energy, restating gravity_sim.energy (scripts/gravity_sim.py:14-18)
    with masses as in scatter_field.energy (scripts/scatter_field.py:759-765)."""
    pos, vel = np.asarray(pos, dtype=np.float64), np.asarray(vel, dtype=np.float64)
    N = len(pos)
    m = np.ones(N) if mass is None else np.broadcast_to(np.asarray(mass, dtype=np.float64), (N,))
    d = pos[None, :, :] - pos[:, None, :]
    r = np.sqrt((d ** 2).sum(-1) + eps ** 2)
    iu = np.triu_indices(N, 1)
    return 0.5 * (m * (vel ** 2).sum(-1)).sum() - (m[iu[0]] * m[iu[1]] / r[iu]).sum()


def momentum(vel, mass):
    """Total momentum, restating scatter_field.momentum (scripts/scatter_field.py:768-769)."""
    vel = np.asarray(vel, dtype=np.float64)
    m = np.broadcast_to(np.asarray(mass, dtype=np.float64), vel.shape[:-1])
    return (m[..., None] * vel).sum(-2)


def kinetic_energy(vel, mass):
    vel = np.asarray(vel, dtype=np.float64)
    m = np.broadcast_to(np.asarray(mass, dtype=np.float64), vel.shape[:-1])
    return 0.5 * (m * (vel ** 2).sum(-1)).sum(-1)


def ewald_params(L, alpha_l=8.0, kmax=None):
    """Split parameter alpha = alpha_l / L and reciprocal cut-off |n| <= kmax
    chosen so both truncation errors are below ~1e-8 of the force."""
    alpha = alpha_l / L
    if kmax is None:
        kmax = int(math.ceil(2 * alpha * math.sqrt(18.5) * L / (2 * math.pi))) + 1
    return alpha, kmax


def ewald_acc(pos, mass, targets, L=1.0, eps=0.0, alpha_l=8.0, kmax=None, nimg=1):
    """Periodic FP64 acceleration (no reference exists; the reference is
    isolated).  Solves the same problem as the periodic PM solve: G=1 gravity
    of point masses in a periodic cube of side L with the mean density removed.
    Real space: images |n|_inf <= nimg with erfc splitting; reciprocal space:
    |n_k|_inf <= kmax.  Plummer softening ``eps`` replaces the Newtonian
    minimum-image pair term: F_soft(r0) + (r0/|r0|^3)(erf(s) - 2s/sqrt(pi) e^-s^2),
    s = alpha |r0|, so softening changes only the nearest image.
    pos (N, 3) -> (T, 3).  O(T * N * (2 nimg+1)^3 + N * kmax^3)."""
    pos = np.asarray(pos, dtype=np.float64)
    N = pos.shape[0]
    m = np.broadcast_to(np.asarray(mass, dtype=np.float64), (N,))
    tgt = np.asarray(targets)
    alpha, kmax = ewald_params(L, alpha_l, kmax)
    V = L ** 3
    out = np.zeros((len(tgt), 3))
    sp = 2.0 / math.sqrt(math.pi)
    # real space
    for ti, i in enumerate(tgt):
        r0 = pos[i][None, :] - pos                     # vector from j to i
        r0 = r0 - L * np.round(r0 / L)                 # minimum image
        acc = np.zeros(3)
        rng = range(-nimg, nimg + 1)
        for nx in rng:
            for ny in rng:
                for nz in rng:
                    r = r0 + L * np.array([nx, ny, nz], dtype=np.float64)
                    rr = np.sqrt((r ** 2).sum(-1))
                    if nx == 0 and ny == 0 and nz == 0:
                        ok = np.arange(N) != i
                        rr_ = rr[ok]
                        s = alpha * rr_
                        g = erf(s) - sp * s * np.exp(-s * s)
                        soft = -r[ok] / ((rr_ ** 2 + eps ** 2) ** 1.5)[:, None]
                        acc += (m[ok][:, None] * (soft + r[ok] * (g / rr_ ** 3)[:, None])).sum(0)
                    else:
                        s = alpha * rr
                        f = (erfc(s) + sp * s * np.exp(-s * s)) / rr ** 3
                        acc += (m[:, None] * (-r * f[:, None])).sum(0)
        out[ti] = acc
    # reciprocal space
    n = np.arange(-kmax, kmax + 1)
    kv = np.stack(np.meshgrid(n, n, n, indexing="ij"), -1).reshape(-1, 3)
    kv = kv[(kv ** 2).sum(-1) > 0] * (2 * math.pi / L)
    k2 = (kv ** 2).sum(-1)
    fac = 4 * math.pi / V * np.exp(-k2 / (4 * alpha ** 2)) / k2
    for s in range(0, len(kv), 512):
        kk = kv[s:s + 512]
        ph = pos @ kk.T                                 # (N, K)
        Sc = (m[:, None] * np.cos(ph)).sum(0)
        Ss = (m[:, None] * np.sin(ph)).sum(0)
        pt = ph[tgt]                                    # (T, K)
        # sum_j m_j sin(k.(x_i - x_j)) = sin(k x_i) Sc - cos(k x_i) Ss
        sinsum = np.sin(pt) * Sc[None, :] - np.cos(pt) * Ss[None, :]
        out += -(sinsum * fac[s:s + 512][None, :]) @ kk
    return out
