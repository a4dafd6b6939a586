"""Seeded 3D scene generators (test infrastructure; the product has its own
copies in augur_amd.scenes, and tests check the two agree).

The reference's samplers are 2D (scripts/pm_baseline_force_error.py:28-65,
scripts/nbody_ic.py:62-97); these are the textbook 3D forms.
"""
import numpy as np


def plummer(n, a, rng, center=(0.5, 0.5, 0.5), rmax=0.45):
    """Plummer sphere: r = a / sqrt(u^(-2/3) - 1), isotropic direction; radii
    >= rmax are redrawn so the sphere stays inside the unit box."""
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
    """Unit periodic box: frac_uniform of the particles uniform, the rest split
    equally over n_clumps Gaussian clumps (uniform-random centres), wrapped."""
    nu = int(round(n * frac_uniform))
    nc = n - nu
    uni = rng.uniform(0.0, 1.0, (nu, 3))
    cen = rng.uniform(0.0, 1.0, (n_clumps, 3))
    which = np.arange(nc) % n_clumps
    cl = cen[which] + rng.normal(0.0, sigma, (nc, 3))
    return np.mod(np.concatenate([uni, cl]), 1.0)
