"""Pins for oracle.pm: each check is chosen so that a plausible mistake in the
restatement (a dropped term, wrong sign, transposed axis, floor vs round,
wrong wrap) fails it.  None of these re-types the oracle's own formula."""
import numpy as np
import pytest

from oracle import pm

F32 = np.float32


# ------------------------------------------------------------------ a1 keys
@pytest.mark.parametrize("D", [2, 3])
def test_cell_keys_on_exact_boundaries(D):
    # positions exactly on cell walls (exact in FP32) land in the upper cell;
    # one ulp below lands in the lower one; outside is clamped; masked -> nc.
    extent, C = 32.0, 8
    cell = extent / C
    rng = np.random.default_rng(1)
    k = rng.integers(0, C, (1, 200, D))
    wall = (-extent / 2 + k * cell).astype(F32)
    below = np.nextafter(wall, F32(-np.inf)).astype(F32)
    mask = np.ones((1, 200))
    key_w, _ = pm.cell_keys(wall, mask, extent, cell, C)
    key_b, _ = pm.cell_keys(below, mask, extent, cell, C)
    exp_w = sum(k[0, :, d] * C ** d for d in range(D))
    assert (key_w == exp_w).all()
    # one ulp below a wall: the FP32 add pos + extent/2 may round back up onto
    # the wall.  Expected cell from the exact FP64 sum rounded once to FP32
    # (a correctly rounded FP32 add) and an exact power-of-two divide.
    s = (below.astype(np.float64) + extent / 2).astype(F32).astype(np.float64)
    kb = np.clip(np.floor(s / cell), 0, C - 1).astype(np.int64)
    exp_b = sum(kb[0, :, d] * C ** d for d in range(D))
    assert (key_b == exp_b).all()
    neg = (k[0] * cell <= extent / 4).all(1)         # there |pos| >= |pos + extent/2|: the add is exact
    assert (key_b[neg] == sum(np.clip(k[0, neg, d] - 1, 0, C - 1) * C ** d for d in range(D))).all()
    far = np.full((1, 3, D), 1e6, F32)
    key_f, _ = pm.cell_keys(far, np.array([[1, 0, 1]]), extent, cell, C)
    assert key_f[0] == C ** D - 1 and key_f[1] == C ** D and key_f[2] == C ** D - 1


def test_stable_sort_matches_bruteforce():
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 17, 500)
    perm = pm.stable_sort_by_cell(keys)
    assert perm.tolist() == sorted(range(500), key=lambda i: (keys[i], i))


def test_cic_base_centre_and_wall():
    h, off = 0.5, 16.0
    k = np.arange(-3, 60)
    centre = (-off + (k + 0.5) * h).astype(F32)[None, :, None]
    wall = (-off + k * h).astype(F32)[None, :, None]
    i0, w = pm.cic_base(centre, off, h)
    assert (i0[0, :, 0] == k).all() and (w == 0).all()
    i0, w = pm.cic_base(wall, off, h)
    assert (i0[0, :, 0] == k - 1).all() and (w == 0.5).all()


# --------------------------------------------------------------- a2 deposit
@pytest.mark.parametrize("D,periodic", [(2, False), (2, True), (3, True), (3, False)])
def test_deposit_matches_tent_bruteforce(D, periodic):
    rng = np.random.default_rng(3)
    G, h = 8, 0.25
    off = 0.0 if periodic else G * h / 2
    lo, hi = (0.0, G * h) if periodic else (-G * h / 2 + 0.3, G * h / 2 - 0.3)
    pos = rng.uniform(lo, hi, (2, 40, D)).astype(F32)
    mass = rng.uniform(0.5, 2.0, (2, 40))
    rho = pm.cic_deposit(pos, mass, np.ones((2, 40)), G, h, off, periodic)
    ref = pm.cic_deposit_bruteforce(pos, mass, G, h, off, periodic)
    # the oracle's weights are FP32 (as the reference computes them), the
    # brute force uses FP64 positions: agreement to FP32 weight rounding
    np.testing.assert_allclose(rho, ref, rtol=2e-5, atol=1e-6 * ref.max())
    np.testing.assert_allclose(rho.reshape(2, -1).sum(1) * h ** D, mass.sum(1), rtol=1e-12)


def test_deposit_velocity_channels_and_open_drop():
    # one particle exactly on a cell centre: all of m, m*v in that cell;
    # a particle beyond the open edge contributes nothing.
    G, extent = 8, 8.0
    h = extent / G
    pos = np.array([[[-4 + 2.5 * h, -4 + 5.5 * h], [100.0, 0.0]]], F32)
    vel = np.array([[[2.0, -3.0], [1.0, 1.0]]])
    out = pm.cic_deposit(pos, np.ones((1, 2)), np.ones((1, 2)), G, h, extent / 2, False, vel=vel, in_scale=2.0)
    s = 2.0 / h ** 2
    assert out[0, 0, 5, 2] == pytest.approx(s) and out[0, 1, 5, 2] == pytest.approx(2 * s)
    assert out[0, 2, 5, 2] == pytest.approx(-3 * s)
    assert np.count_nonzero(out) == 3


# ---------------------------------------------------------------- a6 gather
def test_gather_reproduces_linear_field_open_2d():
    # reference test style: tests/test_staggered_double_grid.py:43-52
    G, extent = 32, 16.0
    h = extent / G
    rng = np.random.default_rng(4)
    pos = rng.uniform(-5, 5, (1, 50, 2)).astype(F32)
    x = (np.arange(G) + 0.5) * h - extent / 2
    f = (2.0 * x[None, :] + 3.0 * x[:, None] + 1.0)[None, None]
    g = pm.cic_gather(f, pos, h, extent / 2, False)[0, :, 0]
    np.testing.assert_allclose(g, 2.0 * pos[0, :, 0] + 3.0 * pos[0, :, 1] + 1.0, atol=1e-5)


def test_gather_is_adjoint_of_deposit_periodic_3d():
    rng = np.random.default_rng(5)
    G, h = 8, 1.0 / 8
    pos = rng.uniform(0, 1, (1, 64, 3)).astype(F32)
    m = rng.uniform(0.1, 1.0, (1, 64))
    f = rng.normal(size=(1, 1, G, G, G))
    rho = pm.cic_deposit(pos, m, np.ones((1, 64)), G, h, 0.0, True)
    lhs = (rho * f).sum() * h ** 3
    rhs = (m[0] * pm.cic_gather(f, pos, h, 0.0, True)[0, :, 0]).sum()
    assert lhs == pytest.approx(rhs, rel=1e-12)


# -------------------------------------------------------------- a4/a5 2D open
def test_open_potential_equals_direct_convolution():
    rng = np.random.default_rng(6)
    G, h = 12, 0.5
    rho = rng.uniform(0, 1, (2, 1, G, G))
    np.testing.assert_allclose(pm.open_potential_2d(rho, h), pm.open_potential_2d_bruteforce(rho, h), rtol=1e-10, atol=1e-12)


def test_fd_gradient_of_quadratic():
    G, h = 16, 0.5
    x = np.arange(G) * h
    phi = (x[None, :] ** 2 + 3 * x[:, None])[None, None]
    g = pm.neg_grad_fd(phi, h)
    np.testing.assert_allclose(g[0, 0, 1:-1, 1:-1], -2 * np.broadcast_to(x[None, 1:-1], (G - 2, G - 2)), atol=1e-12)
    np.testing.assert_allclose(g[0, 1, 1:-1, 1:-1], -3.0, atol=1e-12)


def test_pm_acc_2d_newton_third_law_interior():
    # CIC deposit/gather adjointness plus an even kernel gives sum m a = 0
    # (docs/debugging/mesh-momentum-investigation.md:5-16)
    rng = np.random.default_rng(7)
    pos = rng.normal(0, 3, (1, 200, 2)).astype(F32)
    a = pm.pm_acc_2d_open(pos, np.ones((1, 200)), np.ones((1, 200)), 64, 32.0)
    assert np.abs(a.sum(1)).max() < 1e-10 * np.abs(a).sum()


# ------------------------------------------------------- a4/a5 3D periodic
@pytest.mark.parametrize("n", [(1, 0, 0), (0, 2, 0), (1, 1, 3), (3, 0, 2)])
def test_periodic_poisson_single_mode(n):
    G, L = 16, 2.0
    z, y, x = np.meshgrid(*[np.arange(G) * L / G] * 3, indexing="ij")
    k = 2 * np.pi * np.array(n) / L
    A = 0.7
    ph = k[0] * x + k[1] * y + k[2] * z
    rho = 3.0 + A * np.cos(ph)
    acc, phi = pm.poisson_periodic_accel_grids(rho, L, Gconst=1.3)
    k2 = (k ** 2).sum()
    np.testing.assert_allclose(phi, -4 * np.pi * 1.3 * A * np.cos(ph) / k2, atol=1e-10)
    for d in range(3):
        np.testing.assert_allclose(acc[d], -4 * np.pi * 1.3 * A * k[d] * np.sin(ph) / k2, atol=1e-10)


def test_periodic_gradient_nyquist_zeroed():
    G, L = 8, 1.0
    z, y, x = np.meshgrid(*[np.arange(G) * L / G] * 3, indexing="ij")
    rho = np.cos(np.pi * G * x)          # pure Nyquist mode along x
    acc, phi = pm.poisson_periodic_accel_grids(rho, L)
    assert np.abs(acc).max() < 1e-12 and np.abs(phi).max() > 1e-3


# --------------------------------------------------------------- a13 metric
def test_errs_hand_example():
    at = np.array([[[3.0, 4.0], [0.0, 0.0], [6.0, 8.0]]])
    a = np.array([[[3.0, 5.0], [1.0, 0.0], [6.0, 8.0]]])
    mask = np.array([[1.0, 0.0, 1.0]])
    e, rel = pm.errs(a, at, mask)
    rms = np.sqrt((25 + 100) / 2)
    np.testing.assert_allclose(e, [1 / rms, 0.0])
    np.testing.assert_allclose(rel, [0.2, 0.0])
    s = pm.stats(np.array([1.0, 2.0, 3.0, 4.0]))
    assert s == {"median": 2.5, "p90": pytest.approx(3.7), "mean": 2.5}
