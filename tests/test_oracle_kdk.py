"""Pins for the oracle's KDK rollout (oracle.pm.kdk_pm_rollout), the checker of
every GPU stepping test, and for the PyTorch-CPU port that the CPU baseline
times (oracle.torch_cpu).

Each check is one a plausible mistake fails: a dropped or doubled half kick
breaks time reversibility or the harmonic invariant, a wrong kick/drift order
differs from the integrator pinned by the reference's own rollout fixture
(tests/golden/truth_ref.json, through oracle.direct.kdk_rollout), a wrong wrap
breaks the free drift, a wrong force scale fails the two-body Ewald check.
``test_checks_catch_planted_mistakes`` plants such mistakes in a copy of the
integrator and shows the checks reject them."""
import numpy as np
import torch

from augur_amd import scenes
from oracle import direct
from oracle import pm as opm
from oracle import torch_cpu as tc


def _kdk(pos, vel, force, dt, steps, L, closing=0.5, opening=0.5, drift_first=False):
    """A copy of the integrator with knobs to plant mistakes (defaults = correct)."""
    pos, vel = pos.copy(), vel.copy()
    a = force(pos)
    for _ in range(steps):
        if drift_first:
            pos = np.mod(pos + dt * vel, L)
            vel += opening * dt * a
        else:
            vel += opening * dt * a
            pos = np.mod(pos + dt * vel, L)
        a = force(pos)
        vel += closing * dt * a
    return pos, vel


def _scene(n=256, seed=11):
    rng = np.random.default_rng(seed)
    pos = scenes.to_box_f32(scenes.clumpy(n, rng, n_clumps=4, sigma=0.08)).astype(np.float64)
    vel = 0.05 * rng.normal(size=(n, 3))
    return pos, vel


def _reversal_error(integrate, pos, vel, steps):
    p1, v1 = integrate(pos, vel, steps)
    p2, v2 = integrate(p1, -v1, steps)
    dp = np.abs((p2 - pos + 0.5) % 1.0 - 0.5).max()
    return dp, np.abs(-v2 - vel).max()


def test_free_drift_on_uniform_lattice():
    # one particle per cell at the same offset: the CIC density is uniform, so the
    # PM force vanishes and KDK is a pure drift, wrapped into [0, 1)
    G = 16
    g = (np.arange(G) + 0.25) / G
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    v = np.array([1 / 8, -3 / 16, 5 / 32])
    vel = np.broadcast_to(v, pos.shape).copy()
    dt, steps = 1 / 64, 9          # every increment exact in FP32 and FP64
    a0 = opm.pm_acc_3d_periodic(pos, 1.0 / len(pos), G)
    assert np.abs(a0).max() < 1e-10
    p, vv = opm.kdk_pm_rollout(pos, vel, 1.0 / len(pos), G, dt, steps)
    expect = np.mod(pos + steps * dt * v, 1.0)
    assert np.abs(p - expect).max() < 1e-10
    assert ((p >= 0) & (p < 1)).all()
    assert np.abs(vv - vel).max() < 1e-10


def test_kdk_time_reversible():
    pos, vel = _scene()
    m, G = 1.0 / len(pos), 16
    dp, dv = _reversal_error(lambda p, v, k: opm.kdk_pm_rollout(p, v, m, G, 2e-3, k), pos, vel, 6)
    assert dp < 1e-11 and dv < 1e-9


def test_kdk_same_steps_as_the_pinned_direct_integrator():
    # oracle.direct.kdk_rollout reproduces the reference's truth_ref.json; with the
    # same force and substeps=1 the PM rollout must take bit-identical steps
    rng = np.random.default_rng(3)
    pos = rng.uniform(10.0, 20.0, (40, 3))
    vel = rng.normal(size=(40, 3))
    eps, dt = 0.5, 0.05
    p, v = opm.kdk_pm_rollout(pos, vel, None, None, dt, 7, L=1e6, force=lambda x: direct.direct_acc(x, eps))
    ps, vs = direct.kdk_rollout(pos, vel, 7, dt=dt, substeps=1, eps=eps)
    assert np.array_equal(p, ps[-1]) and np.array_equal(v, vs[-1])


def test_kdk_harmonic_invariant():
    # velocity Verlet on a = -w^2 (x - c) conserves v^2 + w^2 (x - c)^2 (1 - (w dt / 2)^2) exactly
    w, dt, c = 3.0, 0.07, 500.0
    x0 = c + np.array([[1.0, -0.5, 0.25]])
    v0 = np.array([[0.3, 0.8, -1.1]])
    inv = lambda x, v: (v ** 2 + w ** 2 * (x - c) ** 2 * (1 - (w * dt / 2) ** 2)).sum()
    x, v = opm.kdk_pm_rollout(x0, v0, None, None, dt, 200, L=1e6, force=lambda p: -w ** 2 * (p - c))
    assert abs(inv(x, v) - inv(x0, v0)) < 1e-10 * inv(x0, v0)
    # and the exact discrete orbit: x_n - c = (x_0 - c) cos(n t) + dt v_0 sin(n t) / sin t, cos t = 1 - (w dt)^2 / 2
    th = np.arccos(1 - (w * dt) ** 2 / 2)
    np.testing.assert_allclose(x - c, (x0 - c) * np.cos(200 * th) + dt * v0 * np.sin(200 * th) / np.sin(th),
                               atol=1e-11)


def test_pm_force_scale_against_ewald():
    # two equal masses 16 cells apart: periodic PM within 1.5% of the FP64 Ewald force
    pos = np.array([[0.375, 0.5, 0.5], [0.625, 0.5, 0.5]])
    a_pm = opm.pm_acc_3d_periodic(pos, 0.5, 64)
    a_ew = direct.ewald_acc(pos, np.full(2, 0.5), [0, 1])
    assert np.abs(a_pm - a_ew).max() < 0.015 * np.abs(a_ew).max()
    assert a_pm[0, 0] > 0 and a_pm[1, 0] < 0


def test_checks_catch_planted_mistakes():
    pos, vel = _scene()
    m, G = 1.0 / len(pos), 16
    f = lambda p: opm.pm_acc_3d_periodic(p, m, G)
    ok = _reversal_error(lambda p, v, k: _kdk(p, v, f, 2e-3, k, 1.0), pos, vel, 6)
    assert ok[0] < 1e-11
    dropped = _reversal_error(lambda p, v, k: _kdk(p, v, f, 2e-3, k, 1.0, closing=0.0), pos, vel, 6)
    assert dropped[1] > 1e-6
    swapped = _reversal_error(lambda p, v, k: _kdk(p, v, f, 2e-3, k, 1.0, drift_first=True), pos, vel, 6)
    assert max(swapped) > 1e-8
    # doubled kicks stay reversible but break the harmonic invariant
    w, dt, c = 3.0, 0.07, 500.0
    x0, v0 = c + np.array([[1.0, -0.5, 0.25]]), np.array([[0.3, 0.8, -1.1]])
    inv = lambda x, v: (v ** 2 + w ** 2 * (x - c) ** 2 * (1 - (w * dt / 2) ** 2)).sum()
    x, v = _kdk(x0, v0, lambda p: -w ** 2 * (p - c), dt, 200, 1e6, closing=1.0, opening=1.0)
    assert abs(inv(x, v) - inv(x0, v0)) > 1e-3 * inv(x0, v0)
    # a wrap into [0, 2) instead of [0, 1) breaks the free drift
    g = (np.arange(16) + 0.25) / 16
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    vv = np.broadcast_to(np.array([1 / 8, -3 / 16, 5 / 32]), lat.shape).copy()
    p_bad, _ = _kdk(lat, vv, lambda p: np.zeros_like(p), 1 / 64, 9, 2.0)
    assert np.abs(p_bad - np.mod(lat + 9 / 64 * vv, 1.0)).max() > 0.5


# ------------------------------------------------------- the PyTorch-CPU port
def _port_scene(n=4096, seed=5):
    rng = np.random.default_rng(seed)
    pos = scenes.to_box_f32(scenes.clumpy(n, rng, n_clumps=8, sigma=0.05))
    vel = (0.02 * rng.normal(size=(n, 3))).astype(np.float32)
    return pos, vel


def test_torch_cpu_port_force_matches_oracle():
    pos, _ = _port_scene()
    n, G = len(pos), 32
    a_port = tc.pm_acc(torch.from_numpy(pos), torch.full((n,), 1.0 / n), G).double().numpy()
    a_ref = opm.pm_acc_3d_periodic(pos, 1.0 / n, G)
    assert opm.rel_rms(a_port, a_ref) < 1e-5


def test_torch_cpu_port_step_matches_oracle():
    pos, vel = _port_scene()
    n, G, dt = len(pos), 32, 2e-3
    m = torch.full((n,), 1.0 / n)
    p = torch.from_numpy(pos)
    a = tc.pm_acc(p, m, G)
    p1, v1, a1 = tc.kdk_step(p, torch.from_numpy(vel), a, m, G, dt)
    rp, rv = opm.kdk_pm_rollout(pos, vel, 1.0 / n, G, dt, 1)
    assert np.abs((p1.double().numpy() - rp + 0.5) % 1.0 - 0.5).max() < 1e-6
    assert opm.rel_rms(v1.double().numpy(), rv) < 1e-5
    assert opm.rel_rms(a1.double().numpy(), opm.pm_acc_3d_periodic(rp, 1.0 / n, G)) < 1e-4


def test_torch_cpu_port_2d_matches_oracle():
    # the 2D reference-shape plain PM (scripts/pm_baseline_force_error.py:68-76)
    rng = np.random.default_rng(8)
    n, G, ext = 2000, 64, 32.0
    pos = rng.normal(0.0, 4.0, (1, n, 2)).astype(np.float32)
    a_port = tc.pm_acc_2d(torch.from_numpy(pos[0]), torch.ones(n), G, ext).double().numpy()
    a_ref = opm.pm_acc_2d_open(pos, np.ones((1, n)), np.ones((1, n)), G, ext)[0]
    assert opm.rel_rms(a_port, a_ref) < 1e-5


def test_conv3d_periodic_is_the_brute_force_correlation():
    # the FP64 reference of every GPU conv test: out[co, z, y, x] = b[co] + sum over
    # ci and taps (dz, dy, dx) of w[co, ci, dz, dy, dx] x[ci, z+dz-1, y+dy-1, x+dx-1]
    # (indices mod G); a flipped kernel or a wrong tap axis fails it
    g = torch.Generator().manual_seed(4)
    G, ci, co = 4, 2, 3
    x = torch.randn(1, ci, G, G, G, generator=g, dtype=torch.float64)
    w = torch.randn(co, ci, 3, 3, 3, generator=g, dtype=torch.float64)
    b = torch.randn(co, generator=g, dtype=torch.float64)
    y = tc.conv3d_periodic(x, w, b)
    ref = torch.zeros_like(y)
    for z in range(G):
        for yy in range(G):
            for xx in range(G):
                for dz in range(3):
                    for dy in range(3):
                        for dx in range(3):
                            v = x[0, :, (z + dz - 1) % G, (yy + dy - 1) % G, (xx + dx - 1) % G]
                            ref[0, :, z, yy, xx] += w[:, :, dz, dy, dx] @ v
    ref += b.view(1, co, 1, 1, 1)
    torch.testing.assert_close(y, ref, rtol=1e-12, atol=1e-12)


def test_localnet3d_composition():
    # identity input and output convs, zero middle convs with bias 0.3: every
    # residual block adds gelu(0.3), so y = gelu(x) + (layers - 2) gelu(0.3)
    # exactly (tanh GELU); a block without its residual, or one too many, fails
    C, G, layers = 3, 4, 5
    eye = torch.zeros(C, C, 3, 3, 3, dtype=torch.float64)
    eye[range(C), range(C), 1, 1, 1] = 1.0
    ws = [eye] + [torch.zeros(C, C, 3, 3, 3, dtype=torch.float64) for _ in range(layers - 2)] + [eye]
    bs = [torch.zeros(C, dtype=torch.float64)] + [torch.full((C,), 0.3, dtype=torch.float64)
                                                  for _ in range(layers - 2)] + [torch.zeros(C, dtype=torch.float64)]
    x = torch.randn(1, C, G, G, G, generator=torch.Generator().manual_seed(6), dtype=torch.float64)
    y = tc.localnet3d(x, ws, bs)
    gel = lambda t: torch.nn.functional.gelu(t, approximate="tanh")
    torch.testing.assert_close(y, gel(x) + (layers - 2) * gel(torch.tensor(0.3, dtype=torch.float64)))
