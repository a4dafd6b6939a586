"""GPU tests of the learned force path: the split KDK step it runs on, the
adjoint of the CIC gather it trains through, the gain of the trained
correction over plain PM against the FP64 Ewald oracle, and the fallback to
plain PM when the residual check fails."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return torch.device("cuda:0")


def clumpy_scene(seed, n, dev, K=None, G=64):
    """Seeded clumpy scene; with K, K target indices and their FP64 Ewald
    accelerations (softening one cell)."""
    from augur_amd import ops, scenes
    rng = np.random.default_rng(seed)
    nc = int(rng.integers(4, 32))
    p = scenes.to_box_f32(scenes.clumpy(n, rng, n_clumps=nc, sigma=float(rng.uniform(0.02, 0.06))))
    pos = torch.from_numpy(p).to(dev)
    if K is None:
        return pos
    tg = torch.from_numpy(np.sort(rng.choice(n, K, replace=False))).to(dev)
    return pos, tg, ops.ewald_f64(pos, 1.0 / n, tg, L=1.0, eps=1.0 / G)


def test_pm_step_halves_equal_kdk(dev):
    # step_begin / step_end in a loop are the same launches as augur_pm_kdk
    from augur_amd.pm import PeriodicPM
    n, G = 20000, 64
    pos = clumpy_scene(1, n, dev)
    vel = 0.5 * torch.randn(n, 3, device=dev)
    pm = PeriodicPM(G, 1.0, 1.0, dev)
    a, b = pm.load(pos, vel), pm.load(pos, vel)
    pm.kdk(a, 1e-3, 5)
    for s in range(5):
        pm.step_begin(b, 1e-3, s == 0)
        pm.step_end(b, 1e-3, s == 4)
    for x, y in ((a.pos, b.pos), (a.vel, b.vel), (a.acc, b.acc), (a.ids, b.ids)):
        assert torch.equal(x, y)


def test_grids_view_holds_pm_force(dev):
    # the workspace views after step_begin equal the functional PM fields
    from augur_amd.pm import PeriodicPM
    n, G = 20000, 64
    pos = clumpy_scene(2, n, dev)
    pm = PeriodicPM(G, 1.0, 1.0, dev)
    st = pm.load(pos, torch.zeros_like(pos))
    pm.step_begin(st, 0.0, True)          # dt = 0: no motion
    rho, force = pm.grids(st)
    rho2, force2 = pm.fields(st.pos, mass_value=1.0 / n)
    assert torch.equal(rho, rho2)
    assert (force - force2).abs().max() <= 1e-6 * force2.abs().max()


def test_gather_adjoint_is_the_deposit(dev):
    # <gather(F), g> = <F, deposit(g)>: the training gradient path
    from augur_amd import ops
    G, n = 32, 5000
    gen = torch.Generator(device="cpu").manual_seed(5)
    F = torch.randn(1, 3, G, G, G, generator=gen).to(dev)
    pos = torch.rand(1, n, 3, generator=gen).to(dev)
    g = torch.randn(1, n, 3, generator=gen).to(dev)
    h = 1.0 / G
    lhs = (ops.cic_gather(F, pos, h, 0.0, True).double() * g.double()).sum()
    dF = ops.cic_deposit(pos, None, G, h, 0.0, True, vel=g, scale=1.0, mass_value=1.0)[:, 1:4]
    rhs = (F.double() * dF.double()).sum()
    assert abs(float(lhs - rhs)) <= 1e-5 * float((F.double().abs() * dF.double().abs()).sum())


def test_trained_correction_beats_plain_pm(dev):
    # trained on the GPU by the project's own forward/backward kernels
    # against FP64 Ewald forces, the correction lowers the force rel-RMS of
    # held-out scenes well below plain PM's at the same grid
    from augur_amd.surrogate import PMSurrogate, rel_rms
    G, n, K = 64, 1 << 15, 4096
    train = [clumpy_scene(s, n, dev, K) for s in range(100, 106)]
    test = [clumpy_scene(s, n, dev, K) for s in range(900, 902)]
    sur = PMSurrogate(G, device=dev, mode="correction")
    losses = [sur.train_step(*train[it % len(train)], mass_value=1.0 / n, lr=1e-3) for it in range(300)]
    assert np.isfinite(losses).all()
    for pos, tg, at in test:
        f, force = sur.field(pos, mass_value=1.0 / n)
        pt = pos[tg].contiguous()
        e_pm, e_sur = rel_rms(sur.gather(force, pt), at), rel_rms(sur.gather(f, pt), at)
        assert e_sur < 0.6 * e_pm, (e_sur, e_pm)


def test_untrained_correction_is_plain_pm(dev):
    # the output layer starts at zero: the surrogate step is the PM step bitwise
    from augur_amd.surrogate import PMSurrogate
    n, G = 20000, 64
    pos = clumpy_scene(3, n, dev)
    vel = 0.3 * torch.randn(n, 3, device=dev)
    sur = PMSurrogate(G, device=dev, mode="correction")
    a, b = sur.load(pos, vel), sur.load(pos, vel)
    sur.pm.kdk(a, 1e-3, 4)
    fell_back, res = sur.kdk(b, 1e-3, 4, probe=torch.arange(0, n, 7, device=dev))
    assert not fell_back and res == 0.0
    assert torch.equal(a.pos, b.pos) and torch.equal(a.vel, b.vel)


def test_residual_check_falls_back_to_plain_pm(dev):
    # random (untrained, non-zero) weights move the forces far from PM: the
    # probe residual exceeds the threshold and the step is plain PM
    from augur_amd.surrogate import PMSurrogate
    n, G = 20000, 64
    pos = clumpy_scene(4, n, dev)
    vel = 0.3 * torch.randn(n, 3, device=dev)
    sur = PMSurrogate(G, device=dev, mode="replace", init_std=0.5)
    probe = torch.arange(0, n, 5, device=dev)
    a, fell_back, res = sur.accel(pos, mass_value=1.0 / n, probe=probe)
    assert fell_back and res > sur.check_threshold
    a_pm = sur.gather(sur.pm.fields(pos, mass_value=1.0 / n)[1], pos)
    assert torch.equal(a, a_pm)
    s1, s2 = sur.load(pos, vel), sur.load(pos, vel)
    sur.pm.kdk(s1, 1e-3, 3)
    fell_back, res = sur.kdk(s2, 1e-3, 3, probe=probe)
    assert fell_back
    assert torch.equal(s1.pos, s2.pos) and torch.equal(s1.vel, s2.vel)


def test_train_step_on_grids_equals_train_step(dev):
    """train_step_grids on grids made once gives the loss and the weights that
    train_step (which makes the grids itself) gives, bit for bit."""
    from augur_amd.surrogate import PMSurrogate
    n, G = 30000, 64
    pos, tg, a_ex = clumpy_scene(5, n, dev, K=512, G=G)
    s1 = PMSurrogate(G, 1.0, 1.0, dev, mode="correction", init_std=0.05)
    s2 = PMSurrogate(G, 1.0, 1.0, dev, mode="correction", init_std=0.05)
    rho, force = s2.pm.fields(pos, mass_value=1.0 / n)
    pt = pos[tg].contiguous()
    for _ in range(3):
        l1 = s1.train_step(pos, tg, a_ex, mass_value=1.0 / n, lr=1e-3)
        l2 = s2.train_step_grids(rho, force, pt, a_ex, lr=1e-3)
        assert l1 == l2
    for w1, w2 in zip(s1.net.parameters(), s2.net.parameters()):
        assert torch.equal(w1, w2)


def test_correction_fallback_trips_at_a_later_step(dev):
    """Correction mode with a non-zero net: a per-step threshold that passes
    step 0 and fails step 1 (forced from the host) makes the call the
    surrogate's step 0 followed by plain PM, bit for bit; the gate records the
    step and every residual, on the device."""
    from augur_amd.surrogate import PMSurrogate
    n, G, dt = 20000, 64, 1e-3
    pos = clumpy_scene(6, n, dev)
    vel = 0.3 * torch.randn(n, 3, device=dev)
    sur = PMSurrogate(G, device=dev, mode="correction", init_std=0.02)
    probe = torch.arange(0, n, 3, device=dev)
    a, b = sur.load(pos, vel), sur.load(pos, vel)
    g0 = sur.kdk(a, dt, 1, probe=probe, threshold=float("inf"))
    sur.pm.kdk(a, dt, 2)
    gate = sur.kdk(b, dt, 3, probe=probe, threshold=[float("inf"), 0.0, 0.0])
    assert not g0.fell_back and g0.residual > 0
    assert gate.fell_back and gate.fallback_step == 1
    res = gate.residuals
    assert res[0] == g0.residual and res[1] > 0 and res[2] == 0.0
    for x, y in ((a.pos, b.pos), (a.vel, b.vel), (a.acc, b.acc)):
        assert torch.equal(x, y)


def test_kdk_with_gate_is_capturable(dev):
    """A 2-step surrogate call with its device-side residual check captured in
    a HIP graph: every replay equals the eager call on the same state."""
    from augur_amd.surrogate import PMSurrogate
    n, G, dt = 20000, 64, 1e-3
    pos = clumpy_scene(7, n, dev)
    vel = 0.3 * torch.randn(n, 3, device=dev)
    sur = PMSurrogate(G, device=dev, mode="correction", init_std=0.02)
    probe = torch.arange(0, n, 3, device=dev)
    st = sur.load(pos, vel)
    sur.kdk(st, dt, 2, probe=probe)          # warm-up: calibration, allocations
    torch.cuda.synchronize()
    saved = [t.clone() for t in (st.pos, st.vel, st.acc, st.ids, st.ws)]
    eager = sur.kdk(st, dt, 2, probe=probe)
    want = [t.clone() for t in (st.pos, st.vel, st.acc)]
    want_res = eager.residuals
    graph = torch.cuda.CUDAGraph()
    for t, s in zip((st.pos, st.vel, st.acc, st.ids, st.ws), saved):
        t.copy_(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(graph):
        cap = sur.kdk(st, dt, 2, probe=probe)
    for _ in range(3):
        for t, s in zip((st.pos, st.vel, st.acc, st.ids, st.ws), saved):
            t.copy_(s)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for x, y in zip((st.pos, st.vel, st.acc), want):
            assert torch.equal(x, y)
        assert cap.residuals == want_res


@pytest.mark.parametrize("reuse,steps", [(3, 3), (2, 4)])
def test_correction_reuse_adds_the_stored_correction(dev, reuse, steps):
    """kdk(reuse=k): the net runs on steps 0, k, 2k, ... only and its stored
    correction is added to the PM force planes of the steps in between; the
    call equals that sequence done step by step with the stepper's pieces, bit
    for bit, and the device residual check still logs every step."""
    from augur_amd.surrogate import PMSurrogate
    n, G, dt = 20000, 64, 1e-3
    pos = clumpy_scene(8, n, dev)
    vel = 0.3 * torch.randn(n, 3, device=dev)
    sur = PMSurrogate(G, device=dev, mode="correction", init_std=0.02)
    rho, force = sur.pm.fields(pos, mass_value=1.0 / n)
    sur.calibrate(rho, force)
    a, b = sur.load(pos, vel), sur.load(pos, vel)
    probe = torch.arange(0, n, 3, device=dev)
    gate = sur.kdk(b, dt, steps, probe=probe, threshold=float("inf"), reuse=reuse)
    corr = torch.empty(3, G, G, G, dtype=torch.float32, device=dev)
    md = sur._mass_total(a) / sur.L ** 3
    for s in range(steps):
        sur.pm.step_begin(a, dt, s == 0)
        rho, force = sur.pm.grids(a)
        if s % reuse == 0:
            sur.net.forward_into(sur.packed_features(rho, force, md), corr, sur.force_scale, accumulate=False)
        force.add_(corr)
        sur.pm.step_end(a, dt, s == steps - 1)
    assert not gate.fell_back and all(r > 0 for r in gate.residuals)
    for x, y in ((a.pos, b.pos), (a.vel, b.vel), (a.acc, b.acc)):
        assert torch.equal(x, y)
