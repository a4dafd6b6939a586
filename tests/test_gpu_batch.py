"""GPU tests of batched training on ragged scenes: the ragged deposit (all
scenes in one launch) against per-scene deposits, the batch gradient against
the sum of single-scene gradients, and the batch gradient step captured in a
HIP graph and replayed with new batches against the same steps run eagerly."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return torch.device("cuda:0")


def scenes_np(seed, sizes):
    from augur_amd import scenes
    rng = np.random.default_rng(seed)
    out = []
    for k, n in enumerate(sizes):
        if k % 2:
            p = scenes.plummer(n, float(rng.uniform(0.05, 0.1)), rng)
        else:
            p = scenes.clumpy(n, rng, n_clumps=8, sigma=0.05)
        out.append(scenes.to_box_f32(p))
    return out


def test_ragged_deposit_matches_per_scene(dev):
    from augur_amd import ops
    from augur_amd.surrogate import ragged_batch
    G, sizes = 32, [1000, 7000, 20011]
    sc = scenes_np(1, sizes)
    pos, starts, mass = ragged_batch(sc, dev)
    rho = ops.cic_deposit_ragged(pos, starts, G, 1.0 / G, 0.0, True, mass=mass)[:, 0]
    for b, p in enumerate(sc):
        ref = ops.cic_deposit(torch.from_numpy(p)[None].to(dev), None, G, 1.0 / G, 0.0, True,
                              mass_value=1.0 / len(p))[0, 0]
        assert (rho[b] - ref).abs().max() <= 1e-6 * ref.abs().max(), b
        assert abs(float(rho[b].double().sum()) / G ** 3 - 1.0) < 1e-6


def make_batch(sur, sc, T, seed, dev):
    """ragged batch, targets and smooth stand-in target forces (a PM field
    gathered at the targets, scaled): only the gradient plumbing is tested"""
    from augur_amd import ops
    from augur_amd.surrogate import ragged_batch, _batch_grids
    pos, starts, mass = ragged_batch(sc, dev)
    rng = np.random.default_rng(seed)
    tidx = [int(starts[b]) + rng.choice(len(p), T, replace=False) for b, p in enumerate(sc)]
    tpos = torch.stack([pos[torch.from_numpy(t).to(dev)] for t in tidx]).contiguous()
    _, force = _batch_grids(sur, pos, starts, mass)
    a_true = (1.1 * ops.cic_gather(force, tpos, sur.h, 0.0, True)).double()
    inv_norm = (1.0 / (a_true ** 2).sum()).reshape(1)
    dens = torch.ones(len(sc), dtype=torch.float32, device=dev)
    return pos, starts, mass, tpos, a_true, inv_norm, dens


def test_batch_gradient_is_sum_of_scene_gradients(dev):
    from augur_amd.surrogate import PMSurrogate, batch_grads, _batch_grids
    G, T, sizes = 32, 256, [3000, 9000, 17000]
    sc = scenes_np(2, sizes)
    sur = PMSurrogate(G, device=dev, mode="correction", init_std=0.05)
    batch = make_batch(sur, sc, T, 3, dev)
    rho, force = _batch_grids(sur, *batch[:3])
    sur.calibrate(rho[0], force[0])
    loss, grads = batch_grads(sur, *batch)
    gsum, lsum = None, 0.0
    pos, starts, mass, tpos, a_true, inv_norm, dens = batch
    for b in range(len(sc)):
        lo, hi = int(starts[b]), int(starts[b + 1])
        st1 = torch.tensor([0, hi - lo], dtype=torch.int64, device=dev)
        l1, g1 = batch_grads(sur, pos[lo:hi].contiguous(), st1, mass[lo:hi].contiguous(), tpos[b:b + 1].contiguous(),
                             a_true[b:b + 1].contiguous(), inv_norm, dens[b:b + 1].contiguous())
        lsum += float(l1)
        g1 = [(dw.clone(), db.clone()) for dw, db in g1]
        gsum = g1 if gsum is None else [(a + c, b_ + d) for (a, b_), (c, d) in zip(gsum, g1)]
    assert abs(float(loss) - lsum) <= 1e-6 * abs(lsum)
    for (dw, db), (sw, sb) in zip(grads, gsum):
        assert (dw - sw).abs().max() <= 1e-4 * sw.abs().max() + 1e-12
        assert (db - sb).abs().max() <= 1e-4 * sb.abs().max() + 1e-12


def test_batch_graph_replays_match_eager(dev):
    from augur_amd.surrogate import PMSurrogate, BatchGraph, batch_grads, _batch_grids
    G, T, sizes = 32, 128, [2000, 6000, 11000, 4000]
    s1 = PMSurrogate(G, device=dev, mode="correction", init_std=0.05)
    s2 = PMSurrogate(G, device=dev, mode="correction", init_std=0.05)
    batches = [make_batch(s1, scenes_np(10 + k, sizes), T, 20 + k, dev) for k in range(3)]
    rho, force = _batch_grids(s1, *batches[0][:3])
    for s in (s1, s2):
        s.calibrate(rho[0], force[0])
    gr = BatchGraph(s2, *batches[0])
    for k in range(6):
        b = batches[k % 3]
        l1, g1 = batch_grads(s1, *b)
        s1.net.adam_step(g1, lr=1e-3, loss=l1)
        l2 = gr.step(*b, lr=1e-3)
        torch.cuda.synchronize()
        assert torch.equal(l1, l2), k
        for w1, w2 in zip(s1.net.w + s1.net.b, s2.net.w + s2.net.b):
            assert torch.equal(w1, w2), k


# ------------------------------------------------- batched ragged stepping
def _ragged_state(sizes, seed, dev):
    rng = np.random.default_rng(seed)
    total = int(sum(sizes))
    pos = torch.from_numpy(rng.random((total, 3), dtype=np.float32)).to(dev)
    vel = torch.from_numpy((0.05 * rng.standard_normal((total, 3))).astype(np.float32)).to(dev)
    acc = torch.from_numpy(rng.standard_normal((total, 3)).astype(np.float32)).to(dev)
    starts = torch.tensor([0] + list(np.cumsum(sizes)), dtype=torch.int64, device=dev)
    return pos, vel, acc, starts


def test_persistent_step_matches_single_scene_launches(dev):
    """Every phase of the ragged launch gives each scene the bits it gets
    stepped alone (empty and one-particle scenes included); the gather equals
    the generic CIC gather of the scene's own field, the kicks and drift the
    KDK formulas."""
    from augur_amd import batch as RB
    from augur_amd import ops
    G, L, dt = 32, 1.0, 2e-3
    sizes = [1000, 0, 3001, 1, 2500, 0]
    pos, vel, acc0, starts = _ragged_state(sizes, 11, dev)
    g = torch.Generator(device=dev).manual_seed(1)
    force = torch.randn(len(sizes), 3, G, G, G, device=dev, generator=g)
    out = {}
    for phase in (RB.PHASE_GATHER, RB.PHASE_OPEN, RB.PHASE_CLOSE, RB.PHASE_CLOSE_OPEN):
        p, v, a = pos.clone(), vel.clone(), acc0.clone()
        RB.persistent_step(p, v, a, starts, force, G, L, dt, phase)
        out[phase] = (p, v, a)
        lo = 0
        for b, n in enumerate(sizes):
            sl = slice(lo, lo + n)
            lo += n
            if n == 0:
                continue
            pb, vb, ab = pos[sl].clone(), vel[sl].clone(), acc0[sl].clone()
            RB.persistent_step(pb, vb, ab, torch.tensor([0, n], dtype=torch.int64, device=dev),
                               force[b:b + 1].contiguous(), G, L, dt, phase)
            assert torch.equal(p[sl], pb) and torch.equal(v[sl], vb) and torch.equal(a[sl], ab), (phase, b)
    # the gather is each scene's CIC gather of its own field
    a = out[RB.PHASE_GATHER][2]
    lo = 0
    for b, n in enumerate(sizes):
        sl = slice(lo, lo + n)
        lo += n
        if n:
            ref = ops.cic_gather(force[b:b + 1], pos[sl][None].contiguous(), 1.0 / G, 0.0, True)[0]
            assert (a[sl] - ref).abs().max() <= 1e-6 * ref.abs().max()
    a_g = a.double()
    hdt = 0.5 * dt
    # OPEN: v + dt/2 acc, then x + dt v wrapped; CLOSE: v + dt/2 a, acc = a
    p1, v1, _ = out[RB.PHASE_OPEN]
    v_exp = vel.double() + hdt * acc0.double()
    assert (v1.double() - v_exp).abs().max() < 1e-6
    dx = ((p1.double() - pos.double() - dt * v1.double()) + 0.5) % 1.0 - 0.5
    assert dx.abs().max() < 1e-6 and bool(((p1 >= 0) & (p1 < 1)).all())
    p2, v2, a2 = out[RB.PHASE_CLOSE]
    assert torch.equal(p2, pos) and torch.equal(a2, a)
    assert (v2.double() - (vel.double() + hdt * a_g)).abs().max() < 1e-6
    # CLOSE_OPEN = CLOSE followed by OPEN, bit for bit
    p3, v3, _ = out[RB.PHASE_CLOSE_OPEN]
    pc, vc, ac = p2.clone(), v2.clone(), a2.clone()
    RB.persistent_step(pc, vc, ac, starts, None, G, L, dt, RB.PHASE_OPEN)
    assert torch.equal(pc, p3) and torch.equal(vc, v3)


def test_ragged_kdk_matches_oracle_per_scene(dev):
    """A KDK rollout of three scenes of uneven sizes stepped as one batch
    against the FP64 oracle's PM rollout of each scene alone."""
    from augur_amd import batch as RB
    from augur_amd.surrogate import ragged_batch
    from oracle import pm as opm
    G, dt, steps = 32, 1e-3, 3
    sizes = [3000, 700, 5000]
    sc = scenes_np(7, sizes)
    rng = np.random.default_rng(8)
    vels = [(0.02 * rng.standard_normal((n, 3))).astype(np.float32) for n in sizes]
    pos, starts, mass = ragged_batch(sc, dev)
    vel = torch.from_numpy(np.concatenate(vels)).to(dev)
    st = RB.RaggedStepper(G, device=dev)
    acc = st.accel(pos, starts, mass)
    lo = 0
    for p, n in zip(sc, sizes):
        ref = opm.pm_acc_3d_periodic(p, 1.0 / n, G)
        a = acc[lo: lo + n].cpu().numpy()
        assert np.sqrt(((a - ref) ** 2).sum() / (ref ** 2).sum()) < 1e-5
        lo += n
    st.kdk(pos, vel, acc, starts, mass, dt, steps)
    lo = 0
    for p, v, n in zip(sc, vels, sizes):
        rp, rv = opm.kdk_pm_rollout(p, v, 1.0 / n, G, dt, steps)
        dp = np.abs(((pos[lo: lo + n].cpu().numpy() - rp) + 0.5) % 1.0 - 0.5).max()
        dv = np.abs(vel[lo: lo + n].cpu().numpy() - rv).max() / np.abs(rv).max()
        assert dp < 1e-5 and dv < 1e-4, (n, dp, dv)
        lo += n


def test_planned_deposit_resorts_on_drift_and_keeps_the_bits(dev):
    """The stepping loop's planned ragged deposit sorts on the first step of
    a call and re-sorts on the device only when the tracked drift may pass a
    cell: fast particles (about 0.4 cells per step) force re-sorts inside a
    12-step call, and the call gives the bits of 12 one-step calls (every
    deposit freshly sorted), with empty and one-particle scenes."""
    from augur_amd import batch as RB
    G, dt, steps = 64, 2e-3, 12
    sizes = [6000, 0, 2500, 1]
    pos, vel, acc, starts = _ragged_state(sizes, 41, dev)
    vel = vel * 60.0                                         # |v| ~ 3: 0.4 cells per step at G = 64
    total = pos.shape[0]
    mass = torch.from_numpy(np.random.default_rng(42).uniform(0.5, 1.5, total).astype(np.float32)).to(dev)
    mass /= total
    st = RB.RaggedStepper(G, device=dev)
    acc = st.accel(pos, starts, mass)
    P, V, A = pos.clone(), vel.clone(), acc.clone()
    st.kdk(pos, vel, acc, starts, mass, dt, steps)
    for _ in range(steps):
        st.kdk(P, V, A, starts, mass, dt, 1)
    assert torch.equal(pos, P) and torch.equal(vel, V) and torch.equal(acc, A)
    # the long call re-sorted on the way: the drift accumulated since its last
    # sort (0 if the last step re-sorted) is below a cell, where 12 steps
    # without one would be ~5 cells
    plan = st.plan(total, starts)
    off = plan.drift_word - plan.ws.data_ptr() - 4        # PmDev: cum_cells just before the drift word
    st.kdk(pos, vel, acc, starts, mass, dt, steps)
    torch.cuda.synchronize(dev)
    cum = float(plan.ws[off: off + 4].view(torch.float32)[0])
    assert 0.0 <= cum < 1.0, cum


@pytest.mark.parametrize("with_net", [False, True])
def test_ragged_kdk_graph_replays_match_eager(dev, with_net):
    """A 2-step call of the batched stepper (plain PM, or PM with the
    correction net at batch B) captured once in a HIP graph: three replays,
    each from a new state with eager work between, equal eager calls bit for
    bit; the batched surrogate field of a scene matches the scene alone."""
    from augur_amd import batch as RB
    from augur_amd.surrogate import PMSurrogate, ragged_batch
    G, dt = 32, 1e-3
    sizes = [2000, 333, 4100, 64]
    sc = scenes_np(9, sizes)
    pos, starts, mass = ragged_batch(sc, dev)
    sur = None
    if with_net:
        sur = PMSurrogate(G, 1.0, 1.0, dev, mode="correction", init_std=0.05)
        rho, force = RB.RaggedStepper(G, device=dev).grids(pos, starts, mass)
        sur.calibrate(rho[0], force[0])
    st = RB.RaggedStepper(G, device=dev, surrogate=sur)
    rng = np.random.default_rng(10)
    vel = torch.from_numpy((0.02 * rng.standard_normal((pos.shape[0], 3))).astype(np.float32)).to(dev)
    acc = st.accel(pos, starts, mass)
    if with_net:   # scene 2 of the batch against the same scene alone
        f_b = st.force_field(pos, starts, mass)[2]
        f_1, _ = sur.field(torch.from_numpy(sc[2]).to(dev), mass_value=1.0 / sizes[2])
        assert float((f_b - f_1).pow(2).sum().sqrt() / f_1.pow(2).sum().sqrt()) < 1e-3
    P, V, A = pos.clone(), vel.clone(), acc.clone()
    graph = st.capture(P, V, A, starts, mass, dt, 2)
    state = (pos, vel, acc)
    for r in range(3):
        P.copy_(state[0])
        V.copy_(state[1])
        A.copy_(state[2])
        graph.replay()
        torch.cuda.synchronize()
        p, v, a = (t.clone() for t in state)
        st.kdk(p, v, a, starts, mass, dt, 2)
        torch.cuda.synchronize()
        assert torch.equal(P, p) and torch.equal(V, v) and torch.equal(A, a), r
        state = (p, v, a)


def test_ragged_sort_keeps_scenes_and_bits(dev):
    """Sorting a ragged batch into cell order permutes particles only within
    their scenes, and stepping the sorted batch gives every particle the bits
    it gets unsorted (the fixed-point deposit does not depend on order)."""
    from augur_amd import batch as RB
    from augur_amd.surrogate import ragged_batch
    G, dt = 32, 1e-3
    sizes = [2000, 333, 4100]
    sc = scenes_np(12, sizes)
    pos, starts, mass = ragged_batch(sc, dev)
    rng = np.random.default_rng(13)
    vel = torch.from_numpy((0.02 * rng.standard_normal((pos.shape[0], 3))).astype(np.float32)).to(dev)
    st = RB.RaggedStepper(G, device=dev)
    perm, P, V, M = st.sort(pos, starts, vel, mass)
    n = pos.shape[0]
    assert torch.equal(torch.sort(perm).values, torch.arange(n, device=dev))
    scene = torch.searchsorted(starts[1:].contiguous(), torch.arange(n, device=dev), right=True)
    assert torch.equal(scene[perm], scene)
    cell = torch.floor(P.double() * G - 0.5).long().remainder(G)
    key = ((scene * G + cell[:, 2]) * G + cell[:, 1]) * G + cell[:, 0]
    assert bool((key[1:] >= key[:-1]).all())
    A0 = st.accel(pos, starts, mass)
    A1 = st.accel(P, starts, M)
    assert torch.equal(A1, A0[perm])
    st.kdk(pos, vel, A0, starts, mass, dt, 2)
    st.kdk(P, V, A1, starts, M, dt, 2)
    assert torch.equal(P, pos[perm]) and torch.equal(V, vel[perm]) and torch.equal(A1, A0[perm])


def test_batched_solve_is_each_grids_solve(dev):
    """augur_pm_solve_grids: every grid of the batch gets the bits the
    single-grid solve gives it, and the planes match the oracle's periodic
    Poisson solve with spectral gradient and the generic rfftn path."""
    from augur_amd import _lib, ops
    from augur_amd.ops import _p, _stream, _ws
    from oracle import pm as opm
    G, B, L = 32, 3, 1.0
    g = torch.Generator(device=dev).manual_seed(5)
    rho = torch.rand(B, G, G, G, device=dev, generator=g) * 3.0
    f = ops.pm_solve_grids(rho, L, 1.0)
    ws = _ws(_lib.size("augur_pm_workspace_bytes", 1, G), dev)
    for b in range(B):
        one = torch.empty(3, G, G, G, dtype=torch.float32, device=dev)
        _lib.call("augur_pm_solve_grid", _p(rho[b].contiguous()), _p(one), G, L, 1.0, _p(ws), ws.numel(),
                  _stream(dev))
        assert torch.equal(f[b], one), b
        ref, _ = opm.poisson_periodic_accel_grids(rho[b].double().cpu().numpy(), L)
        assert opm.rel_rms(f[b].double().cpu().numpy(), ref) < 1e-5
    spec = ops.rfftn(rho, s=(G, G, G))
    gen = ops.irfftn(ops.poisson_green_grad(spec, G, L, 1.0, scale=1.0 / G ** 3), s=(G, G, G), scale=1.0)
    rel = float(((f - gen).double().pow(2).sum() / gen.double().pow(2).sum()).sqrt())
    assert rel < 1e-5, rel


@pytest.mark.parametrize("G", [32, 64])
def test_tiled_ragged_deposit_is_the_atomic_deposit_bitwise(dev, G):
    """The ragged deposit on the tiled LDS kernels (scene-major tiles) gives
    the global-atomic fixed-point deposit's bits: mixed scenes with per-particle
    masses, an empty and a one-particle scene, particles on the box faces, and
    a dense clump whose tiles take several work items (the int64 image path)."""
    from augur_amd import ops
    rng = np.random.default_rng(31 + G)
    sc = scenes_np(7, [3000, 9000])
    clump = (0.5 + 0.004 * rng.standard_normal((40000, 3))).astype(np.float32) % np.float32(1.0)
    edge = rng.uniform(0, 1, (500, 3)).astype(np.float32)
    edge[:100, 0] = 0.0
    edge[100:200, 1] = np.float32(1.0) - np.float32(2.0 ** -24)
    sc = [sc[0], np.zeros((0, 3), np.float32), clump, edge, rng.uniform(0, 1, (1, 3)).astype(np.float32), sc[1]]
    sizes = [len(p) for p in sc]
    pos = torch.from_numpy(np.concatenate(sc).astype(np.float32)).to(dev)
    starts = torch.tensor([0] + list(np.cumsum(sizes)), dtype=torch.int64, device=dev)
    mass = torch.from_numpy(rng.uniform(0.5, 2.0, pos.shape[0]).astype(np.float32) / 1000.0).to(dev)
    h = 1.0 / G
    for m in (mass, None):
        a = ops.cic_deposit_ragged(pos, starts, G, h, 0.0, True, mass=m, mass_value=0.37, tiled=True)
        b = ops.cic_deposit_ragged(pos, starts, G, h, 0.0, True, mass=m, mass_value=0.37, tiled=False)
        assert a.shape == b.shape == (len(sc), 1, G, G, G)
        assert torch.equal(a, b)
        assert float(a[1].abs().max()) == 0.0
