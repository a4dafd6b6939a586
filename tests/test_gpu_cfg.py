"""GPU tests at the shapes of the benchmark configurations:

* the Plummer cold collapse (N = 65,536, 64^3, 100 KDK steps of dt = 1e-3)
  against the FP64 oracle's rollout;
* the clumpy N = 2^24, 256^3 scene, ten KDK steps of dt = 5e-4, against the
  same steps of the pinned PyTorch-CPU port;
* the surrogate network's conv tiles at 256^3 and 512^3 (every family and
  every zc up to 128, the tiles the autotuner picks there), each against the
  float64 conv on slabs of output planes, bits equal across zc within a family;
* the surrogate on that Plummer scene's 256^3 grid: the network step against
  its unfused form, and the per-step fallback to plain PM at the 2e-2
  residual bound.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return torch.device("cuda:0")


def _wrapdiff(a, b):
    return np.abs((a - b + 0.5) % 1.0 - 0.5)


def test_plummer_cold_collapse_100_steps(dev):
    """cfg1 end to end: 100 KDK steps through the stepper (device re-sorts as the
    sphere collapses) against the FP64 oracle.  FP32 vs FP64 roundoff grows
    through the collapse; measured on the CPU port: max |dx| 1.3e-5, velocity
    rel-RMS 4e-5 after 100 steps."""
    from augur_amd import scenes
    from augur_amd.pm import PeriodicPM
    from oracle import pm as opm
    pos, vel = scenes.cfg1()
    n, G, dt = len(pos), 64, 1e-3
    pm = PeriodicPM(G, 1.0, 1.0, dev)
    st = pm.load(torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev), mass_value=1.0 / n)
    P = pos.astype(np.float64)
    V = vel.astype(np.float64)
    done = 0
    for upto, tol_x, tol_v in ((10, 2e-6, 2e-5), (100, 1e-4, 5e-4)):
        pm.kdk(st, dt, upto - done)
        P, V = opm.kdk_pm_rollout(P, V, 1.0 / n, G, dt, upto - done)
        done = upto
        p, v, a = (t.cpu().double().numpy() for t in pm.unload(st))
        dx = _wrapdiff(p, P)
        relv = np.sqrt(((v - V) ** 2).sum() / (V ** 2).sum())
        print(f"step {upto}: max |dx| {dx.max():.2e}, velocity rel-RMS {relv:.2e}", flush=True)
        assert dx.max() < tol_x and relv < tol_v
        # Lagrangian radii of the collapsing sphere
        r, rr = np.linalg.norm(p - 0.5, axis=1), np.linalg.norm(P - 0.5, axis=1)
        for q in (0.1, 0.5, 0.9):
            assert abs(np.quantile(r, q) - np.quantile(rr, q)) < tol_x
    # the PM force is antisymmetric: total momentum stays at FP32 roundoff
    assert np.abs(v.sum(0)).max() / n < 1e-5 * np.sqrt((V ** 2).sum(1).mean())


def test_cfg2_shape_ten_kdk_steps(dev):
    """cfg2 end to end: N = 2^24 (50% uniform, 64 Gaussian clumps), 256^3, ten KDK
    steps of dt = 5e-4 through the stepper, against the same steps of the
    PyTorch-CPU FP32 port (pinned against the FP64 oracle's force and step);
    the two differ only by FP32 summation order in the deposit (measured: max
    |dx| 1.2e-7, velocity rel-RMS 1.9e-7).  The fused FP64 momentum sums drift
    by 2e-11 against a mean |v| of 2.4e-2."""
    from augur_amd import scenes
    from augur_amd.pm import PeriodicPM
    from oracle import torch_cpu as tc
    pos, vel = scenes.cfg2()
    n, G, dt, steps = len(pos), 256, 5e-4, 10
    pm = PeriodicPM(G, 1.0, 1.0, dev)
    st = pm.load(torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev), mass_value=1.0 / n)
    diag = pm.kdk(st, dt, steps, diag=True).cpu().numpy()
    p, v, _ = (t.cpu().numpy() for t in pm.unload(st))
    P, V = torch.from_numpy(pos), torch.from_numpy(vel)
    m = torch.full((n,), 1.0 / n)
    A = tc.pm_acc(P, m, G)
    for _ in range(steps):
        P, V, A = tc.kdk_step(P, V, A, m, G, dt)
    P, V = P.numpy(), V.numpy()
    dx = _wrapdiff(p, P).max()
    relv = np.sqrt(((v.astype(np.float64) - V) ** 2).sum() / (V.astype(np.float64) ** 2).sum())
    print(f"cfg2 after {steps} steps: max |dx| {dx:.2e}, velocity rel-RMS {relv:.2e}", flush=True)
    assert dx < 1e-6 and relv < 2e-6
    # momentum: sum m v over all particles after each step, against sum m |v|
    pmag = (np.abs(vel.astype(np.float64)).sum() / n)
    drift = np.abs(diag[:, :3] - diag[0, :3]).max()
    print(f"momentum drift over the steps {drift:.2e} (mean |v| {pmag:.2e})", flush=True)
    assert drift < 1e-8 * pmag
    assert np.isfinite(diag).all() and (diag[:, 3] > 0).all()


# ------------------------------------------------------------ conv at 256^3
def _slab_ref(x_cl, w_taps, bias, z):
    """float64 conv of output plane z from input planes z-1..z+1 (periodic in
    all axes).  x_cl: BF16 (1, G, G, G, cin) on the device."""
    G = x_cl.shape[1]
    idx = torch.tensor([(z - 1) % G, z, (z + 1) % G], device=x_cl.device)
    slab = x_cl[0].index_select(0, idx).double().cpu().permute(3, 0, 1, 2)[None]   # (1, cin, 3, G, G)
    slab = F.pad(slab, (1, 1, 1, 1, 0, 0), mode="circular")
    from augur_amd.surrogate import from_taps
    return F.conv3d(slab, from_taps(w_taps).double().cpu(), None if bias is None else bias.double().cpu())[0, :, 0]


@pytest.mark.parametrize("cin,mode,cout", [(4, 1, 32), (16, 1, 32), (32, 1, 32), (32, 2, 32), (32, 0, 3)])
def test_conv_tiles_at_256(dev, cin, mode, cout):
    from augur_amd import autotune as A
    from augur_amd import surrogate as S
    B, G = 1, 256
    g = torch.Generator(device=dev).manual_seed(7 * cin + mode)
    x = torch.randn(B, G, G, G, cin, device=dev, generator=g).to(torch.bfloat16)
    w = (0.1 * torch.randn(27, 32, cin, device=dev, generator=g)).to(torch.bfloat16).float()
    w[:, cout:] = 0
    b = 0.1 * torch.randn(32, device=dev, generator=g)
    wk = S.kernel_weights(w, cin)
    cands = A.candidates(cin, mode, cout, G)
    assert max(zc for _, zc in cands) == 128
    planes = (0, 97, 255)
    refs = {}
    for z in planes:
        r = _slab_ref(x, w, b, z)[:cout]                       # (cout, G, G)
        if mode in (1, 2):
            r = F.gelu(r, approximate="tanh")
            if mode == 2:
                r = r + x[0, z].double().cpu().permute(2, 0, 1)
        refs[z] = r
    first = {}
    for tile in cands:
        y = S.conv3_fwd(x, wk, b, mode, cout, tile=tile)
        for z in planes:
            got = (y[0, :, z] if mode == 0 else y[0, z].permute(2, 0, 1)).double().cpu()
            tol = (1e-4 if mode == 0 else 1.5e-2) * refs[z].abs().max()
            assert (got - refs[z]).abs().max() < tol, (tile, z)
        fam = tile[0]
        if fam in first:
            assert torch.equal(y, first[fam]), tile
        else:
            first[fam] = y.clone()
        del y


@pytest.mark.parametrize("cin,mode,cout", [(4, 1, 32), (32, 2, 32), (32, 0, 3)])
def test_conv_tiles_at_512(dev, cin, mode, cout):
    """The large workload's 512^3 grid (the surrogate leg of N = 2^30): the
    input, residual and output layers on every candidate tile, two planes each
    against the float64 conv, the checked planes bitwise equal across zc within
    a family."""
    from augur_amd import autotune as A
    from augur_amd import surrogate as S
    B, G = 1, 512
    g = torch.Generator(device=dev).manual_seed(11 * cin + mode)
    x = torch.randn(B, G, G, G, cin, device=dev, generator=g).to(torch.bfloat16)
    w = (0.1 * torch.randn(27, 32, cin, device=dev, generator=g)).to(torch.bfloat16).float()
    w[:, cout:] = 0
    b = 0.1 * torch.randn(32, device=dev, generator=g)
    wk = S.kernel_weights(w, cin)
    planes = (0, 301)
    refs = {}
    for z in planes:
        r = _slab_ref(x, w, b, z)[:cout]
        if mode in (1, 2):
            r = F.gelu(r, approximate="tanh")
            if mode == 2:
                r = r + x[0, z].double().cpu().permute(2, 0, 1)
        refs[z] = r
    first = {}
    for tile in A.candidates(cin, mode, cout, G):
        y = S.conv3_fwd(x, wk, b, mode, cout, tile=tile)
        got = {z: (y[0, :, z] if mode == 0 else y[0, z].permute(2, 0, 1)).double().cpu() for z in planes}
        del y
        for z in planes:
            tol = (1e-4 if mode == 0 else 1.5e-2) * refs[z].abs().max()
            assert (got[z] - refs[z]).abs().max() < tol, (tile, z)
        fam = tile[0]
        if fam in first:
            assert all(torch.equal(got[z], first[fam][z]) for z in planes), tile
        else:
            first[fam] = got


def test_conv_into_field_at_256(dev):
    """The output layer written into a force field (fused correction add):
    alpha (conv + bias) + field, alpha (conv + bias), and a skipped launch."""
    from augur_amd import _lib
    from augur_amd import autotune as A
    from augur_amd import surrogate as S
    from augur_amd.ops import _p, _stream
    B, G, cout = 1, 256, 3
    g = torch.Generator(device=dev).manual_seed(5)
    x = torch.randn(B, G, G, G, 32, device=dev, generator=g).to(torch.bfloat16)
    w = (0.1 * torch.randn(27, 32, 32, device=dev, generator=g)).to(torch.bfloat16).float()
    w[:, cout:] = 0
    b = 0.1 * torch.randn(32, device=dev, generator=g)
    wk = S.kernel_weights(w, 32)
    y = S.conv3_fwd(x, wk, b, S.MODE_OUT_CF, cout, tile=(A.TILE_NARROW, 16))
    field0 = torch.randn(B, cout, G, G, G, device=dev, generator=g)
    alpha = 0.375          # exact in FP32: field + alpha y is then exact in FP64, and fmaf rounds it once
    skip = torch.zeros(1, dtype=torch.int32, device=dev)
    for zc in (16, 128):
        f = field0.clone()
        _lib.call("augur_conv3_bf16_fwd_into", _p(x), 32, _p(wk), _p(b), B, G, cout, alpha, 1, _p(f), _p(skip), zc,
                  _stream(dev))
        exp = field0.double() + alpha * y.double()
        assert ((f.double() - exp).abs() <= 2.0 ** -24 * exp.abs() + 1e-38).all(), zc
        _lib.call("augur_conv3_bf16_fwd_into", _p(x), 32, _p(wk), _p(b), B, G, cout, alpha, 0, _p(f), None, zc,
                  _stream(dev))
        assert torch.equal(f, alpha * y)
    skip.fill_(1)
    f = field0.clone()
    _lib.call("augur_conv3_bf16_fwd_into", _p(x), 32, _p(wk), _p(b), B, G, cout, alpha, 1, _p(f), _p(skip), 0,
              _stream(dev))
    assert torch.equal(f, field0)


# ---------------------------------------------- surrogate on the Plummer grid
def _plummer_state(dev):
    from augur_amd import scenes
    pos, vel = scenes.cfg1()
    n = len(pos)
    return pos, vel, n


def test_surrogate_step_on_plummer_256(dev):
    """One surrogate step of the cfg1 scene at 256^3 (the packed 4-channel input
    and the output layer adding into the PM field) against the same step with
    the 16-channel features and a separate add; the net is non-zero."""
    from augur_amd.surrogate import PMSurrogate, pack_cl
    pos, vel, n = _plummer_state(dev)
    G, dt = 256, 1e-3
    sur = PMSurrogate(G, 1.0, 1.0, dev, mode="correction", init_std=0.02)
    p, v = torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev)
    a, b, c = (sur.load(p, v, mass_value=1.0 / n) for _ in range(3))
    gate = sur.kdk(a, dt, 1)
    assert not gate.fell_back
    sur.pm.kdk(c, dt, 1)            # plain PM: the correction is a.acc - c.acc
    # the unfused step: 16-channel features, FP32 net output, separate add
    sur.pm.step_begin(b, dt, True)
    rho, force = sur.pm.grids(b)
    x = sur.features_from_grids(rho, force)
    xc4 = sur.packed_features(rho, force, 1.0)
    y16 = sur.net.forward(x)[0]
    y4 = sur.net.forward_packed(xc4)[0]
    # the two input packings agree to BF16 resolution
    assert float((y4 - y16).norm() / y16.norm()) < 2e-2
    force.add_(y16 * sur.force_scale)
    sur.pm.step_end(b, dt, True)
    assert torch.equal(a.pos, b.pos) and torch.equal(a.ids, b.ids)
    ca, cb = (a.acc - c.acc).double(), (b.acc - c.acc).double()
    assert float(cb.norm() / c.acc.double().norm()) > 1e-3      # the correction is not negligible
    assert float((ca - cb).norm() / cb.norm()) < 3e-2


def test_replace_net_falls_back_at_2e2_on_plummer_256(dev):
    """The replacement net (density -> force field, cfg3's architecture)
    untrained is far from PM: the first step's residual exceeds 2e-2 and the
    call runs plain PM bit for bit, with no host synchronisation inside."""
    from augur_amd.surrogate import PMSurrogate
    pos, vel, n = _plummer_state(dev)
    G, dt = 256, 1e-3
    sur = PMSurrogate(G, 1.0, 1.0, dev, mode="replace", init_std=0.05)
    assert sur.check_threshold == 2e-2
    p, v = torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev)
    a, b = sur.load(p, v, mass_value=1.0 / n), sur.load(p, v, mass_value=1.0 / n)
    probe = torch.arange(0, n, 16, device=dev)
    gate = sur.kdk(a, dt, 3, probe=probe)
    sur.pm.kdk(b, dt, 3)
    assert gate.fell_back and gate.fallback_step == 0
    res = gate.residuals
    assert res[0] > 2e-2 and res[1] == 0.0 and res[2] == 0.0
    for ta, tb in ((a.pos, b.pos), (a.vel, b.vel), (a.acc, b.acc)):
        assert torch.equal(ta, tb)
