"""GPU tests of the compact-state large-N driver (augur_amd.large.CompactPM):
with FP32 state it is the oracle's KDK rollout; its density is bitwise the PM
stepper's; the chunk size changes no bits; compact encodings track the FP32
state; the grid-weighted mean acceleration is the particles' mean."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return torch.device("cuda:0")


def _scene(n, seed=21):
    from augur_amd import scenes
    rng = np.random.default_rng(seed)
    pos = scenes.to_box_f32(scenes.clumpy(n, rng, n_clumps=16, sigma=0.04))
    vel = (0.3 * rng.normal(size=(n, 3))).astype(np.float32)
    return pos, vel


def _wrapdiff(a, b):
    return np.abs((a - b + 0.5) % 1.0 - 0.5)


def test_f32_state_is_the_kdk_rollout(dev):
    from augur_amd.large import CompactPM
    from oracle import pm as opm
    n, G, dt, steps = 20000, 64, 1e-3, 5
    pos, vel = _scene(n)
    d = CompactPM(G, device=dev, pos_dtype="f32", vel_dtype="f32", momfix=True)
    st = d.load(torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev))
    d.step(st, dt, steps)
    d.finish(st)
    p, v = (t.double().cpu().numpy() for t in d.decode(st))
    rp, rv = opm.kdk_pm_rollout(pos, vel, 1.0 / n, G, dt, steps)
    assert _wrapdiff(p, rp).max() < 1e-6
    assert opm.rel_rms(v, rv) < 1e-5


def test_density_is_the_steppers_bitwise(dev):
    from augur_amd.large import CompactPM
    from augur_amd.pm import PeriodicPM
    n, G = 30000, 64
    pos, vel = _scene(n, 22)
    d = CompactPM(G, device=dev, pos_dtype="f32", vel_dtype="f32", chunk=4099)
    st = d.load(torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev))
    d._field(st, 0.0)
    rho = PeriodicPM(G, 1.0, 1.0, dev).density(torch.from_numpy(pos).to(dev), mass_value=1.0 / n)
    assert torch.equal(d.density(), rho)


def test_chunking_changes_no_bits(dev):
    from augur_amd.large import CompactPM
    n, G = 25000, 64
    pos, vel = _scene(n, 23)
    out = []
    for chunk in (1 << 26, 7919):
        d = CompactPM(G, device=dev, pos_dtype="fix32", vel_dtype="f16", chunk=chunk)
        st = d.load(torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev))
        d.step(st, 1e-3, 3)
        out.append((st.pos.clone(), st.vel.clone()))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])


# FP16 velocities carry 2^-11 relative error, BF16 2^-8; a step moves a particle by dt M,
# so the positions follow the velocity error times dt (the fastest particles set the max)
@pytest.mark.parametrize("pd,vd,tol_x,tol_v", [("fix32", "f16", 5e-5, 2e-3), ("fix32", "bf16", 5e-4, 1.5e-2),
                                               ("fix16", "f16", 1e-4, 5e-3)])
def test_compact_state_tracks_f32(dev, pd, vd, tol_x, tol_v):
    from augur_amd.large import CompactPM
    from oracle import pm as opm
    n, G, dt, steps = 20000, 64, 2e-3, 5
    pos, vel = _scene(n, 24)
    res = []
    for p_, v_ in (("f32", "f32"), (pd, vd)):
        d = CompactPM(G, device=dev, pos_dtype=p_, vel_dtype=v_)
        st = d.load(torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev))
        d.step(st, dt, steps)
        d.finish(st)
        res.append([t.double().cpu().numpy() for t in d.decode(st)])
        assert st.state_bytes == n * 3 * (st.pos.element_size() + st.vel.element_size())
    (p0, v0), (p1, v1) = res
    print(f"{pd}/{vd}: max |dx| {_wrapdiff(p1, p0).max():.2e}, velocity rel-RMS {opm.rel_rms(v1, v0):.2e}",
          flush=True)
    assert _wrapdiff(p1, p0).max() < tol_x
    assert opm.rel_rms(v1, v0) < tol_v


def test_grid_mean_acceleration_is_the_particles_mean(dev):
    from augur_amd import ops
    from augur_amd.large import CompactPM
    n, G = 40000, 64
    pos, vel = _scene(n, 25)
    d = CompactPM(G, device=dev, pos_dtype="f32", vel_dtype="f32")
    st = d.load(torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev))
    d._field(st, 0.0)
    a = ops.cic_gather(d.force.view(1, 3, G, G, G), torch.from_numpy(pos).to(dev)[None], 1.0 / G, 0.0, True)[0]
    mean = a.double().mean(0).cpu()
    abar = d.mean_acceleration().cpu()
    scale = float(a.double().pow(2).sum(1).mean().sqrt())
    # the gather is the transpose of the deposit: the two agree to FP32 round-off of the gather
    assert float((abar - mean).abs().max()) < 1e-6 * scale
    assert float(abar.abs().max()) < 1e-5 * scale     # periodic PM: no net force beyond round-off
