"""HIP graph capture of the PM step, the counterpart of the reference's
captured force evaluation (scripts/scatter_field.py:567-589): one KDK step
captured once and replayed matches the eager step bitwise on every replay,
also when the captured re-sort kernels open their gate from device state.
The library allocates nothing and has no host synchronisation or memset in
a step, so the whole step captures as plain kernel nodes."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return torch.device("cuda:0")


def test_pm_step_graph_replays_match_eager(dev):
    from augur_amd.pm import PeriodicPM
    rng = np.random.default_rng(41)
    n, G, dt, reps = 30000, 64, 2e-3, 6
    pos = np.mod(np.concatenate([rng.uniform(0, 1, (n // 2, 3)), 0.4 + 0.05 * rng.normal(size=(n - n // 2, 3))]),
                 1.0).astype(np.float32)
    pos = np.where(pos >= 1.0, np.float32(0.0), pos).astype(np.float32)
    vel = (3.0 * rng.normal(size=(n, 3))).astype(np.float32)     # about a cell per step at G = 64: re-sorts
    pm = PeriodicPM(G, 1.0, 1.0, dev)
    p, v = torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev)
    a = pm.load(p, v, mass_value=1.0 / n)
    b = pm.load(p, v, mass_value=1.0 / n)
    ids0 = b.ids.clone()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        pm.kdk(b, dt, 1)
    torch.cuda.synchronize()
    for r in range(reps):
        pm.kdk(a, dt, 1)
        g.replay()
        torch.cuda.synchronize()
        for x, y in ((a.pos, b.pos), (a.vel, b.vel), (a.acc, b.acc), (a.ids, b.ids)):
            assert torch.equal(x, y), r
    assert not torch.equal(b.ids, ids0)   # the replays re-sorted the state on the device


def test_pm_multi_step_graph_replays_match_eager(dev):
    # a 3-step call (fused kicks between steps) captured once; replays
    # interleaved with eager work on another state
    from augur_amd.pm import PeriodicPM
    rng = np.random.default_rng(42)
    n, G, dt = 20000, 32, 1e-3
    pos = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    vel = (0.5 * rng.normal(size=(n, 3))).astype(np.float32)
    pm = PeriodicPM(G, 1.0, 1.0, dev)
    p, v = torch.from_numpy(pos).to(dev), torch.from_numpy(vel).to(dev)
    a, b, c = (pm.load(p, v, mass_value=1.0 / n) for _ in range(3))
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        pm.kdk(b, dt, 3)
    for r in range(3):
        pm.kdk(a, dt, 3)
        pm.kdk(c, 2 * dt, 1)      # unrelated eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(a.pos, b.pos) and torch.equal(a.vel, b.vel) and torch.equal(a.acc, b.acc), r
