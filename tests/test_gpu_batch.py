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
        s1.net.adam_step(g1, lr=1e-3)
        l2 = gr.step(*b, lr=1e-3)
        torch.cuda.synchronize()
        assert torch.equal(l1, l2), k
        for w1, w2 in zip(s1.net.w + s1.net.b, s2.net.w + s2.net.b):
            assert torch.equal(w1, w2), k
