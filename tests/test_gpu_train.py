"""GPU tests of the surrogate's training step robustness: the gradient clip at
norm 1.0 and Adam against torch (clip_grad_norm_ + torch.optim.Adam), the
non-finite skip decided on the device (a NaN gradient injected from the
host leaves weights, moments and the step count untouched), and a training
run stopped by SIGTERM and resumed from its checkpoint ending on the same
bits as an uninterrupted run (scripts/train_scatter_field.py:164-251)."""
import os
import signal

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return torch.device("cuda:0")


def _grads(net, scale, seed, nan=False):
    g = torch.Generator(device="cpu").manual_seed(seed)
    out = []
    for w, b in zip(net.w, net.b):
        dw = scale * torch.randn(w.shape, generator=g)
        db = scale * torch.randn(w.shape[1], generator=g)
        out.append((dw.to(w.device), db.to(w.device)))
    if nan:
        out[2][0][3, 1, 0] = float("nan")
    return out


def _torch_ref(net):
    params = [p.detach().cpu().double().clone().requires_grad_(True) for p in net.w + net.b]
    return params, torch.optim.Adam(params, lr=1e-3)


def _torch_step(params, opt, grads, nw):
    for k, (dw, db) in enumerate(grads):
        params[k].grad = dw.detach().cpu().double()
        gb = torch.zeros(params[nw + k].shape, dtype=torch.float64)
        gb[: db.shape[0]] = db.detach().cpu().double()
        params[nw + k].grad = gb
    gn = torch.nn.utils.clip_grad_norm_(params, 1.0)
    if torch.isfinite(gn):
        opt.step()
    opt.zero_grad()


def test_clip_and_adam_match_torch(dev):
    from augur_amd.surrogate import LocalNet3D
    net = LocalNet3D(cin=4, cout=3, layers=5, device=dev, init_std=0.05)
    params, opt = _torch_ref(net)
    nw = len(net.w)
    for s in range(3):
        grads = _grads(net, 50.0, s)
        net.adam_step(grads, lr=1e-3)
        _torch_step(params, opt, grads, nw)
        for p, q in zip(net.w + net.b, params):
            assert float((p.double().cpu() - q.detach()).abs().max()) < 2e-6
    assert net.optimizer_steps == (3, 0)
    norm = float(net.gacc[1])
    assert norm > 1.0       # the clip was active


def test_nonfinite_gradient_skips_the_step(dev):
    from augur_amd.surrogate import LocalNet3D
    net = LocalNet3D(cin=4, cout=3, layers=5, device=dev, init_std=0.05)
    params, opt = _torch_ref(net)
    nw = len(net.w)
    g1 = _grads(net, 0.3, 1)
    net.adam_step(g1, lr=1e-3)
    _torch_step(params, opt, g1, nw)
    before = [t.clone() for t in net.w + net.b]
    moments = [t.clone() for mv in net.adam for pair in mv for t in pair]
    net.adam_step(_grads(net, 0.3, 2, nan=True), lr=1e-3)            # injected from the host
    assert all(torch.equal(a, b) for a, b in zip(before, net.w + net.b))
    assert all(torch.equal(a, b) for a, b in zip(moments, [t for mv in net.adam for pair in mv for t in pair]))
    assert net.optimizer_steps == (1, 1)
    # a NaN loss skips too
    net.adam_step(_grads(net, 0.3, 3), lr=1e-3, loss=torch.tensor([float("inf")], device=dev))
    assert net.optimizer_steps == (1, 2)
    g4 = _grads(net, 0.3, 4)
    net.adam_step(g4, lr=1e-3)
    _torch_step(params, opt, g4, nw)
    assert net.optimizer_steps == (2, 2)
    for p, q in zip(net.w + net.b, params):
        assert float((p.double().cpu() - q.detach()).abs().max()) < 2e-6


def _scenes(dev, G, n=20000, K=256):
    from augur_amd import ops, scenes
    out = []
    for seed in (31, 32):
        rng = np.random.default_rng(seed)
        p = scenes.to_box_f32(scenes.clumpy(n, rng, n_clumps=8, sigma=0.04))
        pos = torch.from_numpy(p).to(dev)
        tg = torch.from_numpy(np.sort(rng.choice(n, K, replace=False))).to(dev)
        out.append((pos, tg, ops.ewald_f64(pos, 1.0 / n, tg, L=1.0, eps=1.0 / G), 1.0 / n))
    return out


def test_sigterm_resume_equals_uninterrupted_training(dev, tmp_path):
    from augur_amd import train
    from augur_amd.surrogate import PMSurrogate
    G = 32
    data = _scenes(dev, G)
    a = PMSurrogate(G, device=dev, mode="correction", init_std=0.02)
    train.surrogate_trainer(a, data, 6, str(tmp_path / "a.pt"), lr=1e-3, log=lambda s: None, log_every=1)
    ck = str(tmp_path / "b.pt")

    def log(s):
        if s.startswith("it 2 "):
            os.kill(os.getpid(), signal.SIGTERM)

    b = PMSurrogate(G, device=dev, mode="correction", init_std=0.02)
    with pytest.raises(SystemExit) as e:
        train.surrogate_trainer(b, data, 6, ck, lr=1e-3, log=log, log_every=1)
    assert e.value.code == 143 and train.load_checkpoint(ck)["it"] == 3
    c = PMSurrogate(G, device=dev, mode="correction", init_std=0.5)      # other weights: all come from the checkpoint
    logs = []
    train.surrogate_trainer(c, data, 6, ck, lr=1e-3, log=logs.append, log_every=1)
    assert "resumed from it 3" in logs
    for x, y in zip(a.net.w + a.net.b, c.net.w + c.net.b):
        assert torch.equal(x, y)
    assert a.net.optimizer_steps == c.net.optimizer_steps == (6, 0)
