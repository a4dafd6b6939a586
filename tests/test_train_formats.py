"""CPU tests of the training loop's on-disk conventions and failure handling
(augur_amd.train), in the reference's own test style
(tests/test_staggered_double_grid.py:92-139): resume plus validation log,
SIGTERM forced from the host, the plateau guard; and of the exported-weights
format (augur_amd.formats) on the reference's shipped weights."""
import os
import signal

import numpy as np
import pytest
import torch

from augur_amd import formats, train

GOLD = os.path.join(os.path.dirname(__file__), "golden", "scatter2d")


class Toy:
    """A CPU stand-in for the GPU step: one parameter vector, Adam-free update."""

    def __init__(self):
        self.w = torch.zeros(3)
        self.steps = 0

    def step(self, it, lr):
        self.w += lr
        self.steps += 1
        return float(self.w.sum())

    def state(self):
        return {"w": self.w.clone()}

    def load(self, st):
        self.w = st["w"].clone()

    def opt(self):
        return {"steps": self.steps}

    def load_opt(self, st):
        self.steps = int(st["steps"])


def _run(toy, ck, iters, logs, **kw):
    return train.train_loop(toy.step, toy.state, toy.load, toy.opt, toy.load_opt, iters, ck, lr=0.1,
                            log=logs.append, log_every=1, val_fn=lambda: float(toy.w[0]), val_every=2, **kw)


def test_train_resume_and_val_log(tmp_path):
    ck = str(tmp_path / "m.pt")
    logs = []
    _run(Toy(), ck, 4, logs)
    t2 = Toy()
    _run(t2, ck, 8, logs)
    assert any(s.startswith("resumed from it 4") for s in logs)
    rows = train.read_val_log(ck + ".val.jsonl")
    assert [r["it"] for r in rows] == [0, 2, 4, 6]
    st = train.load_checkpoint(ck)
    assert st["it"] == 8 and st["opt"]["steps"] == 8
    # the learning rate followed each run's cosine schedule (over 4, then over 8 iterations)
    exp = sum(train.cosine_lr(0.1, i / 3) for i in range(4)) + sum(train.cosine_lr(0.1, i / 7) for i in range(4, 8))
    assert float(st["model"]["w"][0]) == pytest.approx(exp, rel=1e-6)
    assert not os.path.exists(ck + ".tmp")


def test_sigterm_checkpoint_then_exit(tmp_path):
    ck = str(tmp_path / "s.pt")
    toy = Toy()
    seen = []

    def log(s):
        seen.append(s)
        if s.startswith("it 2 "):
            os.kill(os.getpid(), signal.SIGTERM)

    before = signal.getsignal(signal.SIGTERM)
    with pytest.raises(SystemExit) as e:
        train.train_loop(toy.step, toy.state, toy.load, toy.opt, toy.load_opt, 10, ck, log=log, log_every=1)
    assert e.value.code == 143
    assert train.load_checkpoint(ck)["it"] == 3
    assert signal.getsignal(signal.SIGTERM) == before     # the caller's handler is back


def test_nonfinite_weights_are_not_checkpointed(tmp_path):
    ck = str(tmp_path / "n.pt")
    toy = Toy()

    def step(it, lr):
        toy.w += float("nan") if it == 3 else lr
        return 0.0

    train.train_loop(step, toy.state, toy.load, toy.opt, toy.load_opt, 6, ck, log=lambda s: None, log_every=1)
    assert train.load_checkpoint(ck)["it"] == 3     # the last save with finite weights


def test_plateau_guard():
    flat = [{"it": i, "val_loss": 1.0} for i in range(0, 100, 2)]
    assert train.plateau_guard(flat, 100)["status"] == "PLATEAUED"
    falling = [{"it": i, "val_loss": 2.0 - i / 100} for i in range(0, 100, 2)]
    assert train.plateau_guard(falling, 100)["status"] == "UNDERTRAINED"
    dup = falling + [{"it": 90, "val_loss": 0.0}]
    assert train.plateau_guard(dup, 100)["status"] == "UNDERTRAINED"
    assert train.plateau_guard(flat[:5], 100)["status"] == "NO-DATA"


def test_cosine_lr_endpoints():
    assert train.cosine_lr(1.0, 0.0) == pytest.approx(1.0)
    assert train.cosine_lr(1.0, 1.0) == pytest.approx(0.05)
    assert train.cosine_lr(1.0, 0.5) == pytest.approx(0.525)


def test_web_weights_read_matches_manifest_and_round_trips(tmp_path):
    sd, cfg, man = formats.read_web_weights(os.path.join(GOLD, "weights.json"), os.path.join(GOLD, "weights.bin"))
    assert cfg["grid"] == 128 and cfg["pp"] == 2.0 and man["iters"] == 8087
    assert len(sd) == 60 and tuple(sd["net.dec.0.0.weight"].shape) == (32, 64, 3, 3)
    # offsets are float32 counts: the last tensor ends where the blob ends
    raw = np.fromfile(os.path.join(GOLD, "weights.bin"), dtype="<f4")
    assert sum(t.numel() for t in sd.values()) == raw.size == 260419
    assert float(sd["net.out.bias"][0]) == float(raw[260418])
    jp, bp = str(tmp_path / "w.json"), str(tmp_path / "w.bin")
    formats.write_web_weights(sd, cfg, jp, bp, checkpoint=man["checkpoint"], iters=man["iters"])
    sd2, cfg2, man2 = formats.read_web_weights(jp, bp)
    assert cfg2 == cfg and man2["tensors"] == man["tensors"]
    assert all(torch.equal(sd[k], sd2[k]) for k in sd)


def test_web_weights_load_into_the_drop_in_model():
    # the drop-in's parameter names are the reference's: the shipped state_dict loads with no key mismatch
    model = formats.scatter_field_from_web(os.path.join(GOLD, "weights.json"), os.path.join(GOLD, "weights.bin"),
                                           device="cpu")
    assert model.grid == 128 and model.verlet and model.kernel and model.pp == 2.0
    assert isinstance(model.net, torch.nn.Module) and model.net.levels == 5
