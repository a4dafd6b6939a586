"""The training loop around the surrogate's GPU step, with the reference's
on-disk conventions (scripts/train_scatter_field.py:134-136,158-255,
scripts/plateau_guard.py:13-24):

* checkpoints are ``{"model", "opt", "it"}`` dicts written atomically
  (a ``.tmp`` file then ``os.replace``) every ``log_every`` iterations while
  the weights are finite, and at the end;
* a run resumes from its checkpoint when one exists;
* the learning rate follows lr (0.05 + 0.95 (1 + cos(pi frac)) / 2);
* a validation loss every ``val_every`` iterations is appended to
  ``<ckpt>.val.jsonl`` as ``{"it", "val_loss"}``, which ``plateau_guard``
  reads;
* SIGTERM saves a checkpoint after the current iteration and exits with
  status 143.

The step itself (clip at norm 1.0, non-finite skip, Adam) is the caller's
``step_fn``: for the surrogate it is ``PMSurrogate.train_step`` or a
``BatchGraph.step``, whose skip decision stays on the device.
"""
import json
import math
import os
import signal
import time

import torch


def save_atomic(obj, path):
    """torch.save to path + '.tmp', then os.replace (scripts/train_scatter_field.py:134-136)."""
    torch.save(obj, path + ".tmp")
    os.replace(path + ".tmp", path)


def load_checkpoint(path, device="cpu"):
    """A checkpoint written by save_atomic; tensors, lists, dicts and numbers
    only (loaded with weights_only=True: nothing in the file is executed)."""
    return torch.load(path, map_location=device, weights_only=True)


def cosine_lr(base, frac):
    """scripts/train_scatter_field.py:191."""
    return base * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * frac)))


def _finite(state):
    for v in state.values():
        if torch.is_tensor(v) and v.is_floating_point() and not bool(torch.isfinite(v).all()):
            return False
        if isinstance(v, (list, tuple)) and not _finite({i: t for i, t in enumerate(v)}):
            return False
        if isinstance(v, dict) and not _finite(v):
            return False
    return True


def train_loop(step_fn, model_state, load_model_state, opt_state, load_opt_state, iters, ckpt_path, lr=3e-4,
               log=print, log_every=100, val_fn=None, val_every=0, save_on_sigterm=True, time_budget=0.0):
    """Runs ``step_fn(it, lr)`` (returns the loss) from the checkpoint's
    iteration (or 0) to ``iters``.  ``model_state()`` / ``opt_state()``
    return picklable states and ``load_*`` restore them.  Returns
    (iterations done, seconds)."""
    start = 0
    stop = []
    old = None
    if save_on_sigterm:
        old = signal.signal(signal.SIGTERM, lambda *_: stop.append(1))
    try:
        if os.path.exists(ckpt_path):
            st = load_checkpoint(ckpt_path)
            load_model_state(st["model"])
            load_opt_state(st["opt"])
            start = int(st["it"])
            log(f"resumed from it {start}")

        def save(it_next):
            ms = model_state()
            if _finite(ms if isinstance(ms, dict) else {"m": ms}):
                save_atomic({"model": ms, "opt": opt_state(), "it": it_next}, ckpt_path)
                return True
            return False

        t_start = time.time()
        it = start - 1
        while True:
            it += 1
            if time_budget > 0:
                frac = (time.time() - t_start) / time_budget
                if frac >= 1.0:
                    break
            else:
                if it >= iters:
                    break
                frac = it / max(1, iters - 1)
            cur = cosine_lr(lr, frac)
            loss = step_fn(it, cur)
            if log_every and it % log_every == 0:
                log(f"it {it} loss {float(loss):.6f} lr {cur:.3e} t {time.time() - t_start:.0f}s")
                save(it + 1)
            if val_fn is not None and val_every > 0 and it % val_every == 0:
                vl = float(val_fn())
                log(f"val it {it} loss {vl:.6f}")
                with open(f"{ckpt_path}.val.jsonl", "a") as f:
                    f.write(json.dumps({"it": it, "val_loss": vl}) + "\n")
            if stop:
                save(it + 1)
                log(f"SIGTERM: checkpoint saved at it {it + 1}, exiting")
                raise SystemExit(143)
        save(max(it, start))
        log(f"trained {it} iters in {time.time() - t_start:.0f}s")
        return it, time.time() - t_start
    finally:
        if save_on_sigterm:
            signal.signal(signal.SIGTERM, old)


FRAC, DROP = 0.10, 0.03


def plateau_guard(rows, total):
    """The plateau test of scripts/plateau_guard.py:13-24 on .val.jsonl rows:
    UNDERTRAINED when the mean validation loss over the last 10% of the
    iterations is more than 3% below the mean over the 10% before, PLATEAUED
    otherwise, NO-DATA with fewer than 3 points in either window (the last
    value logged for an iteration counts)."""
    last = {}
    for r in rows:
        last[r["it"]] = r["val_loss"]
    lo2, lo1 = total * (1 - 2 * FRAC), total * (1 - FRAC)
    a = [v for it, v in last.items() if lo2 <= it < lo1]
    b = [v for it, v in last.items() if it >= lo1]
    if len(a) < 3 or len(b) < 3:
        return {"status": "NO-DATA", "n_prev": len(a), "n_last": len(b)}
    ma, mb = sum(a) / len(a), sum(b) / len(b)
    drop = (ma - mb) / ma
    return {"status": "UNDERTRAINED" if drop > DROP else "PLATEAUED", "mean_prev": ma, "mean_last": mb,
            "drop": drop, "n_prev": len(a), "n_last": len(b)}


def read_val_log(path):
    with open(path) as f:
        return [json.loads(line) for line in f if line.strip()]


def surrogate_trainer(sur, scenes, iters, ckpt_path, lr=3e-4, val=None, **kw):
    """train_loop for a PMSurrogate on (pos, targets, a_true, mass_value)
    scenes, cycling through them; ``val`` (the same tuples) gives the mean
    loss of held-out scenes without a step.  Checkpoint: {"model": the
    surrogate's state_dict, "opt": the net's Adam state, "it"}."""
    def step(it, cur_lr):
        pos, tg, at, mv = scenes[it % len(scenes)]
        return sur.train_step(pos, tg, at, mass_value=mv, lr=cur_lr)

    def val_fn():
        from .surrogate import rel_rms
        tot = 0.0
        for pos, tg, at, mv in val:
            f, _ = sur.field(pos, mass_value=mv)
            tot += rel_rms(sur.gather(f, pos[tg].contiguous()), at) ** 2
        return tot / len(val)

    return train_loop(step, sur.state_dict, sur.load_state_dict, sur.net.opt_state_dict, sur.net.load_opt_state_dict,
                      iters, ckpt_path, lr=lr, val_fn=val_fn if val else None, **kw)
