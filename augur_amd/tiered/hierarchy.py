"""Tier-0 adapters with the interfaces of model/tiered/hierarchy.py:12-49."""
import torch


class ZeroTier0:
    """model/tiered/hierarchy.py:12-25."""

    dt = 0.1
    summary_ch = 0

    def init_state(self, B, device, dtype=torch.float32):
        return None

    def force(self, pos, vel, mass, mask, state):
        return torch.zeros_like(pos), state

    def summary(self):
        return None


class ScatterFieldTier0:
    """Black-box adapter around augur_amd.scatter_field.ScatterField
    (model/tiered/hierarchy.py:28-44); touches only force/init_field."""

    summary_ch = 0

    def __init__(self, model):
        self.model, self.dt = model, model.dt

    def init_state(self, B, device, dtype=torch.float32):
        f = self.model.init_field(B, device, dtype)
        return f[0] if isinstance(f, tuple) else f

    def force(self, pos, vel, mass, mask, state):
        return self.model.force(pos, vel, mass, mask, state)

    def summary(self):
        return None


def remove_mean(a, mass, mask):
    """model/tiered/hierarchy.py:47-49."""
    mm = (mass * mask)[..., None]
    return (a - (mm * a).sum(1, keepdim=True) / mm.sum(1, keepdim=True).clamp_min(1e-9)) * mask[..., None]
