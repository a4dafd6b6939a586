"""Plain PM force and the force-error metric with the reference signatures
(scripts/pm_baseline_force_error.py:68-96), on libaugur_hip.so."""
import torch

PM_EPS = 0.5


@torch.no_grad()
def pm_acc(model, pos, vel, mass, mask):
    """Plain PM acceleration for a ScatterField-like model
    (scripts/pm_baseline_force_error.py:68-76).

    2D (pos (B, N, 2)): the reference algorithm -- CIC deposit, zero-padded
    FFT convolution with -1/sqrt(r^2 + 0.25) on the (2G)^2 domain, h^2 scale,
    crop, zero-padded central-difference gradient, CIC gather, times mask.
    3D (pos (B, N, 3)) with ``model.periodic``: periodic Green's function and
    spectral gradient."""
    rho = model.scatter(pos, vel, mass, mask)[:, :1]
    acc = model.pm_grid_acc(rho)
    return model.gather(acc, pos) * mask[..., None]


def errs(a, at, mask):
    """|a - a_true| / rms(|a_true|) per particle and |a - a_true| / |a_true|
    (scripts/pm_baseline_force_error.py:88-92)."""
    rms = ((at.norm(dim=-1) ** 2 * mask).sum(1) / mask.sum(1)).sqrt()[:, None]
    e = ((a - at).norm(dim=-1) / rms)[mask > 0]
    rel = ((a - at).norm(dim=-1) / at.norm(dim=-1).clamp_min(1e-12))[mask > 0]
    return e, rel


def stats(x):
    """scripts/pm_baseline_force_error.py:95-96."""
    return {"median": float(x.median()), "p90": float(x.quantile(0.9)), "mean": float(x.mean())}


def rel_rms(a, at):
    """sqrt(mean |a - a_t|^2) / sqrt(mean |a_t|^2) over the given rows."""
    a, at = a.double(), at.double()
    return float((((a - at) ** 2).sum(-1).mean() / (at ** 2).sum(-1).mean()).sqrt())
