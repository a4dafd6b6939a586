"""FP64 direct-sum oracle helpers with the reference signatures
(scripts/scatter_scaling_ops.py), running on the GPU kernels of
libaugur_hip.so, plus the periodic (Ewald) oracle for the 3D periodic path."""
import torch

from . import ops


@torch.no_grad()
def exact_acc(pos, eps, mass=None, elems=3.0e8, targets=None):
    """Softened all-pairs acceleration, G=1 (scripts/scatter_scaling_ops.py:56-71).
    pos (N, D) float64 or float32 on the GPU; ``elems`` is accepted for
    signature compatibility (the kernel tiles sources through LDS instead of
    chunking).  ``targets`` restricts the rows (returns (T, D))."""
    del elems
    return ops.direct_sum_f64(pos, eps, mass, targets)


@torch.no_grad()
def exact_rollout(pos, vel, steps, dt, substeps, eps):
    """KDK with substeps and the FP64 direct-sum force (scripts/scatter_scaling_ops.py:74-87)."""
    h = dt / substeps
    pos, vel = pos.double(), vel.double()
    ps, vs = [pos], [vel]
    a = exact_acc(pos, eps)
    for _ in range(steps):
        for _ in range(substeps):
            vel = vel + 0.5 * h * a
            pos = pos + h * vel
            a = exact_acc(pos, eps)
            vel = vel + 0.5 * h * a
        ps.append(pos)
        vs.append(vel)
    return torch.stack(ps), torch.stack(vs)


def pair_potential(pos, eps, mass=None):
    """sum_{i<j} m_i m_j / sqrt(r^2 + eps^2) in FP64."""
    phi = ops.direct_potential_f64(pos, eps, mass)
    m = torch.ones(pos.shape[0], dtype=torch.float64, device=pos.device) if mass is None else mass.double()
    return 0.5 * (m * phi).sum()


@torch.no_grad()
def exact_energy(pos, vel, eps, elems=3.0e8):
    """Total energy for unit masses, G=1, softened (scripts/scatter_scaling_ops.py:90-103)."""
    del elems
    return float(0.5 * (vel.double() ** 2).sum() - pair_potential(pos, eps))


@torch.no_grad()
def ewald_acc(pos, mass, targets, L=1.0, eps=0.0, alpha_l=8.0, kmax=None):
    """FP64 periodic accelerations (mean density removed) at the targets."""
    return ops.ewald_f64(pos, mass, targets, L, eps, alpha_l, kmax)


def retarget(model, grid, extent):
    """scripts/scatter_scaling_ops.py:50-53."""
    model.grid, model.extent, model.h = grid, float(extent), float(extent) / grid
    return model
