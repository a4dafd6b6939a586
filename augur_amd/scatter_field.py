"""Drop-in ``ScatterField`` for the PM / scatter-field path, on the HIP library.

Keeps the reference's constructor and method signatures
(scripts/scatter_field.py:76-688): ``scatter``, ``gather``, ``kernel_acc``,
``neg_grad``, ``init_field``, ``force``, ``step``, ``_corners``,
``_cell_keys`` and the module-level ``rollout``, ``apply_opts``, ``energy``,
``momentum``.  ``grid``, ``extent`` and ``h`` stay plain mutable attributes
read on every call (the protocol scripts/scatter_tiling.py:36-44 relies on).

Extensions for the 3D periodic problem (keyword-only, defaults keep the
reference's 2D open-boundary behaviour):
  dim=3        positions (B, N, 3), grids (B, C, G, G, G)
  periodic     wrap CIC corners mod G instead of dropping them
  pm           grid force from the periodic Green's function -4 pi G/k^2 with
               a spectral gradient (the plain PM force) instead of the
               learned kernel
  surrogate    an ``augur_amd.surrogate.LocalNet3D`` evaluated on the
               deposited density (its 3 output channels are the force field)

Everything numeric runs through libaugur_hip.so; the small learned-kernel
MLP (``kmlp``) and momentum fix are torch tensor algebra on the same GPU.
"""
import math

import torch
import torch.nn as nn

from . import ops

CENTER = 500.0
H0 = 0.5
EPS0 = 0.5
DT0 = 0.1
PM_EPS = 0.5


class ScatterField(nn.Module):
    def __init__(self, grid=64, extent=64.0, net="local", recurrent=True, hidden_ch=0, dt=0.1, width=32, layers=8,
                 levels=5, momfix=True, in_scale=1.0, potential=False, kernel=False, pp=0.0, split=False, nonet=False,
                 verlet=False, prior=False, knn=16, eps=0.5, cu=False, epsin=False, norm="", massfeat=False,
                 deposit="cic", dual=False, dtin=False, pairctx=False, *, dim=2, periodic=False, pm=False,
                 Gconst=1.0, surrogate=None):
        super().__init__()
        assert deposit == "cic", "only the CIC deposit is on this path"
        if pp or prior or dual or massfeat or cu or epsin or norm or dtin or pairctx:
            raise NotImplementedError("pair term / dual / mass features / norms are outside the PM path")
        assert dim in (2, 3)
        self.dim, self.periodic, self.pm, self.Gconst = dim, bool(periodic), bool(pm), float(Gconst)
        self.deposit, self.dual = deposit, dual
        self.potential, self.kernel, self.pp, self.split, self.nonet = potential, kernel, pp, split, nonet
        self.verlet = verlet
        self.substeps = 1
        self.kpad = "zero"
        self.eps_cur = EPS0
        self.opt = set()
        self.t0 = None
        self.grid, self.extent, self.h = grid, extent, extent / grid
        self.recurrent, self.dt, self.momfix, self.in_scale = recurrent, dt, momfix, in_scale
        self.cf = (1 if potential else dim) + hidden_ch
        self.surrogate = surrogate
        if surrogate is not None:
            self.cf = 3
        if kernel:
            self.kmlp = nn.Sequential(nn.Linear(2, 64), nn.GELU(), nn.Linear(64, 64), nn.GELU(), nn.Linear(64, 1))
            nn.init.zeros_(self.kmlp[-1].weight)
            nn.init.zeros_(self.kmlp[-1].bias)
        if not nonet and surrogate is None and not pm:
            raise NotImplementedError("the 2D grid nets (LocalNet/UNet) are not part of this build; use nonet=True, "
                                      "pm=True or a 3D surrogate")

    # ------------------------------------------------------------ geometry
    @property
    def off(self):
        return self.extent / 2

    def _corners(self, pos):
        """scripts/scatter_field.py:179-182: (i0 long, w) of u = (pos + extent/2)/h - 0.5."""
        i0, w = ops.cic_base(pos, self.off, self.h)
        return i0.long(), w

    def _cell_keys(self, pos, mask, cell=None):
        """scripts/scatter_field.py:302-311 with cell size ``cell`` (the pair-term
        cell ppe() in the reference).  Returns (C, pf, mk, c, b, nc, key, counts)."""
        B, N = mask.shape
        D = pos.shape[-1]
        cell = self.h if cell is None else cell
        C = int(math.ceil(self.extent / cell))
        key, counts = ops.cell_keys(pos, mask, self.off, cell, C, counts=True)
        pf, mk = pos.reshape(B * N, D), mask.reshape(-1)
        b = torch.arange(B, device=pos.device).repeat_interleave(N)
        nc = B * C ** D
        lin = key - b * C ** D
        c = torch.stack([(lin // C ** d) % C for d in range(D)], -1)
        return C, pf, mk, c, b, nc, key, counts

    # ------------------------------------------------------ deposit / gather
    def scatter(self, pos, vel, mass, mask):
        """CIC deposit of [m, m v_0, .., m v_{D-1}] * in_scale / h^D
        (scripts/scatter_field.py:237-250) -> (B, 1+D, G, .., G)."""
        D = pos.shape[-1]
        return ops.cic_deposit(pos, mass, self.grid, self.h, self.off, self.periodic, vel=vel, mask=mask,
                               scale=float(self.in_scale / self.h ** D))

    def gather(self, field, pos):
        """CIC gather (scripts/scatter_field.py:252-261) -> (B, N, C)."""
        return ops.cic_gather(field, pos, self.h, self.off, self.periodic)

    def neg_grad(self, phi, pad="zero"):
        """-grad by zero-padded central differences (scripts/scatter_field.py:551-555), 2D."""
        if phi.dim() != 4:
            raise NotImplementedError("finite-difference gradient is the 2D reference path; 3D uses the spectral one")
        g = ops.neg_grad_fd(phi, self.h)
        return g if pad != "none" else g[..., 1:-1, 1:-1]

    # -------------------------------------------------------- grid solvers
    def _feat(self, r):
        return torch.stack([r, torch.log(r + self.h)], -1)

    def _kernel_fft(self, rho):
        G = self.grid
        idx = torch.arange(2 * G, device=rho.device, dtype=torch.float32)
        d = torch.minimum(idx, 2 * G - idx) * self.h
        r = torch.sqrt(d[:, None] ** 2 + d[None, :] ** 2)
        K = self.kmlp(self._feat(r))[..., 0]
        return ops.rfftn(K.float(), s=(2 * G, 2 * G))

    def kernel_acc(self, rho):
        """Learned-kernel open-boundary convolution (scripts/scatter_field.py:275-289), 2D."""
        G = self.grid
        Kf = self._kernel_fft(rho)
        rf = ops.rfftn(rho, s=(2 * G, 2 * G))
        full = ops.irfftn(ops.complex_mul(rf, Kf), s=(2 * G, 2 * G), out_shape=(G, G),
                          scale=self.h ** 2 / (4 * G * G))
        return self.neg_grad(full)

    def pm_grid_acc(self, rho):
        """Plain PM acceleration grids from a density grid.  3D periodic:
        Green's function + spectral gradient -> (B, 3, G, G, G).  2D open:
        the reference pm_acc convolution (scripts/pm_baseline_force_error.py:
        71-75) -> (B, 2, G, G)."""
        G = self.grid
        if rho.dim() == 5:
            L = self.extent
            spec = ops.rfftn(rho[:, 0], s=(G, G, G))
            acc = ops.poisson_green_grad(spec, G, L, self.Gconst, scale=1.0 / G ** 3)
            return ops.irfftn(acc, s=(G, G, G), scale=1.0)
        K = ops.open_kernel_2d(G, self.h, PM_EPS, rho.device)
        Kf = ops.rfftn(K, s=(2 * G, 2 * G))
        rf = ops.rfftn(rho, s=(2 * G, 2 * G))
        full = ops.irfftn(ops.complex_mul(rf, Kf), s=(2 * G, 2 * G), out_shape=(G, G),
                          scale=self.h ** 2 / (4 * G * G))
        return self.neg_grad(full)

    # --------------------------------------------------------------- force
    def init_field(self, B, device, dtype=torch.float32):
        """scripts/scatter_field.py:557-559."""
        f = torch.zeros((B, self.cf) + (self.grid,) * self.dim, device=device, dtype=dtype)
        return (f, None) if self.verlet else f

    def force(self, pos, vel, mass, mask, field):
        """(a, field) as in scripts/scatter_field.py:561-618 (grid part + momfix)."""
        x = self.scatter(pos, vel, mass, mask)
        acc = None
        if self.pm:
            acc = self.pm_grid_acc(x[:, :1])
        if self.kernel:
            ak = self.kernel_acc(x[:, :1])
            acc = ak if acc is None else acc + ak
        if self.surrogate is not None:
            field = self.surrogate(x[:, :1])
            acc = field if acc is None else acc + field
        elif self.nonet:
            field = x.new_zeros((x.shape[0], self.cf) + (self.grid,) * self.dim)
        if acc is None:
            acc = torch.zeros((x.shape[0], pos.shape[-1]) + x.shape[2:], device=x.device)
        a_tok = self.gather(acc, pos)
        if self.momfix:
            mm = (mass * mask)[..., None]
            a_tok = a_tok - (mm * a_tok).sum(1, keepdim=True) / mm.sum(1, keepdim=True).clamp_min(1e-9)
        return a_tok * mask[..., None], field

    def step(self, pos, vel, mass, mask, field):
        """scripts/scatter_field.py:668-688."""
        if self.substeps == 1:
            return self._step(pos, vel, mass, mask, field, self.dt)
        dvt = 0
        for _ in range(self.substeps):
            pos, vel, field, dv = self._step(pos, vel, mass, mask, field, self.dt / self.substeps)
            dvt = dvt + dv
        return pos, vel, field, dvt

    def _wrap(self, pos):
        if not self.periodic:
            return pos
        return torch.remainder(pos + self.off, self.extent) - self.off

    def _step(self, pos, vel, mass, mask, field, dt):
        if not self.verlet:
            a_tok, field = self.force(pos, vel, mass, mask, field)
            dv = a_tok * dt
            return self._wrap(pos + dt * (vel + 0.5 * dv)), vel + dv, field, dv
        field, a_prev = field
        if a_prev is None:
            a_prev, field = self.force(pos, vel, mass, mask, field)
        pos_new = self._wrap(pos + dt * vel + 0.5 * dt ** 2 * a_prev)
        a_new, field = self.force(pos_new, vel + dt * a_prev, mass, mask, field)
        dv = 0.5 * dt * (a_prev + a_new)
        return pos_new, vel + dv, (field, a_new), dv


def apply_opts(model, spec, kind="infer"):
    """scripts/scatter_field.py:1241: the speed flags are no-ops here (the HIP
    path has one implementation); the set is kept for callers that read it."""
    if isinstance(spec, str):
        spec = [s for s in spec.split(",") if s]
    model.opt = set(spec or ())
    return model


def _step_m(model, pos, vel, mass, mask, field):
    p, v, f, dv = model.step(pos, vel, mass, mask, field)
    return p, v, mass, f, dv


@torch.no_grad()
def rollout(model, pos, vel, mass, mask, steps):
    """scripts/scatter_field.py:745-756."""
    field = model.init_field(pos.shape[0], pos.device, pos.dtype)
    ps, vs = [pos], [vel]
    for _ in range(steps):
        pos, vel, mass, field, _ = _step_m(model, pos, vel, mass, mask, field)
        ps.append(pos)
        vs.append(vel)
    return torch.stack(ps, 1), torch.stack(vs, 1)


def momentum(vel, mass, mask):
    """scripts/scatter_field.py:768-769."""
    return ((mass * mask)[..., None] * vel).sum(-2)


def energy(pos, vel, mass, mask, eps):
    """Softened total energy (scripts/scatter_field.py:759-765) via the FP64
    direct-sum potential is O(N^2); kinetic part through the FP64 reduction."""
    from .scatter_scaling_ops import pair_potential
    mm = (mass * mask)
    out = []
    for b in range(pos.shape[0]):
        ke = 0.5 * (mm[b].double() * (vel[b].double() ** 2).sum(-1)).sum()
        out.append(ke - pair_potential(pos[b], eps, mm[b]))
    return torch.stack(out)
