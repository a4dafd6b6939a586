"""Drop-in ``ScatterField`` for the PM / scatter-field path, on the HIP library.

Keeps the reference's constructor and method signatures
(scripts/scatter_field.py:76-688): ``scatter``, ``gather``, ``kernel_acc``,
``neg_grad``, ``init_field``, ``force``, ``step``, ``pp_acc``, ``_corners``,
``_cell_keys`` and the module-level ``rollout``, ``apply_opts``, ``energy``,
``momentum``; the submodules ``net`` (LocalNet / UNet), ``kmlp`` and ``ppmlp``
carry the reference's parameter names, so a reference state_dict loads as is.
``grid``, ``extent`` and ``h`` stay plain mutable attributes read on every
call (the protocol scripts/scatter_tiling.py:36-44 relies on).

The 2D reference model (open boundary) runs every step in the library's
kernels: the CIC deposit, the learned kernel grid and its FFT convolution,
the grid net's FP32 convolutions (augur_conv2d_f32), the finite-difference
gradient, the gather, the cell-list kNN pair term, the momentum fix and the
Verlet / plain updates.  Torch only allocates, concatenates channel blocks
and keeps the parameters.

Extensions for the 3D periodic problem (keyword-only, defaults keep the
reference's 2D open-boundary behaviour):
  dim=3        positions (B, N, 3), grids (B, C, G, G, G)
  periodic     wrap CIC corners mod G instead of dropping them
  pm           grid force from the periodic Green's function -4 pi G/k^2 with
               a spectral gradient (the plain PM force) instead of the
               learned kernel
  surrogate    an ``augur_amd.surrogate.LocalNet3D`` evaluated on the
               deposited density (its 3 output channels are the force field)
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


class LocalNet(nn.Module):
    """The reference's LocalNet (scripts/scatter_field.py:15-26): gelu(inp(x)),
    then h + gelu(conv(h)) per middle layer, then out(h); 3x3 zero-padded
    convolutions, run by augur_conv2d_f32 (the nn.Conv2d modules hold the
    parameters under the reference's names)."""

    def __init__(self, cin, cout, width=32, layers=8):
        super().__init__()
        self.inp = nn.Conv2d(cin, width, 3, padding=1)
        self.mid = nn.ModuleList(nn.Conv2d(width, width, 3, padding=1) for _ in range(layers - 1))
        self.out = nn.Conv2d(width, cout, 3, padding=1)

    def forward(self, x):
        h = ops.conv2d(x, self.inp.weight, self.inp.bias, act=1)
        for c in self.mid:
            h = ops.conv2d(h, c.weight, c.bias, act=2, res=h)
        return ops.conv2d(h, self.out.weight, self.out.bias)


def _block(cin, cout):
    return nn.Sequential(nn.Conv2d(cin, cout, 3, padding=1), nn.GELU(), nn.Conv2d(cout, cout, 3, padding=1), nn.GELU())


class UNet(nn.Module):
    """The reference's UNet (scripts/scatter_field.py:33-67): the 2x2 average
    pool of each encoder level and the nearest x2 upsampling plus skip
    concatenation of each decoder level are read by the next convolution's
    input loads (augur_conv2d_f32 amode 1 / 2 with a second source)."""

    def __init__(self, cin, cout, width=32, levels=4):
        super().__init__()
        self.levels = levels
        self.inp = nn.Conv2d(cin, width, 3, padding=1)
        self.enc = nn.ModuleList(_block(width, width) for _ in range(levels + 1))
        self.dec = nn.ModuleList(_block(2 * width, width) for _ in range(levels))
        self.out = nn.Conv2d(width, cout, 3, padding=1)

    @staticmethod
    def _run_block(blk, h, amode=0, skip=None):
        h = ops.conv2d(h, blk[0].weight, blk[0].bias, act=1, amode=amode, b=skip)
        return ops.conv2d(h, blk[2].weight, blk[2].bias, act=1)

    def forward(self, x):
        h = self._run_block(self.enc[0], ops.conv2d(x, self.inp.weight, self.inp.bias, act=1))
        skips = [h]
        for lv in range(1, self.levels + 1):
            h = self._run_block(self.enc[lv], h, amode=1)
            skips.append(h)
        for lv in reversed(range(self.levels)):
            h = self._run_block(self.dec[lv], h, amode=2, skip=skips[lv])
        return ops.conv2d(h, self.out.weight, self.out.bias)


def _mlp(cin):
    m = nn.Sequential(nn.Linear(cin, 64), nn.GELU(), nn.Linear(64, 64), nn.GELU(), nn.Linear(64, 1))
    nn.init.zeros_(m[-1].weight)
    nn.init.zeros_(m[-1].bias)
    return m


class ScatterField(nn.Module):
    def __init__(self, grid=64, extent=64.0, net="local", recurrent=True, hidden_ch=0, dt=0.1, width=32, layers=8,
                 levels=5, momfix=True, in_scale=1.0, potential=False, kernel=False, pp=0.0, split=False, nonet=False,
                 verlet=False, prior=False, knn=16, eps=0.5, cu=False, epsin=False, norm="", massfeat=False,
                 deposit="cic", dual=False, dtin=False, pairctx=False, *, dim=2, periodic=False, pm=False,
                 Gconst=1.0, surrogate=None):
        super().__init__()
        assert deposit in ("cic", "quad")
        if prior or dual or massfeat or cu or epsin or norm or dtin or pairctx:
            raise NotImplementedError("prior / dual / mass features / norms / pair context are outside this path")
        assert dim in (2, 3)
        assert not (pp and split) or kernel, "split tapers the learned kernel"
        self.dim, self.periodic, self.pm, self.Gconst = dim, bool(periodic), bool(pm), float(Gconst)
        self.deposit, self.dual = deposit, dual
        self.potential, self.kernel, self.pp, self.split, self.nonet = potential, kernel, pp, split, nonet
        self.verlet = verlet
        self.knn, self.eps = knn, eps
        self.substeps = 1
        self.kpad = "zero"
        self.eps_cur = EPS0
        self.opt = set()
        self.t0 = None
        self.cell_max_occ = None
        self._kf = None
        self.grid, self.extent, self.h = grid, extent, extent / grid
        self.recurrent, self.dt, self.momfix, self.in_scale = recurrent, dt, momfix, in_scale
        self.cf = (1 if potential else dim) + hidden_ch
        self.surrogate = surrogate
        if surrogate is not None:
            self.cf = 3
        if pp:
            self.ppmlp = _mlp(2)
        if kernel:
            self.kmlp = _mlp(2)
        if not nonet and surrogate is None and not pm:
            if dim != 2:
                raise NotImplementedError("the 3D grid net is augur_amd.surrogate.LocalNet3D (surrogate=)")
            cin = 3 + (self.cf if recurrent else 0) + (2 if kernel else 0)
            if net == "local":
                self.net = LocalNet(cin, self.cf, width, layers)
            else:
                self.net = UNet(cin, self.cf, width, min(levels, int(math.log2(grid)) - 2))
            nn.init.zeros_(self.net.out.weight)
            nn.init.zeros_(self.net.out.bias)

    # ------------------------------------------------------------ geometry
    @property
    def off(self):
        return self.extent / 2

    def ppe(self):
        return self.pp

    def _corners(self, pos):
        """scripts/scatter_field.py:179-182: (i0 long, w) of u = (pos + extent/2)/h - 0.5."""
        i0, w = ops.cic_base(pos, self.off, self.h)
        return i0.long(), w

    def _cell_keys(self, pos, mask, cell=None):
        """scripts/scatter_field.py:302-311 with cell size ``cell`` (default the
        pair-term cell ppe()).  Returns (C, pf, mk, c, b, nc, key, counts)."""
        B, N = mask.shape
        D = pos.shape[-1]
        cell = (self.ppe() if self.pp else self.h) if cell is None else cell
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
        """CIC (or, with deposit="quad", TSC) deposit of [m, m v_0, .., m v_{D-1}]
        * in_scale / h^D (scripts/scatter_field.py:237-250) -> (B, 1+D, G, .., G)."""
        D = pos.shape[-1]
        dep = ops.tsc_deposit if self.deposit == "quad" else ops.cic_deposit
        return dep(pos, mass, self.grid, self.h, self.off, self.periodic, vel=vel, mask=mask,
                   scale=float(self.in_scale / self.h ** D))

    def gather(self, field, pos):
        """CIC (or, with deposit="quad", TSC) gather (scripts/scatter_field.py:252-261) -> (B, N, C)."""
        if self.deposit == "quad":
            return ops.tsc_gather(field, pos, self.h, self.off, self.periodic)
        return ops.cic_gather(field, pos, self.h, self.off, self.periodic)

    def neg_grad(self, phi, pad="zero"):
        """-grad by zero-padded central differences (scripts/scatter_field.py:551-555), 2D."""
        if phi.dim() != 4:
            raise NotImplementedError("finite-difference gradient is the 2D reference path; 3D uses the spectral one")
        if pad == "none":   # unpadded differences of a (G+2)^2 array -> G^2
            return ops.sf_neg_grad(_ext_from_none(phi), phi.shape[-1] - 2, self.h, ext=True)
        return ops.sf_neg_grad(phi, phi.shape[-1], self.h)

    # -------------------------------------------------------- grid solvers
    def _kernel_fft(self, rho):
        """rfft2 of the learned kernel on the (2G)^2 grid (scripts/scatter_field.py:
        263-273), cached per (grid, h, pp, parameter versions) when no
        gradient is recorded (the reference's "kcache", :277-281)."""
        G = self.grid
        key = (rho.device, G, self.h, self.pp, self.split,
               tuple(p._version for p in self.kmlp.parameters()), tuple(p.data_ptr() for p in self.kmlp.parameters()))
        if self._kf is not None and self._kf[0] == key:
            return self._kf[1]
        K = ops.sf_kernel_grid(ops.mlp_params(self.kmlp).to(rho.device), G, self.h, self.pp if self.split else 0.0)
        Kf = ops.rfftn(K, s=(2 * G, 2 * G))
        self._kf = (key, Kf)
        return Kf

    def kernel_acc(self, rho, addend=None):
        """Learned-kernel open-boundary convolution and its -grad
        (scripts/scatter_field.py:275-289), 2D; kpad "ext" differences the
        full convolution across the grid edge (:286-288)."""
        G = self.grid
        Kf = self._kernel_fft(rho)
        rf = ops.rfftn(rho, s=(2 * G, 2 * G))
        prod = ops.complex_mul(rf, Kf)
        scale = self.h ** 2 / (4 * G * G)
        if self.kpad == "ext":
            full = ops.irfftn(prod, s=(2 * G, 2 * G), scale=scale)
            return ops.sf_neg_grad(full, G, self.h, ext=True, addend=addend)
        full = ops.irfftn(prod, s=(2 * G, 2 * G), out_shape=(G, G), scale=scale)
        return ops.sf_neg_grad(full, G, self.h, addend=addend)

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
        return ops.sf_neg_grad(full, G, self.h)

    def pp_acc(self, pos, mass, mask, knn=16, qmask=None, vel=None, acc=None):
        """Learned pair term (scripts/scatter_field.py:500-522) through the cell
        list of pp_acc_cell (:324-359): cells of side pp, the knn nearest
        within pp; the dense, cell and adaptive modes of the reference give
        this same sum.  Added into ``acc`` when given."""
        if pos.shape[-1] != 2:
            raise NotImplementedError("the pair term is 2D")
        if qmask is not None:
            mask = mask * qmask
        out = torch.zeros(pos.shape, dtype=torch.float32, device=pos.device) if acc is None else acc
        ops.pair_knn(pos, mass, mask, self.extent, self.ppe(), min(knn, 16), ops.mlp_params(self.ppmlp).to(pos.device),
                     out)
        return out

    # --------------------------------------------------------------- force
    def init_field(self, B, device, dtype=torch.float32):
        """scripts/scatter_field.py:557-559."""
        f = torch.zeros((B, self.cf) + (self.grid,) * self.dim, device=device, dtype=dtype)
        return (f, None) if self.verlet else f

    def _grid_force(self, pos, vel, mass, mask, field):
        """scripts/scatter_field.py:620-666 (plain configuration): deposit, the
        kernel acceleration, [x, ak, field] through the net, -grad of its
        potential channel plus ak, gather."""
        x = self.scatter(pos, vel, mass, mask)
        acc = ak = None
        if self.pm:
            acc = self.pm_grid_acc(x[:, :1])
        if self.kernel:
            ak = self.kernel_acc(x[:, :1])
        if self.surrogate is not None:
            field = self.surrogate(x[:, :1])
            acc = field if acc is None else acc + field
        elif self.nonet or self.pm:
            field = x.new_zeros((x.shape[0], self.cf) + (self.grid,) * self.dim)
        else:
            parts = [x] + ([ak] if ak is not None else []) + ([field.float()] if self.recurrent else [])
            field = self.net(torch.cat(parts, 1).contiguous())
            if self.potential:
                acc = ops.sf_neg_grad(field[:, :1], self.grid, self.h, addend=ak)
                ak = None
            else:
                acc = field[:, :self.dim]
        if ak is not None:
            acc = ak if acc is None else acc + ak
        if acc is None:
            acc = torch.zeros((x.shape[0], pos.shape[-1]) + x.shape[2:], device=x.device)
        return self.gather(acc.contiguous(), pos), field

    def force(self, pos, vel, mass, mask, field):
        """(a, field) as in scripts/scatter_field.py:591-618: grid force, pair
        term, momfix and the mask."""
        a_tok, field = self._grid_force(pos, vel, mass, mask, field)
        if self.pp:
            self.pp_acc(pos, mass, mask, self.knn, acc=a_tok)
        ops.sf_momfix(a_tok, mass, mask, fix=self.momfix)
        return a_tok, field

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
        pos = pos.float().contiguous()
        vel = vel.float().contiguous()
        if not self.verlet:
            a_tok, field = self.force(pos, vel, mass, mask, field)
            p, v, dv = pos.clone(), vel.clone(), torch.empty_like(pos)
            ops.sf_update(p, v, a_tok, None, dv, dt, 2)
            return self._wrap(p), v, field, dv
        field, a_prev = field
        if a_prev is None:
            a_prev, field = self.force(pos, vel, mass, mask, field)
        p, vh = pos.clone(), torch.empty_like(vel)
        ops.sf_update(p, vel, a_prev, None, vh, dt, 0)
        p = self._wrap(p)
        a_new, field = self.force(p, vh, mass, mask, field)
        v, dv = vel.clone(), torch.empty_like(vel)
        ops.sf_update(p, v, a_prev, a_new, dv, dt, 1)
        return p, v, (field, a_new), dv


def _ext_from_none(phi):
    """A (B, C, G+2, G+2) array differenced unpadded, as the (B, C, 2G, 2G)
    periodic layout sf_neg_grad(ext=True) reads (rows/columns -1 .. G)."""
    B, C, P, _ = phi.shape
    G = P - 2
    full = phi.new_zeros(B, C, 2 * G, 2 * G)
    idx = torch.tensor([2 * G - 1] + list(range(G + 1)), device=phi.device)
    full[:, :, idx[:, None], idx[None, :]] = phi
    return full


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
