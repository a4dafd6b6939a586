"""Plain periodic particle-mesh gravity on one MI355X.

``PeriodicPM`` runs the whole step through the HIP library: particles held
sorted by CIC base cell (tile-major order), a tiled deterministic fixed-point
deposit, the FFT
Poisson solve with Green's function -4 pi G / k^2 and spectral gradient, and
a CIC gather fused with the KDK kick and drift.  The state is re-sorted on the
device whenever the accumulated drift could move a particle by more than one
cell since the last sort.

It replaces, for the 3D periodic problem, the reference's pm_acc
(scripts/pm_baseline_force_error.py:68-76) and its KDK integrators
(scripts/gravity_sim.py:31-44, ScatterField._step verlet at
scripts/scatter_field.py:677-688).
"""
import ctypes

import torch

from . import _lib
from .ops import _dev, _p, _stream, _ws


class PMState:
    """Particles of one scene in sorted (cell) order: pos, vel, acc (n, 3),
    ids (n,) original index, mass (n,) or None with mass_value, and the
    stepper's workspace (cell offsets of the last sort and the drift bound)."""

    def __init__(self, pos, vel, acc, ids, mass, mass_value, ws):
        self.pos, self.vel, self.acc, self.ids, self.mass, self.mass_value = pos, vel, acc, ids, mass, mass_value
        self.ws = ws

    @property
    def n(self):
        return self.pos.shape[0]


def _mass_args(mass, mass_value, n, device):
    if torch.is_tensor(mass):
        return mass.detach().to(device, torch.float32).reshape(-1).contiguous().clone(), 0.0
    if mass is not None:
        return None, float(mass)
    return None, float(1.0 / n if mass_value is None else mass_value)


class PeriodicPM:
    """Periodic cube [0, L)^3 with G^3 cells (G a power of two, 32..512)."""

    def __init__(self, G, L=1.0, Gconst=1.0, device="cuda"):
        if G & (G - 1) or not 32 <= G <= 512:
            raise ValueError("G must be a power of two in [32, 512]")
        self.G, self.L, self.Gconst = int(G), float(L), float(Gconst)
        self.h = self.L / self.G
        self.device = torch.device(device)
        self._scratch = None
        self._scratch_n = -1

    def workspace_bytes(self, n):
        return _lib.size("augur_pm_workspace_bytes", n, self.G)

    def _scratch_ws(self, n):
        if self._scratch is None or self._scratch_n < n:
            self._scratch = _ws(self.workspace_bytes(n), self.device)
            self._scratch_n = n
        return self._scratch

    # ---------------------------------------------------------------- state
    def load(self, pos, vel=None, mass=None, mass_value=None, compute_acc=True):
        """Copy a scene into sorted device state.  pos, vel (n, 3); mass (n,),
        a float, or None (1/n each)."""
        pos = pos.detach().to(self.device, torch.float32).reshape(-1, 3).contiguous().clone()
        _dev(pos)
        n = pos.shape[0]
        vel = (torch.zeros_like(pos) if vel is None else
               vel.detach().to(self.device, torch.float32).reshape(-1, 3).contiguous().clone())
        acc = torch.zeros_like(pos)
        ids = torch.arange(n, dtype=torch.int32, device=self.device)
        m, mv = _mass_args(mass, mass_value, n, self.device)
        st = PMState(pos, vel, acc, ids, m, mv, _ws(self.workspace_bytes(n), self.device))
        self.sort(st)
        if compute_acc:   # a zero-length step through the state's own workspace: acc = PM(x), x and v unchanged
            self.step_begin(st, 0.0, False)
            self.step_end(st, 0.0, True)
        return st

    def sort(self, st):
        """Stable physical reorder of the state by CIC base cell."""
        _lib.call("augur_pm_prepare", _p(st.pos), _p(st.vel), _p(st.acc), _p(st.mass), _p(st.ids), st.n, self.G,
                  self.L, _p(st.ws), st.ws.numel(), _stream(self.device))

    def kdk(self, st, dt, steps, diag=False):
        """`steps` KDK leapfrog steps in place (v += dt/2 a; x += dt v; a = PM(x); v += dt/2 a).
        With diag, returns a float64 (steps, 4) tensor of [sum m v (3), sum m |v|^2 / 2]
        after each step, summed in the fused gather (the steps are the same)."""
        if diag:
            out = torch.empty(int(steps), 4, dtype=torch.float64, device=self.device)
            _lib.call("augur_pm_kdk_diag", _p(st.pos), _p(st.vel), _p(st.acc), _p(st.mass), _p(st.ids),
                      st.mass_value, st.n, self.G, self.L, self.Gconst, float(dt), int(steps), _p(out), _p(st.ws),
                      st.ws.numel(), _stream(self.device))
            return out
        _lib.call("augur_pm_kdk", _p(st.pos), _p(st.vel), _p(st.acc), _p(st.mass), _p(st.ids), st.mass_value, st.n,
                  self.G, self.L, self.Gconst, float(dt), int(steps), _p(st.ws), st.ws.numel(),
                  _stream(self.device))
        return st

    def step_begin(self, st, dt, first):
        """First half of one KDK step: (with `first`) the opening half kick and
        drift, the gated re-sort, the deposit and the FFT solve; the density
        and force planes are then in grids(st)."""
        _lib.call("augur_pm_step_begin", _p(st.pos), _p(st.vel), _p(st.acc), _p(st.mass), _p(st.ids), st.mass_value,
                  st.n, self.G, self.L, self.Gconst, float(dt), int(bool(first)), _p(st.ws), st.ws.numel(),
                  _stream(self.device))

    def step_end(self, st, dt, last):
        """Second half: gather of grids(st)'s force planes with the kicks (and,
        unless `last`, the next drift)."""
        _lib.call("augur_pm_step_end", _p(st.pos), _p(st.vel), _p(st.acc), st.n, self.G, self.L, float(dt),
                  int(bool(last)), _p(st.ws), st.ws.numel(), _stream(self.device))

    def grids(self, st):
        """Views into the state's workspace: density (G, G, G) and force
        planes (3, G, G, G), valid between step_begin and step_end."""
        ro, fo = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.call("augur_pm_grid_offsets", st.n, self.G, ctypes.byref(ro), ctypes.byref(fo))
        G3 = self.G ** 3
        base = st.ws.view(torch.uint8)
        rho = base[ro.value: ro.value + 4 * G3].view(torch.float32).view(self.G, self.G, self.G)
        force = base[fo.value: fo.value + 12 * G3].view(torch.float32).view(3, self.G, self.G, self.G)
        return rho, force

    def unload(self, st):
        """(pos, vel, acc) in the original particle order."""
        idx = st.ids.long()
        out = []
        for a in (st.pos, st.vel, st.acc):
            o = torch.empty_like(a)
            o[idx] = a
            out.append(o)
        return tuple(out)

    # ------------------------------------------------------------ functional
    def accel(self, pos, mass=None, mass_value=None):
        """PM acceleration at pos (n, 3) in the given order (no reorder)."""
        pos = pos.detach().to(self.device, torch.float32).reshape(-1, 3).contiguous()
        _dev(pos)
        n = pos.shape[0]
        m, mv = _mass_args(mass, mass_value, n, self.device)
        acc = torch.empty_like(pos)
        ws = self._scratch_ws(n)
        _lib.call("augur_pm_accel", _p(pos), _p(acc), _p(m), mv, n, self.G, self.L, self.Gconst, _p(ws), ws.numel(),
                  _stream(self.device))
        return acc

    def density(self, pos, mass=None, mass_value=None):
        """Periodic CIC mass density (G, G, G) by the sort-by-cell tiled deposit."""
        pos = pos.detach().to(self.device, torch.float32).reshape(-1, 3).contiguous()
        _dev(pos)
        n = pos.shape[0]
        m, mv = _mass_args(mass, mass_value, n, self.device)
        rho = torch.empty(self.G, self.G, self.G, dtype=torch.float32, device=self.device)
        ws = self._scratch_ws(n)
        _lib.call("augur_pm_density", _p(pos), _p(m), mv, n, self.G, self.L, _p(rho), _p(ws), ws.numel(),
                  _stream(self.device))
        return rho

    def fields(self, pos, mass=None, mass_value=None):
        """PM grids of particles in any order: density (G, G, G) and force
        (3, G, G, G), the acceleration field the gather interpolates."""
        pos = pos.detach().to(self.device, torch.float32).reshape(-1, 3).contiguous()
        _dev(pos)
        n = pos.shape[0]
        m, mv = _mass_args(mass, mass_value, n, self.device)
        rho = torch.empty(self.G, self.G, self.G, dtype=torch.float32, device=self.device)
        force = torch.empty(3, self.G, self.G, self.G, dtype=torch.float32, device=self.device)
        ws = self._scratch_ws(n)
        _lib.call("augur_pm_fields", _p(pos), _p(m), mv, n, self.G, self.L, self.Gconst, _p(rho), _p(force), _p(ws),
                  ws.numel(), _stream(self.device))
        return rho, force

    def rollout(self, pos, vel, mass=None, dt=1e-3, steps=1):
        """Load, run `steps` KDK steps, unload: returns (pos, vel, acc)."""
        st = self.load(pos, vel, mass)
        self.kdk(st, dt, steps)
        return self.unload(st)
