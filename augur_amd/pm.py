"""Plain periodic particle-mesh gravity on one MI355X.

``PeriodicPM`` keeps its particles sorted by CIC base cell in device memory
and runs the whole step through the HIP library:

    deposit (deterministic fixed-point CIC) -> rfftn -> Green's function
    -4 pi G / k^2 -> spectral gradient -> irfftn -> fused CIC gather + kick
    + drift.

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
    ids (n,) original index, mass (n,) or None with mass_value."""

    def __init__(self, pos, vel, acc, ids, mass, mass_value):
        self.pos, self.vel, self.acc, self.ids, self.mass, self.mass_value = pos, vel, acc, ids, mass, mass_value

    @property
    def n(self):
        return self.pos.shape[0]


class PeriodicPM:
    """Periodic cube [0, L)^3 with G^3 cells (G a power of two, 8..512)."""

    def __init__(self, G, L=1.0, Gconst=1.0, device="cuda"):
        if G & (G - 1) or not 8 <= G <= 512:
            raise ValueError("G must be a power of two in [8, 512]")
        self.G, self.L, self.Gconst = int(G), float(L), float(Gconst)
        self.h = self.L / self.G
        self.device = torch.device(device)
        self._ws = None
        self._ws_n = -1

    def workspace(self, n):
        if self._ws is None or self._ws_n < n:
            nb = _lib.size("augur_pm_workspace_bytes", n, self.G)
            self._ws = _ws(nb, self.device)
            self._ws_n = n
        return self._ws

    def _wsargs(self, n):
        ws = self.workspace(n)
        return _p(ws), ws.numel()

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
        m = None
        if torch.is_tensor(mass):
            m = mass.detach().to(self.device, torch.float32).reshape(-1).contiguous().clone()
            mass_value = 0.0
        elif mass is not None:
            mass_value = float(mass)
        elif mass_value is None:
            mass_value = 1.0 / n
        st = PMState(pos, vel, acc, ids, m, float(mass_value))
        self.sort(st)
        if compute_acc:
            self.accel_state(st)
        return st

    def sort(self, st):
        """Stable physical reorder of the state by CIC base cell."""
        wp, wn = self._wsargs(st.n)
        _lib.call("augur_pm_prepare", _p(st.pos), _p(st.vel), _p(st.acc), _p(st.mass), _p(st.ids), st.n, self.G,
                  self.L, wp, wn, _stream(self.device))

    def accel_state(self, st):
        wp, wn = self._wsargs(st.n)
        _lib.call("augur_pm_accel", _p(st.pos), _p(st.acc), _p(st.mass), st.mass_value, st.n, self.G, self.L,
                  self.Gconst, wp, wn, _stream(self.device))
        return st.acc

    def kdk(self, st, dt, steps):
        """`steps` KDK leapfrog steps in place (v += dt/2 a; x += dt v; a = PM(x); v += dt/2 a)."""
        wp, wn = self._wsargs(st.n)
        _lib.call("augur_pm_kdk", _p(st.pos), _p(st.vel), _p(st.acc), _p(st.mass), st.mass_value, st.n, self.G,
                  self.L, self.Gconst, float(dt), int(steps), wp, wn, _stream(self.device))
        return st

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
        m = None
        if torch.is_tensor(mass):
            m = mass.detach().to(self.device, torch.float32).reshape(-1).contiguous()
            mass_value = 0.0
        elif mass is not None:
            mass_value = float(mass)
        elif mass_value is None:
            mass_value = 1.0 / n
        acc = torch.empty_like(pos)
        wp, wn = self._wsargs(n)
        _lib.call("augur_pm_accel", _p(pos), _p(acc), _p(m), float(mass_value), n, self.G, self.L, self.Gconst, wp,
                  wn, _stream(self.device))
        return acc

    def rollout(self, pos, vel, mass=None, dt=1e-3, steps=1):
        """Load, run `steps` KDK steps, unload: returns (pos, vel, acc)."""
        st = self.load(pos, vel, mass)
        self.kdk(st, dt, steps)
        return self.unload(st)
