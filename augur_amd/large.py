"""Large-N PM driver with compact particle state (the MI355X counterpart of
the reference's billion-body loop, scripts/scatter_bh_1b.py:107-146).

State per particle: positions in FP32, 32-bit or 16-bit fixed point of the
periodic box, velocities in FP32, FP16 or BF16 -- e.g. 3 x 4 + 3 x 2 = 18
bytes instead of 24, 12 with 16-bit positions.  It is kept in the merged KDK
form M = v + (dt/2) a, so a step is two passes over the particles, each run
over chunks of ``chunk`` particles:

  1. drift x += dt M and deposit (fixed-point CIC units into an int64 grid:
     the density is bitwise independent of the chunking);
  2. after the FFT solve, kick M += dt (a(x) - abar), where abar is the FP64
     grid-weighted mean acceleration (sum rho a / sum rho, :119-120), which
     removes the net force that FP32 round-off leaves (momentum conservation).

The kernels decode and encode the state themselves; no FP32 copy of the
particles is made.  16-bit positions resolve L / 65536: they suit scenes
whose particles move more than that per step (the reference's fp16 state at
dt = 0.1), not slow ones.
"""
import ctypes

import torch

from . import _lib
from .ops import _dev, _p, _stream, _ws

POS_ENC = {"f32": (0, torch.float32), "fix32": (1, torch.int32), "fix16": (2, torch.int16)}
VEL_ENC = {"f32": (0, torch.float32), "f16": (1, torch.float16), "bf16": (2, torch.bfloat16)}


class CompactState:
    def __init__(self, pos, vel, n, mass_value, pos_dtype, vel_dtype):
        self.pos, self.vel, self.n, self.mass_value = pos, vel, n, mass_value
        self.pos_dtype, self.vel_dtype = pos_dtype, vel_dtype
        self.started_dt = None       # dt of the half kick that made M (None: M holds v)

    @property
    def state_bytes(self):
        return self.pos.numel() * self.pos.element_size() + self.vel.numel() * self.vel.element_size()


class CompactPM:
    """Periodic PM in [0, L)^3 on G^3 cells with compact particle state."""

    def __init__(self, G, L=1.0, Gconst=1.0, device="cuda", pos_dtype="fix32", vel_dtype="f16", chunk=1 << 26,
                 momfix=True):
        self.G, self.L, self.Gconst = int(G), float(L), float(Gconst)
        self.device = torch.device(device)
        self.penc, self.pt = POS_ENC[pos_dtype]
        self.venc, self.vt = VEL_ENC[vel_dtype]
        self.pos_dtype, self.vel_dtype = pos_dtype, vel_dtype
        self.chunk, self.momfix = int(chunk), bool(momfix)
        G3 = self.G ** 3
        self.img = torch.zeros(G3, dtype=torch.int64, device=self.device)   # kept zero between steps
        self.rho = torch.empty(G3, dtype=torch.float32, device=self.device)
        self.force = torch.empty(3 * G3, dtype=torch.float32, device=self.device)
        self.ws = _ws(_lib.size("augur_pm_workspace_bytes", 1, self.G), self.device)
        self.gm_ws = _ws(_lib.size("augur_grid_mean_workspace_bytes"), self.device)
        self.abar = torch.zeros(4, dtype=torch.float64, device=self.device)

    # --------------------------------------------------------------- state
    def load(self, pos, vel=None, mass_value=None):
        """Encodes FP32 (n, 3) positions in [0, L) and velocities."""
        pos = pos.detach().to(self.device, torch.float32).reshape(-1, 3).contiguous()
        _dev(pos)
        n = pos.shape[0]
        vel = (torch.zeros_like(pos) if vel is None else
               vel.detach().to(self.device, torch.float32).reshape(-1, 3).contiguous())
        st = CompactState(torch.empty(n, 3, dtype=self.pt, device=self.device),
                          torch.empty(n, 3, dtype=self.vt, device=self.device), n,
                          float(1.0 / n if mass_value is None else mass_value), self.pos_dtype, self.vel_dtype)
        _lib.call("augur_lp_encode", _p(st.pos), self.penc, _p(st.vel), self.venc, n, self.L, _p(pos), _p(vel),
                  _stream(self.device))
        return st

    def decode(self, st):
        """(pos, M) in FP32; M is the velocity half a kick ahead once started."""
        p = torch.empty(st.n, 3, dtype=torch.float32, device=self.device)
        v = torch.empty_like(p)
        _lib.call("augur_lp_decode", _p(st.pos), self.penc, _p(st.vel), self.venc, st.n, self.L, _p(p), _p(v),
                  _stream(self.device))
        return p, v

    def _chunks(self, n):
        for lo in range(0, n, self.chunk):
            yield lo, min(n, lo + self.chunk)

    # ---------------------------------------------------------- one pass
    def _field(self, st, dt):
        """Pass 1 (drift by dt, deposit) and the solve: force planes of the drifted state."""
        s = _stream(self.device)
        for lo, hi in self._chunks(st.n):
            _lib.call("augur_lp_drift_deposit", _p(st.pos), self.penc, _p(st.vel), self.venc, lo, hi, st.n,
                      st.mass_value, float(dt), self.G, self.L, _p(self.img), s)
        _lib.call("augur_lp_finish", _p(self.img), self.G, st.n, st.mass_value, self.L, _p(self.rho), s)
        _lib.call("augur_pm_solve_grid", _p(self.rho), _p(self.force), self.G, self.L, self.Gconst, _p(self.ws),
                  self.ws.numel(), s)
        if self.momfix:
            _lib.call("augur_grid_mean_acc_f64", _p(self.rho), _p(self.force), self.G, _p(self.abar),
                      _p(self.gm_ws), self.gm_ws.numel(), s)

    def _kick(self, st, kick):
        s = _stream(self.device)
        for lo, hi in self._chunks(st.n):
            _lib.call("augur_lp_kick", _p(st.pos), self.penc, _p(st.vel), self.venc, lo, hi, float(kick),
                      _p(self.force), _p(self.abar) if self.momfix else None, self.G, self.L, s)

    def start(self, st, dt):
        """The opening half kick: M = v + (dt/2) a(x)."""
        self._field(st, 0.0)
        self._kick(st, 0.5 * dt)
        st.started_dt = float(dt)

    def step(self, st, dt, steps=1):
        """`steps` merged KDK steps (start() first): drift, deposit, solve, kick."""
        if st.started_dt is None:
            self.start(st, dt)
        assert st.started_dt == float(dt), "the merged state holds a half kick of another dt"
        for _ in range(int(steps)):
            self._field(st, dt)
            self._kick(st, dt)

    def finish(self, st):
        """The closing half kick back: M - (dt/2) a(x) = the synchronised velocity
        (the state is then unstarted again)."""
        if st.started_dt is None:
            return
        self._field(st, 0.0)
        self._kick(st, -0.5 * st.started_dt)
        st.started_dt = None

    def density(self):
        G = self.G
        return self.rho.view(G, G, G)

    def mean_acceleration(self):
        """The last solve's grid-weighted mean acceleration (FP64, 3)."""
        return self.abar[:3] / self.abar[3]
