"""Many scenes of uneven particle counts stepped together on one GPU.

A ragged batch concatenates the scenes' particles (``ragged_batch`` in
``augur_amd.surrogate``): scene b is rows ``starts[b]:starts[b+1]`` of
``pos``/``vel``/``acc``.  One KDK step of the whole batch is a fixed sequence
of launches with no host synchronisation:

1. the fixed-point CIC deposit of every scene in one launch
   (``augur_cic_deposit_ragged``);
2. the batched FFT Poisson solve with Green's function -4 pi G / k^2 and the
   spectral gradient: the single-scene stepper's fused passes, B grids per
   launch (``augur_pm_solve_grids``);
3. optionally the scatter-field correction net on the B grids
   (``PMSurrogate`` in correction mode, BF16 MFMA convolutions at batch B);
4. one persistent launch of the CIC gather of each scene's force planes with
   the half kicks and the drift (``augur_persistent_step``).

``RaggedStepper.capture`` records a whole k-step call in a HIP graph; the
replay then advances the captured state tensors in place.  The reference
steps a padded, masked (B, N, D) batch (``ScatterField._step`` and
``rollout``, scripts/scatter_field.py:677-688, 745); here scenes keep their
own sizes and each scene's particles see only its own field.
"""
import torch

from . import _lib, ops
from .ops import _dev, _p, _stream
from .surrogate import pack_cl

PHASE_GATHER, PHASE_OPEN, PHASE_CLOSE, PHASE_CLOSE_OPEN = 0, 1, 2, 3


def persistent_step(pos, vel, acc, starts, force, G, L, dt, phase, plan=None):
    """One ragged launch (phases as augur_persistent_step) on (total, 3) FP32
    tensors updated in place; force (B, 3, G, G, G) FP32 or None for OPEN.
    With a planned deposit's ``plan`` (ops.RaggedPlan) the launch also
    records its drifts' largest |v| for that deposit's re-sort decision."""
    dev = _dev(pos)
    B = starts.shape[0] - 1
    for t in (pos, vel, acc, force):
        if t is not None:
            assert t.dtype == torch.float32 and t.is_contiguous()
    assert starts.dtype == torch.int64 and starts.is_contiguous()
    if force is not None:
        assert tuple(force.shape) == (B, 3, G, G, G), force.shape
    if plan is None:
        _lib.call("augur_persistent_step", _p(pos), _p(vel), _p(acc), _p(starts), B, pos.shape[0], _p(force), int(G),
                  float(L), float(dt), int(phase), _stream(dev))
    else:
        _lib.call("augur_persistent_step_tracked", _p(pos), _p(vel), _p(acc), _p(starts), B, pos.shape[0], _p(force),
                  int(G), float(L), float(dt), int(phase), plan.drift_word, _stream(dev))


class RaggedStepper:
    """Periodic PM (optionally with the surrogate's correction net) for a
    ragged batch of scenes in the box [0, L)^3 on a G^3 grid.

    surrogate: a calibrated ``PMSurrogate`` in correction mode whose net adds
    its correction to every scene's PM force field, or None for plain PM."""

    def __init__(self, G, L=1.0, Gconst=1.0, device="cuda", surrogate=None):
        if G & (G - 1) or not 2 <= G <= 1024:
            raise ValueError("G must be a power of two in [2, 1024]")
        self.G, self.L, self.Gconst = int(G), float(L), float(Gconst)
        self.h = self.L / self.G
        self.device = torch.device(device)
        self.sur = surrogate
        self._plan = None  # the planned deposit's workspace of the last batch size stepped
        if surrogate is not None:
            assert surrogate.mode == "correction" and surrogate.rho_scale is not None, \
                "the surrogate must be a calibrated correction net"
            assert surrogate.G == self.G and surrogate.L == self.L

    # ------------------------------------------------------------ fields
    def grids(self, pos, starts, mass, planned=None):
        """(rho (B, G, G, G), PM force (B, 3, G, G, G)) of every scene.
        planned: (plan, drift_dt, force_sort) to deposit through the planned
        ragged deposit (the stepping loop), else a deposit sorted per call."""
        G = self.G
        if planned is None:
            rho = ops.cic_deposit_ragged(pos, starts, G, self.h, 0.0, True, mass=mass)[:, 0]
        else:
            plan, drift_dt, force_sort = planned
            rho = ops.cic_deposit_ragged_planned(pos, starts, G, self.h, plan, drift_dt, force_sort, mass=mass)[:, 0]
        return rho, ops.pm_solve_grids(rho, self.L, self.Gconst)

    def plan(self, total, starts):
        """The planned deposit's workspace for a batch of these sizes (kept
        and reused while the sizes stay the same)."""
        key = (starts.shape[0] - 1, self.G, int(total))
        if self._plan is None or self._plan.key != key:
            self._plan = ops.RaggedPlan(*key, self.device)
        return self._plan

    def force_field(self, pos, starts, mass, scene_density=None, planned=None):
        """(B, 3, G, G, G) force field of every scene: plain PM, plus the
        net's correction when a surrogate is attached (scene_density (B,):
        each scene's mean density, its mass / L^3)."""
        rho, force = self.grids(pos, starts, mass, planned)
        if self.sur is None:
            return force
        sur = self.sur
        if scene_density is None:
            scene_density = self.scene_density(starts, mass, pos.shape[0])
        contrast = (rho / scene_density.view(-1, 1, 1, 1) - 1.0) * sur.rho_scale
        x = torch.cat([contrast[:, None], force / sur.force_scale], 1).contiguous()
        # the output layer adds force_scale * net(x) into the PM field in its epilogue
        return sur.net.forward_into(pack_cl(x, sur.net.cpack), force, sur.force_scale, accumulate=True)

    def scene_density(self, starts, mass, total):
        """(B,) mass / L^3 of every scene, on the device: differences of the
        FP64 prefix sum of the masses at the scene offsets (a fixed order, so
        the same bits on every call and replay)."""
        cs = torch.cat([torch.zeros(1, dtype=torch.float64, device=mass.device), mass.double().cumsum(0)])
        return ((cs[starts[1:]] - cs[starts[:-1]]) / self.L ** 3).float()

    # ------------------------------------------------------------ ordering
    def sort(self, pos, starts, *arrays):
        """Permutation putting every scene's particles in cell order (scene
        ranges unchanged): a stable radix sort by scene x G^3 + CIC base cell
        (floor(x / h - 1/2), wrapped), so that consecutive particles share
        their eight corners (the deposit merges such runs within a wave before
        its atomics) and the gather's accesses fall in nearby cells.  Returns (perm, pos[perm], *(a[perm] for a in
        arrays)); None arrays stay None.  Order only affects speed: the
        fixed-point deposit is order-independent and each particle's step is
        its own."""
        G = self.G
        total = pos.shape[0]
        B = starts.shape[0] - 1
        idx = torch.arange(total, dtype=torch.int64, device=pos.device)
        scene = torch.searchsorted(starts[1:].contiguous(), idx, right=True)
        c = torch.floor(pos.double() * (G / self.L) - 0.5).long().remainder_(G)
        keys = ((scene * G + c[:, 2]) * G + c[:, 1]) * G + c[:, 0]
        perm = ops.sort_by_cell(keys, B * G ** 3)
        return (perm, pos[perm].contiguous()) + tuple(None if a is None else a[perm].contiguous() for a in arrays)

    # ------------------------------------------------------------ stepping
    def accel(self, pos, starts, mass, acc=None, scene_density=None):
        """acc (total, 3) at pos, each scene in its own field."""
        if acc is None:
            acc = torch.empty_like(pos)
        f = self.force_field(pos, starts, mass, scene_density)
        persistent_step(pos, None, acc, starts, f, self.G, self.L, 0.0, PHASE_GATHER)
        return acc

    def kdk(self, pos, vel, acc, starts, mass, dt, steps, scene_density=None):
        """`steps` KDK steps of the whole batch in place (acc must hold the
        force at the incoming positions; it holds the force at the outgoing
        ones on return).  No host synchronisation.  The deposit keeps its
        sort and plans across the steps of the call (sorted on the first
        step, re-sorted on the device when the drift may pass a cell)."""
        if steps < 1:
            return
        if self.sur is not None and scene_density is None:
            scene_density = self.scene_density(starts, mass, pos.shape[0])
        plan = self.plan(pos.shape[0], starts)
        persistent_step(pos, vel, acc, starts, None, self.G, self.L, dt, PHASE_OPEN, plan=plan)
        for s in range(steps):
            f = self.force_field(pos, starts, mass, scene_density, planned=(plan, dt, s == 0))
            persistent_step(pos, vel, acc, starts, f, self.G, self.L, dt,
                            PHASE_CLOSE if s == steps - 1 else PHASE_CLOSE_OPEN, plan=plan)

    def capture(self, pos, vel, acc, starts, mass, dt, steps):
        """A HIP graph of kdk(...) on these tensors; ``replay()`` advances them
        in place by `steps` steps.  Copy a new batch of the same sizes into
        them to step it."""
        return RaggedGraph(self, pos, vel, acc, starts, mass, dt, steps)


class RaggedGraph:
    """A captured k-step call; it keeps every tensor the graph reads or writes
    alive (the replay uses their addresses)."""

    def __init__(self, stepper, pos, vel, acc, starts, mass, dt, steps):
        self.state = (pos, vel, acc)
        self.inputs = (starts, mass)
        # eager warm-up on copies (allocations and kernel loads), then capture;
        # the scenes' densities are recomputed inside the graph from `mass`
        p, v, a = pos.clone(), vel.clone(), acc.clone()
        stepper.kdk(p, v, a, starts, mass, dt, steps)
        torch.cuda.synchronize(pos.device)
        del p, v, a
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            stepper.kdk(pos, vel, acc, starts, mass, dt, steps)
        torch.cuda.synchronize(pos.device)

    def replay(self):
        self.graph.replay()
