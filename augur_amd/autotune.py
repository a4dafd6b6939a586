"""Conv tile autotuner for the surrogate network's BF16 MFMA convolutions.

A forward convolution (``augur_conv3_bf16_fwd_tile``) runs on a tile: a
kernel family (``ZM``: every wave all 32 output channels with 32x32x16
MFMAs; ``SPLIT``: eight waves, each half the output channels; ``NARROW``:
16x16x32 MFMAs over at most 16 outputs) and ``zc``, the number of z planes a
workgroup marches through.  Which tile is fastest depends on the layer (input
and output channels, epilogue), the grid and the batch: ``zc`` trades the
two-plane halo each workgroup loads and its weight fetch against the number of
workgroups available to fill 256 CUs.  The weight gradient is tuned over its
number of persistent workgroups.

``ConvTuner`` times every candidate that the library reports can run a layer
(``augur_conv3_tile_variants``) with HIP events on the launch stream, keeps
the fastest per (layer signature, batch, grid), and can save its table as
JSON.  A family gives the same bits for every ``zc``; families differ only in
FP32 summation order.  The reference has no counterpart (cuDNN chooses its
own algorithms inside ``nn.Conv2d``, scripts/scatter_field.py:15-26).
"""
import json
import os

import torch

from . import _lib

TILE_DEFAULT, TILE_ZM, TILE_SPLIT, TILE_NARROW, TILE_IN4 = 0, 1, 2, 3, 4
FAMILY = {TILE_ZM: "zm", TILE_SPLIT: "split", TILE_NARROW: "narrow", TILE_IN4: "in4"}
ZC_CANDIDATES = (8, 16, 32, 64, 128)
DEFAULT_ZC = 16


def variants(cin, mode, cout):
    """Kernel families that can run a layer (library's own rule)."""
    mask = _lib.get().augur_conv3_tile_variants(int(cin), int(mode), int(cout))
    return [v for v in FAMILY if (mask >> v) & 1]


def candidates(cin, mode, cout, G):
    """Every (family, zc) tile the library accepts for this layer at grid G."""
    return [(v, zc) for v in variants(cin, mode, cout) for zc in ZC_CANDIDATES if zc <= G]


def fwd_key(cin, mode, cout, B, G, addend, want_y=True):
    """Table key of a forward call; the data gradient that writes dz only
    (want_y=False, no FP32 dh store) is tuned apart from the full call."""
    return f"fwd:cin{cin}:mode{mode}:cout{cout}:add{int(bool(addend))}:B{B}:G{G}" + ("" if want_y else ":noy")


def wgrad_key(cin, B, G):
    return f"wgrad:cin{cin}:B{B}:G{G}"


def _time(fn, reps, stream):
    """Median of `reps` HIP-event timings (ms) of fn() on `stream`, after one warm call."""
    fn()
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    ms.sort()
    return ms[len(ms) // 2]


class ConvTuner:
    """Per-shape table of the fastest conv tiles, filled by timing on the GPU.

    ``fwd_tile(...)`` and ``wgrad_nwg(...)`` return the tuned choice for a
    shape or None when it has not been tuned (callers then use the library
    default).  ``tune_fwd`` / ``tune_wgrad`` time the candidates on the
    caller's tensors and record the winner; ``timings`` keeps every candidate's
    median time for reports."""

    def __init__(self, reps=5, path=None):
        self.reps = int(reps)
        self.path = path
        self.table = {}
        self.timings = {}
        if path and os.path.exists(path):
            self.load(path)

    # ------------------------------------------------------------ lookup
    def fwd_tile(self, cin, mode, cout, B, G, addend=False, want_y=True):
        t = self.table.get(fwd_key(cin, mode, cout, B, G, addend, want_y))
        return None if t is None else (int(t[0]), int(t[1]))

    def wgrad_nwg(self, cin, B, G):
        t = self.table.get(wgrad_key(cin, B, G))
        return None if t is None else int(t)

    # ------------------------------------------------------------ tuning
    def tune_fwd(self, x, w, bias, mode, cout, z_out=None, addend=None, z_in=None, want_y=True):
        """Time every tile for this layer on these tensors; returns the best (family, zc)."""
        from .surrogate import conv3_fwd
        B, G, cin = x.shape[0], x.shape[1], x.shape[-1]
        key = fwd_key(cin, mode, cout, B, G, addend is not None, want_y)
        if key in self.table:
            return self.fwd_tile(cin, mode, cout, B, G, addend is not None, want_y)
        stream = torch.cuda.current_stream(x.device)
        res = {}
        for tile in candidates(cin, mode, cout, G):
            res[tile] = _time(lambda: conv3_fwd(x, w, bias, mode, cout, z_out=z_out, addend=addend, tile=tile,
                                                z_in=z_in, want_y=want_y), self.reps, stream)
        best = min(res, key=res.get)
        self.table[key] = list(best)
        self.timings[key] = {f"{FAMILY[v]}/zc{zc}": ms for (v, zc), ms in res.items()}
        return best

    def tune_wgrad(self, x, dz, n_cu=None):
        """Time the weight gradient over its persistent workgroup count; returns the best."""
        from .surrogate import conv3_wgrad
        B, G, cin = x.shape[0], x.shape[1], x.shape[-1]
        key = wgrad_key(cin, B, G)
        if key in self.table:
            return int(self.table[key])
        if n_cu is None:
            n_cu = torch.cuda.get_device_properties(x.device).multi_processor_count
        ntiles = B * (G // 32) * (G // 4) * (G // 2)
        cands = sorted({max(1, min(ntiles, c)) for c in (n_cu // 2, n_cu, 2 * n_cu)})
        stream = torch.cuda.current_stream(x.device)
        res = {c: _time(lambda: conv3_wgrad(x, dz, nwg=c), self.reps, stream) for c in cands}
        best = min(res, key=res.get)
        self.table[key] = int(best)
        self.timings[key] = {f"nwg{c}": ms for c, ms in res.items()}
        return best

    # ------------------------------------------------------------ storage
    def save(self, path=None):
        path = path or self.path
        with open(path, "w") as f:
            json.dump({"table": self.table, "timings_ms": self.timings}, f, indent=1, sort_keys=True)

    def load(self, path):
        with open(path) as f:
            d = json.load(f)
        self.table.update(d.get("table", {}))
        self.timings.update(d.get("timings_ms", {}))

    def report(self):
        """{key: {"best": ..., "ms": {candidate: median ms}}} of the tuned shapes."""
        out = {}
        for k, v in self.table.items():
            best = f"{FAMILY[v[0]]}/zc{v[1]}" if isinstance(v, list) else f"nwg{v}"
            out[k] = {"best": best, "ms": self.timings.get(k, {})}
        return out
