"""The scatter-field CNN surrogate on the BF16 MFMA convolution kernels.

``LocalNet3D`` is the reference LocalNet topology (scripts/scatter_field.py:
15-26) in 3D with periodic padding: h = gelu(conv_in(x)); for each middle
layer h = h + gelu(conv(h)); out = conv_out(h).  With ``layers=5`` it has six
3^3 convolutions of width 32.  GELU is the tanh form (torch's
``approximate='tanh'``), which the kernels evaluate on the hardware exp2 and
reciprocal; it differs from the erf form by at most 3e-4, below the BF16
resolution of the activations.  Weights are FP32 masters [27][32][cin]
(tap-major) with BF16 copies used by the kernels; activations are BF16
channels-last with FP32 accumulation.

Training (``train_step``) runs the forward with saved pre-activations, the
backward through the project's own kernels (data gradient = forward
convolution with flipped weights, weight gradient = MFMA reduction over
voxels, GELU derivative) and Adam on the FP32 masters.
"""
import math

import numpy as np

import torch

from . import _lib, ops
from .ops import _dev, _p, _stream, _ws

BF16 = torch.bfloat16

MODE_OUT_CF, MODE_GELU, MODE_RES_GELU, MODE_OUT_CL, MODE_OUT_CL_DGELU = 0, 1, 2, 3, 4


def pad16(c):
    return 16 if c <= 16 else 32


def to_taps(w_torch):
    """torch conv3d weight (co, ci, 3, 3, 3) -> (27, co, ci) tap-major."""
    co, ci = w_torch.shape[:2]
    return w_torch.permute(2, 3, 4, 0, 1).reshape(27, co, ci)


def from_taps(w_taps):
    """(27, co, ci) -> torch conv3d layout (co, ci, 3, 3, 3)."""
    co, ci = w_taps.shape[1:]
    return w_taps.reshape(3, 3, 3, co, ci).permute(3, 4, 0, 1, 2)


def kernel_weights(w_taps, cinp):
    """(27, cout, cin) FP32 -> BF16 [27][32][cinp] zero padded (kernel layout)."""
    out = torch.zeros(27, 32, cinp, dtype=BF16, device=w_taps.device)
    out[:, : w_taps.shape[1], : w_taps.shape[2]] = w_taps.to(BF16)
    return out


def conv3_fwd(x, w, bias, mode, cout, z_out=None, addend=None, tile=None, z_in=None, want_y=True):
    """x: BF16 channels-last (B, G, G, G, cin); w: BF16 (27, 32, cin); bias FP32 (32,).
    tile: (kernel family, z planes per workgroup) from the autotuner, or None
    for the library's default tile.  Mode 4 (the data gradient fused with the
    previous layer's GELU backward) reads z_in and returns (y, dz); with
    want_y=False it writes dz only and returns (None, dz)."""
    dev = _dev(x)
    B, G, cin = x.shape[0], x.shape[1], x.shape[-1]
    y = h = None
    if mode == MODE_OUT_CF:
        y = torch.empty(B, cout, G, G, G, dtype=torch.float32, device=dev)
    elif mode == MODE_OUT_CL:
        y = torch.empty(B, G, G, G, cout, dtype=torch.float32, device=dev)
    elif mode == MODE_OUT_CL_DGELU:
        assert z_in is not None and z_in.dtype == BF16 and z_in.shape == (B, G, G, G, 32)
        y = torch.empty(B, G, G, G, 32, dtype=torch.float32, device=dev) if want_y else None
        h = torch.empty(B, G, G, G, 32, dtype=BF16, device=dev)
        z_out = z_in.contiguous()
    else:
        h = torch.empty(B, G, G, G, 32, dtype=BF16, device=dev)
    if tile is None:
        _lib.call("augur_conv3_bf16_fwd", _p(x), cin, _p(w), _p(bias), B, G, mode, cout, _p(z_out), _p(h), _p(y),
                  _p(addend), _stream(dev))
    else:
        _lib.call("augur_conv3_bf16_fwd_tile", _p(x), cin, _p(w), _p(bias), B, G, mode, cout, _p(z_out), _p(h),
                  _p(y), _p(addend), int(tile[0]), int(tile[1]), _stream(dev))
    if mode == MODE_OUT_CL_DGELU:
        return y, h
    return h if h is not None else y


def conv3_wgrad(x, dz, nwg=None):
    """-> (dw (27, 32, cin) FP32, db (32,) FP32)."""
    dev = _dev(x, dz)
    B, G, cin = x.shape[0], x.shape[1], x.shape[-1]
    if nwg is None:   # one persistent workgroup per CU (the LDS tile allows one)
        ntiles = B * (G // 32) * (G // 4) * (G // 2)
        nwg = max(1, min(256, ntiles))
    dw = torch.empty(27, 32, cin, dtype=torch.float32, device=dev)
    db = torch.empty(32, dtype=torch.float32, device=dev)
    wsb = _lib.size("augur_conv3_wgrad_workspace_bytes", cin, B, G, nwg)
    ws = _ws(wsb, dev)
    _lib.call("augur_conv3_bf16_wgrad", _p(x), cin, _p(dz), B, G, _p(dw), _p(db), nwg, _p(ws), wsb, _stream(dev))
    return dw, db


def flip_weights(w, cin_fwd, cout_fwd, cinp):
    dev = _dev(w)
    wt = torch.empty(27, 32, cinp, dtype=BF16, device=dev)
    _lib.call("augur_conv3_flip_weights", _p(w), cin_fwd, cout_fwd, cinp, _p(wt), _stream(dev))
    return wt


def gelu_bwd(dh, z):
    dz = torch.empty(z.shape, dtype=BF16, device=z.device)
    _lib.call("augur_gelu_bwd", _p(dh), _p(z), z.numel(), _p(dz), _stream(_dev(dh, z)))
    return dz


def pack_cl(x, cp, scale=1.0):
    """FP32 (B, C, G, G, G) -> BF16 channels-last (B, G, G, G, cp)."""
    dev = _dev(x)
    x = x.contiguous().float()
    B, C, G = x.shape[0], x.shape[1], x.shape[-1]
    out = torch.empty(B, G, G, G, cp, dtype=BF16, device=dev)
    _lib.call("augur_pack_cl_bf16", _p(x), C, B, G, cp, float(scale), _p(out), _stream(dev))
    return out


def adam_update(p, g, m, v, lr, step, b1=0.9, b2=0.999, eps=1e-8, gscale=1.0, p_bf16=None):
    _lib.call("augur_adam_update", _p(p), _p(g), _p(m), _p(v), p.numel(), float(lr), float(b1), float(b2),
              float(eps), int(step), float(gscale), _p(p_bf16), _stream(_dev(p)))


class LocalNet3D:
    """3D LocalNet: conv_in (cin -> 32) + GELU, (layers - 1) residual GELU
    convs (32 -> 32), conv_out (32 -> cout, zero-initialised as in the
    reference, scripts/scatter_field.py:135-136)."""

    def __init__(self, cin=4, cout=3, width=32, layers=5, device="cuda", seed=4738, init_std=None):
        assert width == 32, "the MFMA kernels are built for width 32"
        self.cin, self.cout, self.layers = cin, cout, layers
        self.cinp = pad16(cin)
        # channels of the packed input image: 4 (the packed-tap input layer) for up to 4 inputs
        self.cpack = 4 if cin <= 4 else self.cinp
        self.device = torch.device(device)
        g = torch.Generator().manual_seed(seed)
        shapes = [(cin, 32)] + [(32, 32)] * (layers - 1) + [(32, cout)]
        self.w, self.b = [], []
        for k, (ci, co) in enumerate(shapes):
            bound = 1.0 / math.sqrt(ci * 27)   # torch's default conv init scale
            std = init_std if init_std is not None else bound
            w = (torch.rand(27, co, ci, generator=g) * 2 - 1) * std
            b = (torch.rand(co, generator=g) * 2 - 1) * std
            if k == len(shapes) - 1 and init_std is None:
                w.zero_()
                b.zero_()
            self.w.append(w.to(self.device))
            self.b.append(torch.zeros(32, device=self.device).index_copy_(0, torch.arange(co, device=self.device),
                                                                          b.to(self.device)))
        self.adam = None
        self.step_count = 0
        self.tuner = None      # augur_amd.autotune.ConvTuner with the tuned tiles (None: library defaults)
        self._tuning = False
        self.sync_bf16()

    # ----------------------------------------------------------- weights
    def sync_bf16(self):
        """BF16 kernel copies of the FP32 masters, rewritten in place once they
        exist (a captured graph keeps reading the same buffers)."""
        wk = [kernel_weights(w, pad16(w.shape[2])) for w in self.w]
        # the input layer's 4-channel copy for 4-channel input images (k_conv3_in4)
        wk.append(kernel_weights(self.w[0], 4) if self.cin <= 4 else None)
        if getattr(self, "wk", None) is not None and all(
                (a is None and b is None) or (a is not None and b is not None and a.shape == b.shape)
                for a, b in zip(self.wk + [self.wk_in4], wk)):
            for a, b in zip(self.wk + [self.wk_in4], wk):
                if a is not None:
                    a.copy_(b)
        else:
            self.wk, self.wk_in4 = wk[:-1], wk[-1]

    def parameters(self):
        return self.w + self.b

    def state_dict(self):
        return {"w": [w.cpu() for w in self.w], "b": [b.cpu() for b in self.b], "cin": self.cin,
                "cout": self.cout, "layers": self.layers, "step": self.optimizer_steps[0]}

    def load_state_dict(self, sd):
        # in place when the shapes match, so a captured graph sees the new values
        for name in ("w", "b"):
            cur, new = getattr(self, name), [t.to(self.device) for t in sd[name]]
            if len(cur) == len(new) and all(a.shape == b.shape for a, b in zip(cur, new)):
                for a, b in zip(cur, new):
                    a.copy_(b)
            else:
                setattr(self, name, new)
        self.step_count = sd.get("step", 0)
        self.sync_bf16()

    def torch_weights(self):
        """[(weight (co, ci, 3, 3, 3), bias (co,))] for a torch reference."""
        return [(from_taps(w), b[: w.shape[1]]) for w, b in zip(self.w, self.b)]

    # ------------------------------------------------------ tuned tiles
    def _conv(self, x, w, b, mode, cout, z_out=None, addend=None, z_in=None, want_y=True):
        tile = None
        if self.tuner is not None:
            if self._tuning:
                tile = self.tuner.tune_fwd(x, w, b, mode, cout, z_out=z_out, addend=addend, z_in=z_in,
                                           want_y=want_y)
            else:
                tile = self.tuner.fwd_tile(x.shape[-1], mode, cout, x.shape[0], x.shape[1], addend is not None,
                                           want_y=want_y)
        return conv3_fwd(x, w, b, mode, cout, z_out=z_out, addend=addend, tile=tile, z_in=z_in, want_y=want_y)

    def _wgrad(self, x, dz):
        nwg = None
        if self.tuner is not None:
            nwg = (self.tuner.tune_wgrad(x, dz) if self._tuning
                   else self.tuner.wgrad_nwg(x.shape[-1], x.shape[0], x.shape[1]))
        return conv3_wgrad(x, dz, nwg=nwg)

    def autotune(self, B, G, tuner=None, reps=5, backward=True, seed=0):
        """Choose every conv call's tile for batch B on a G^3 grid by timing the
        candidates (augur_amd.autotune.ConvTuner) on one forward (and backward)
        pass of random activations; later passes of that shape use the winners.
        Returns the tuner (its ``report()`` lists the candidates' times)."""
        from .autotune import ConvTuner
        self.tuner = tuner if tuner is not None else (self.tuner or ConvTuner(reps=reps))
        dev = self.w[0].device
        g = torch.Generator(device=dev).manual_seed(seed)
        x = torch.randn(B, G, G, G, self.cpack, device=dev, generator=g).to(BF16)
        self._tuning = True
        try:
            self.forward_packed(x, save=backward)
            if backward:
                dy = 1e-3 * torch.randn(B, self.cout, G, G, G, device=dev, generator=g)
                self.backward(dy)
        finally:
            self._tuning = False
            self._saved = None
        torch.cuda.synchronize(dev)
        return self.tuner

    # ----------------------------------------------------------- forward
    def __call__(self, x):
        return self.forward(x)

    def forward(self, x, save=False):
        """x: FP32 (B, cin, G, G, G) -> FP32 (B, cout, G, G, G)."""
        return self.forward_packed(pack_cl(x, self.cpack), save=save)

    def forward_packed(self, xc, save=False):
        """xc: BF16 channels-last (B, G, G, G, cpack) input -> FP32 (B, cout, G, G, G).
        With cin <= 4 the image has 4 channels and the input layer packs 4
        taps x 4 channels into each MFMA's K dimension (k_conv3_in4); its
        weight gradient runs on the same 4 channels (cin 4 of
        augur_conv3_bf16_wgrad).  A cinp-channel image (zero padded) is also
        accepted."""
        h = self._hidden(xc, save)
        y = self._conv(h, self.wk[-1], self.b[-1], MODE_OUT_CF, self.cout)
        return y

    def forward_into(self, xc, field, alpha, accumulate=True, skip=None):
        """Inference straight into a force field: field (B, cout, G, G, G) FP32
        becomes alpha * net(xc) + field (accumulate) or alpha * net(xc); the
        output layer's epilogue does the update (augur_conv3_bf16_fwd_into),
        and does nothing while the int32 word ``skip`` is non-zero."""
        assert self.cout <= 16 and field.dtype == torch.float32 and field.is_contiguous()
        h = self._hidden(xc, False)
        B, G = h.shape[0], h.shape[1]
        assert field.numel() == B * self.cout * G ** 3
        zc = 0
        if self.tuner is not None:
            t = self.tuner.fwd_tile(32, MODE_OUT_CF, self.cout, B, G)
            if t is not None and t[0] == 3:    # the NARROW family: the same kernel, other epilogue
                zc = t[1]
        dev = _dev(h, field)
        _lib.call("augur_conv3_bf16_fwd_into", _p(h), 32, _p(self.wk[-1]), _p(self.b[-1]), B, G, self.cout,
                  float(alpha), int(bool(accumulate)), _p(field), _p(skip), int(zc), _stream(dev))
        return field

    def _hidden(self, xc, save):
        if xc.shape[-1] == 4 and self.wk_in4 is None:
            raise ValueError(f"a 4-channel input image needs cin <= 4 (this net has cin = {self.cin})")
        zs, hs = [], [xc]
        need_z = save
        z = torch.empty(xc.shape[:-1] + (32,), dtype=BF16, device=xc.device) if need_z else None
        w0 = self.wk_in4 if xc.shape[-1] == 4 else self.wk[0]
        h = self._conv(xc, w0, self.b[0], MODE_GELU, 32, z_out=z)
        zs.append(z)
        hs.append(h)
        for k in range(1, self.layers):
            z = torch.empty_like(h) if need_z else None
            h = self._conv(h, self.wk[k], self.b[k], MODE_RES_GELU, 32, z_out=z)
            zs.append(z)
            hs.append(h)
        if save:
            self._saved = (zs, hs)
        return h

    def backward(self, dy):
        """dy: FP32 (B, cout, G, G, G) gradient of the loss w.r.t. the output.
        Returns the FP32 gradients [(dw (27, co, ci), db (co,))] per layer."""
        zs, hs = self._saved
        grads = [None] * (self.layers + 1)
        dyc = pack_cl(dy, 32)
        # output layer
        dw, db = self._wgrad(hs[-1], dyc)
        grads[-1] = (dw[:, : self.cout, :], db[: self.cout])
        wt = flip_weights(self.wk[-1], 32, self.cout, 32)
        # dL/dh of the last hidden layer (FP32 channels-last) and, fused, dL/dz = dh * gelu'(z)
        dh, dz = self._conv(dyc, wt, None, MODE_OUT_CL_DGELU, 32, z_in=zs[-1])
        for k in range(self.layers - 1, 0, -1):
            dw, db = self._wgrad(hs[k], dz)
            grads[k] = (dw, db)
            wt = flip_weights(self.wk[k], 32, 32, 32)
            # residual path adds dh; the fused epilogue gives layer k-1's dz
            # (the last one's dh is not needed: only its dz is written)
            dh, dz = self._conv(dz, wt, None, MODE_OUT_CL_DGELU, 32, addend=dh, z_in=zs[k - 1], want_y=k > 1)
        dw, db = self._wgrad(hs[0], dz)
        grads[0] = (dw[:, :, : self.cin], db)
        self._saved = None
        return grads

    def adam_step(self, grads, lr=3e-4, loss=None, clip=1.0):
        """Adam on the FP32 masters after the reference's gradient clip and
        non-finite skip (scripts/train_scatter_field.py:225-235: clip_grad_norm_
        at 1.0, no step when the loss or the norm is not finite), all decided on
        the device: the step counter, the skip flag and the clip factor never
        reach the host, so a captured training step keeps no host sync.
        ``loss`` (a 1-element FP32/FP64 device tensor, optional) joins the
        finiteness test; ``clip`` <= 0 disables clipping."""
        dev = self.w[0].device
        if self.adam is None:
            self.adam = [((torch.zeros_like(w), torch.zeros_like(w)), (torch.zeros_like(b), torch.zeros_like(b)))
                         for w, b in zip(self.w, self.b)]
            # [steps taken, steps skipped, last step skipped]; [sum g^2, last norm]; clip factor
            self.opt_state = torch.tensor([self.step_count, 0, 0], dtype=torch.int32, device=dev)
            self.gacc = torch.zeros(2, dtype=torch.float64, device=dev)
            self.gscale = torch.ones(1, dtype=torch.float32, device=dev)
        stream = _stream(dev)
        gl = []
        for (dw, db), b in zip(grads, self.b):
            gb = torch.zeros_like(b)
            gb[: db.shape[0]] = db
            gl.append((dw.contiguous(), gb))
            for g in gl[-1]:
                _lib.call("augur_grad_sumsq", _p(g), g.numel(), _p(self.gacc), stream)
        lz = None if loss is None else loss.reshape(-1)[:1].to(torch.float32).contiguous()
        _lib.call("augur_grad_clip_decide", _p(self.gacc), _p(lz), float(clip), _p(self.gscale), _p(self.opt_state),
                  stream)
        for k, ((dw, gb), w, b) in enumerate(zip(gl, self.w, self.b)):
            (mw, vw), (mb, vb) = self.adam[k]
            for p, g, m, v in ((w, dw, mw, vw), (b, gb, mb, vb)):
                _lib.call("augur_adam_update_gated", _p(p), _p(g), _p(m), _p(v), p.numel(), float(lr), 0.9, 0.999,
                          1e-8, _p(self.opt_state), _p(self.gscale), None, stream)
        self.sync_bf16()

    @property
    def optimizer_steps(self):
        """(steps taken, steps skipped) so far (reads the device counters)."""
        if self.adam is None:
            return self.step_count, 0
        st = self.opt_state.cpu().tolist()
        return st[0], st[1]

    def opt_state_dict(self):
        """The optimizer state: Adam moments per parameter and the counters."""
        if self.adam is None:
            return {"m": None, "v": None, "counters": [self.step_count, 0, 0]}
        return {"m": [t.cpu() for mv in self.adam for t in (mv[0][0], mv[1][0])],
                "v": [t.cpu() for mv in self.adam for t in (mv[0][1], mv[1][1])],
                "counters": self.opt_state.cpu().tolist()}

    def load_opt_state_dict(self, sd):
        dev = self.w[0].device
        if sd.get("m") is None:
            self.adam = None
            self.step_count = int(sd["counters"][0])
            return
        ms, vs = [t.to(dev) for t in sd["m"]], [t.to(dev) for t in sd["v"]]
        self.adam = [((ms[2 * k], vs[2 * k]), (ms[2 * k + 1], vs[2 * k + 1])) for k in range(len(self.w))]
        self.opt_state = torch.tensor(sd["counters"], dtype=torch.int32, device=dev)
        self.gacc = torch.zeros(2, dtype=torch.float64, device=dev)
        self.gscale = torch.ones(1, dtype=torch.float32, device=dev)
        self.step_count = int(sd["counters"][0])


# --------------------------------------------------------------- PM surrogate
class SurrogateGate:
    """The device-side record of PMSurrogate.kdk's residual checks: gate words
    [fallback, revert, steps checked, first step that fell back] and the
    residual of every step.  Reading a field synchronises with the device."""

    def __init__(self, gate, log, steps):
        self.gate, self.log, self.steps = gate, log, steps

    @property
    def fell_back(self):
        return bool(self.gate is not None and int(self.gate[0]) != 0)

    @property
    def fallback_step(self):
        """The first step that fell back, or None."""
        if self.gate is None or int(self.gate[0]) == 0:
            return None
        return int(self.gate[3])

    @property
    def residuals(self):
        return None if self.log is None else self.log[: self.steps].cpu().tolist()

    @property
    def residual(self):
        """Largest per-step residual of the call (None without probes)."""
        r = self.residuals
        return max(r) if r else None

    def __iter__(self):   # fell_back, residual = sur.kdk(...)
        return iter((self.fell_back, self.residual))


def rel_rms(a, b):
    """sqrt(sum |a - b|^2 / sum |b|^2) over the rows (the force metric)."""
    a, b = a.double(), b.double()
    return float(((a - b) ** 2).sum().sqrt() / (b ** 2).sum().sqrt().clamp_min(1e-300))


class PMSurrogate:
    """The scatter-field CNN on top of 3D periodic PM (the grid half of
    ScatterField._grid_force, scripts/scatter_field.py:620-666, with the
    LocalNet of :15-26).

    mode "correction": input channels [density contrast, PM force (3)], the
        network predicts a correction added to the PM force field; trained
        against FP64 (Ewald) forces it removes part of PM's sub-grid error.
    mode "replace": input the density contrast alone, the network predicts
        the force field itself (the inference architecture of a pure CNN
        surrogate).

    accel() gathers the field at the particles.  With ``probe`` indices it
    also gathers the plain PM force there and falls back to plain PM for all
    particles when the rel-RMS residual between the two exceeds ``threshold``.
    """

    def __init__(self, G, L=1.0, Gconst=1.0, device="cuda", mode="correction", layers=5, threshold=None,
                 seed=4738, init_std=None):
        from .pm import PeriodicPM
        assert mode in ("correction", "replace")
        self.mode = mode
        self.pm = PeriodicPM(G, L, Gconst, device)
        self.G, self.L, self.h = self.pm.G, self.pm.L, self.pm.h
        self.device = self.pm.device
        cin = 4 if mode == "correction" else 1
        self.net = LocalNet3D(cin=cin, cout=3, layers=layers, device=self.device, seed=seed, init_std=init_std)
        self.threshold = threshold   # None: check_threshold's rule
        self.train_residual = None   # |a_sur - a_pm| / |a_pm| on the training targets (running mean)
        self.rho_scale = None    # 1 / rms of the density contrast
        self.force_scale = None  # rms of the PM force field

    @property
    def check_threshold(self):
        """The residual bound of the fallback check.  A replacement must stay
        within 2e-2 of plain PM.  A correction is meant to move the force by
        about PM's own error, so it is bounded by twice the residual it showed
        on its training targets (at least 2e-2): a larger one means the scene is
        outside what it was trained on.  An explicit ``threshold`` wins."""
        if self.threshold is not None:
            return float(self.threshold)
        if self.mode == "replace":
            return 2e-2
        return 0.25 if self.train_residual is None else max(2e-2, 2.0 * self.train_residual)

    def state_dict(self):
        return {"mode": self.mode, "net": self.net.state_dict(), "rho_scale": self.rho_scale,
                "force_scale": self.force_scale, "threshold": self.threshold,
                "train_residual": self.train_residual}

    def load_state_dict(self, sd):
        self.net.load_state_dict(sd["net"])
        self.rho_scale, self.force_scale = sd["rho_scale"], sd["force_scale"]
        self.threshold = sd["threshold"]
        self.train_residual = sd.get("train_residual")

    def calibrate(self, rho, force):
        """Input/output scales from one scene's grids (kept for later scenes)."""
        contrast = rho / rho.mean() - 1.0
        self.rho_scale = float(1.0 / contrast.pow(2).mean().sqrt().clamp_min(1e-6))
        self.force_scale = float(force.pow(2).mean().sqrt().clamp_min(1e-30))

    def features(self, pos, mass=None, mass_value=None):
        """(network input (1, C, G, G, G), PM force field (3, G, G, G))."""
        rho, force = self.pm.fields(pos, mass, mass_value)
        return self.features_from_grids(rho, force), force

    def packed_features(self, rho, force, mean_density):
        """The correction net's BF16 input straight from the PM grids in one
        pass (same values as pack_cl(features_from_grids(...)) up to the
        rounding of the contrast, computed here as one FMA)."""
        G = self.G
        cp = 4 if self.net.cin <= 4 else self.net.cinp   # 4 channels: the packed-tap input layer
        out = torch.empty(1, G, G, G, cp, dtype=BF16, device=self.device)
        _lib.call("augur_pack_features_bf16", _p(rho), _p(force), 1, G, float(self.rho_scale / mean_density),
                  float(-self.rho_scale), float(1.0 / self.force_scale), cp, _p(out),
                  _stream(self.device))
        return out

    def features_from_grids(self, rho, force):
        if self.rho_scale is None:
            self.calibrate(rho, force)
        contrast = (rho / rho.mean() - 1.0) * self.rho_scale
        if self.mode == "correction":
            return torch.cat([contrast[None], force / self.force_scale], 0)[None].contiguous()
        return contrast[None, None].contiguous()

    def field(self, pos, mass=None, mass_value=None, save=False):
        """(surrogate force field (3, G, G, G), PM force field)."""
        x, force = self.features(pos, mass, mass_value)
        y = self.net.forward(x, save=save)[0] * self.force_scale
        return (force + y if self.mode == "correction" else y), force

    def gather(self, field, pos):
        return ops.cic_gather(field[None], pos.reshape(1, -1, 3), self.h, 0.0, True)[0]

    def accel(self, pos, mass=None, mass_value=None, probe=None):
        """Accelerations (n, 3); with probe indices also (fallback?, residual)."""
        pos = pos.detach().to(self.device, torch.float32).reshape(-1, 3).contiguous()
        f, force = self.field(pos, mass, mass_value)
        a = self.gather(f, pos)
        if probe is None:
            return a
        a_pm = self.gather(force, pos[probe])
        res = rel_rms(a[probe], a_pm)
        if res > self.check_threshold:
            return self.gather(force, pos), True, res
        return a, False, res

    # ---------------------------------------------------------- stepping
    def load(self, pos, vel=None, mass=None, mass_value=None):
        """Sorted device state of a scene (PeriodicPM.load; acc from plain PM)."""
        return self.pm.load(pos, vel, mass, mass_value)

    def kdk(self, st, dt, steps, probe=None, threshold=None, reuse=1):
        """`steps` KDK steps of the state with the surrogate force.  Each step
        runs the PM half (re-sort, deposit, FFT) into the state's workspace,
        the network on those grids with its output layer writing straight into
        the PM force planes (correction: += force_scale * net; replace:
        = force_scale * net), then the fused gather and kicks.

        reuse = k > 1 (correction mode): the network runs on steps 0, k, 2k, ...
        of the call only; its correction field is kept (3 G^3 FP32) and added
        to the PM force planes of the k - 1 steps that follow.  The correction
        then lags the particles by up to k - 1 steps: cheaper steps, a larger
        force error (the bench reports both).

        With `probe` slot indices every step is checked against plain PM on
        the device: the force planes are gathered at the probes before and
        after the network, and when the residual
        sqrt(sum |a - a_pm|^2 / sum |a_pm|^2) exceeds `threshold` (a float, or
        one value per step; default self.threshold) the step's planes are
        re-made from the PM spectra and the rest of the call runs plain PM.
        Nothing synchronises with the host, so the call can be captured.
        Returns a SurrogateGate (its fields synchronise when read)."""
        G = self.G
        dev = self.device
        stream = _stream(dev)
        gate = log = None
        if probe is not None:
            probe = probe.to(dev, torch.int64).contiguous()
            P = probe.numel()
            if not torch.cuda.is_current_stream_capturing():
                ops._check_targets(probe, st.n)
            thr = self.check_threshold if threshold is None else threshold
            if isinstance(thr, (int, float)):   # a fill kernel: capturable
                thr = torch.full((max(1, steps),), float(thr), dtype=torch.float32, device=dev)
            else:
                thr = torch.as_tensor(thr, dtype=torch.float32).reshape(-1).to(dev)
                assert thr.numel() >= steps, "one threshold per step"
            gate = torch.empty(4, dtype=torch.int32, device=dev)
            log = torch.zeros(max(1, steps), dtype=torch.float32, device=dev)
            ref = _ws(_lib.size("augur_sur_gate_scratch_bytes", P), dev)
            _lib.call("augur_sur_gate", 0, None, None, None, 0, G, self.L, None, _p(gate), None, None, 0, stream)
        reuse = int(reuse)
        assert reuse >= 1 and (reuse == 1 or self.mode == "correction"), "reuse needs a correction net"
        corr = torch.empty(3, G, G, G, dtype=torch.float32, device=dev) if reuse > 1 else None
        mean_density = None
        for s in range(steps):
            self.pm.step_begin(st, dt, s == 0)
            rho, force = self.pm.grids(st)
            if gate is not None:
                _lib.call("augur_sur_gate", 1, _p(force), _p(st.pos), _p(probe), P, G, self.L, _p(ref), _p(gate),
                          None, None, 0, stream)
            if self.rho_scale is None:
                self.calibrate(rho, force)
            if self.mode == "correction":
                if corr is None or s % reuse == 0:
                    if mean_density is None:   # total mass / L^3
                        mean_density = self._mass_total(st) / self.L ** 3
                    xc = self.packed_features(rho, force, mean_density)
                    if corr is None:
                        self.net.forward_into(xc, force, self.force_scale, accumulate=True, skip=gate)
                    else:
                        self.net.forward_into(xc, corr, self.force_scale, accumulate=False, skip=gate)
                if corr is not None:
                    _lib.call("augur_field_add", _p(force), _p(corr), 1.0, corr.numel(), _p(gate), stream)
            else:
                xc = pack_cl(self.features_from_grids(rho, force), self.net.cpack)
                self.net.forward_into(xc, force, self.force_scale, accumulate=False, skip=gate)
            if gate is not None:
                _lib.call("augur_sur_gate", 2, _p(force), _p(st.pos), _p(probe), P, G, self.L, _p(ref), _p(gate),
                          _p(thr), _p(log), int(log.numel()), stream)
                _lib.call("augur_pm_force_replay", st.n, G, self.L, _p(gate[1:]), _p(st.ws), st.ws.numel(), stream)
            self.pm.step_end(st, dt, s == steps - 1)
        return SurrogateGate(gate, log, steps)

    def _mass_total(self, st):
        if st.mass is None:
            return st.mass_value * st.n
        if getattr(st, "mass_total", None) is None:   # once per state (one host read)
            st.mass_total = float(st.mass.double().sum())
        return st.mass_total

    def train_step(self, pos, targets, a_true, mass=None, mass_value=None, lr=3e-4):
        """One Adam step on the force loss sum |a - a_true|^2 / sum |a_true|^2
        at the target particles.  The gradient reaches the grid through the
        adjoint of the CIC gather (a CIC deposit of dL/da) and the network
        through its own backward kernels.  Returns the loss."""
        pos = pos.detach().to(self.device, torch.float32).reshape(-1, 3).contiguous()
        rho, force = self.pm.fields(pos, mass, mass_value)
        return self.train_step_grids(rho, force, pos[targets].contiguous(), a_true, lr=lr)

    def train_step_grids(self, rho, force, pt, a_true, lr=3e-4):
        """train_step on a scene's PM grids computed once (rho (G, G, G),
        force (3, G, G, G)) and its target positions pt (T, 3): the particles
        are not needed again, which lets a large scene be dropped after its
        grids are made."""
        x = self.features_from_grids(rho, force)
        y = self.net.forward(x, save=True)[0] * self.force_scale
        f = force + y if self.mode == "correction" else y
        a = self.gather(f, pt).double()
        at = a_true.to(self.device, torch.float64)
        norm = float((at ** 2).sum())
        r = a - at
        loss_t = (r ** 2).sum() / norm
        loss = float(loss_t)
        if self.mode == "correction":
            res = rel_rms(a, self.gather(force, pt))
            self.train_residual = res if self.train_residual is None else 0.9 * self.train_residual + 0.1 * res
        g = (2.0 / norm) * r.float()
        dfield = ops.cic_deposit(pt[None], None, self.G, self.h, 0.0, True, vel=g[None], scale=1.0,
                                 mass_value=1.0)[0, 1:4]
        dy = (dfield * self.force_scale)[None].contiguous()
        grads = self.net.backward(dy)
        self.net.adam_step(grads, lr=lr, loss=loss_t)
        return loss


# ------------------------------------------------------------ batched training
def ragged_batch(scenes, device):
    """Concatenate scenes of uneven sizes: (pos (total, 3) FP32, starts (B+1,)
    int64, mass (total,) with m = 1/N_b in scene b), all on the device."""
    sizes = [int(p.shape[0]) for p in scenes]
    starts = torch.tensor([0] + list(np.cumsum(sizes)), dtype=torch.int64, device=device)
    pos = torch.cat([torch.as_tensor(p, dtype=torch.float32).reshape(-1, 3).to(device) for p in scenes]).contiguous()
    mass = torch.cat([torch.full((n,), 1.0 / n, dtype=torch.float32, device=device) for n in sizes])
    return pos, starts, mass


def _batch_grids(sur, pos, starts, mass):
    """PM grids of a ragged batch: every scene's density in one deposit launch,
    then the batched fused FFT Poisson solve and spectral gradient: (rho (B,
    G, G, G), force (B, 3, G, G, G))."""
    G = sur.G
    rho = ops.cic_deposit_ragged(pos, starts, G, sur.h, 0.0, True, mass=mass)[:, 0]
    return rho, ops.pm_solve_grids(rho, sur.L, sur.pm.Gconst)


def batch_grads(sur, pos, starts, mass, tpos, a_true, inv_norm, scene_density):
    """Loss and parameter gradients of the correction net on a ragged batch
    with no host synchronisation, so it can be captured in a HIP graph.

    pos/starts/mass: the ragged batch (ragged_batch); tpos (B, T, 3) target
    positions; a_true (B, T, 3) FP64 target accelerations; inv_norm a 1-element
    FP64 tensor 1 / sum |a_true|^2; scene_density (B,) mean density of each
    scene (its mass / L^3).  The loss is sum |a - a_true|^2 * inv_norm over
    the whole batch (as train_step, summed over scenes).  Returns (loss (1,)
    FP64 tensor, grads as LocalNet3D.backward)."""
    assert sur.mode == "correction" and sur.rho_scale is not None, "calibrate the surrogate first"
    rho, force = _batch_grids(sur, pos, starts, mass)
    contrast = (rho / scene_density.view(-1, 1, 1, 1) - 1.0) * sur.rho_scale
    x = torch.cat([contrast[:, None], force / sur.force_scale], 1).contiguous()
    y = sur.net.forward(x, save=True)
    f = (force + y * sur.force_scale).contiguous()
    a = ops.cic_gather(f, tpos, sur.h, 0.0, True)
    r = a.double() - a_true
    g = (2.0 * inv_norm * r).float().contiguous()
    # loss by the library's FP64 reduction (sum m |v|^2 / 2 with m = 1)
    loss = 2.0 * inv_norm * ops.energy_momentum_f64(r.float().reshape(-1, 3), None, 1.0)[3:4]
    dfield = ops.cic_deposit(tpos, None, sur.G, sur.h, 0.0, True, vel=g, scale=1.0, mass_value=1.0)[:, 1:4]
    grads = sur.net.backward((dfield * sur.force_scale).contiguous())
    return loss, grads


class BatchGraph:
    """batch_grads captured once in a HIP graph for fixed shapes and replayed
    with each new batch copied into its static inputs, the pattern of the
    reference's captured force evaluation (scripts/scatter_field.py:567-589)
    applied to the training step (scripts/train_scatter_field.py:158-255).
    The Adam update runs after each replay (its step count lives on the host)."""

    def __init__(self, sur, pos, starts, mass, tpos, a_true, inv_norm, scene_density):
        self.sur = sur
        self.inp = [t.clone() for t in (pos, starts, mass, tpos, a_true, inv_norm, scene_density)]
        batch_grads(sur, *self.inp)          # eager warm-up: allocations and kernel loads
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.loss, self.grads = batch_grads(sur, *self.inp)
        torch.cuda.synchronize()

    def grads_for(self, pos, starts, mass, tpos, a_true, inv_norm, scene_density):
        for dst, src in zip(self.inp, (pos, starts, mass, tpos, a_true, inv_norm, scene_density)):
            dst.copy_(src)
        self.graph.replay()
        return self.loss, self.grads

    def step(self, *batch, lr=3e-4):
        loss, grads = self.grads_for(*batch)
        self.sur.net.adam_step(grads, lr=lr, loss=loss)
        return loss
