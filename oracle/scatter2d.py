"""CPU FP64 restatement of the reference's 2D scatter-field step (test
infrastructure only): the model ScatterField of scripts/scatter_field.py in
the configurations its shipped weights use (web/scatter/weights.json:
UNet, learned kernel with the split taper, pair term, potential output,
recurrent field, Verlet, momfix), plus the LocalNet net, the non-Verlet
step, kpad="ext" and the TSC ("quad") deposit.

The arithmetic is torch float64 on the CPU (conv2d, rfft2/irfft2, topk),
which is the reference's own formulation; every function cites the lines it
restates.  It is pinned against the reference's FP64 rollouts
web/scatter/tests/vectors.json (copied to tests/golden/scatter2d/) by
tests/test_oracle_scatter2d.py, and by closed forms for the pieces those
rollouts do not reach (kpad="ext", TSC weights, LocalNet, non-Verlet step).

``params`` is a dict of float64 tensors keyed by the reference's state_dict
names (net.inp.weight, net.enc.0.0.weight, kmlp.0.weight, ppmlp.4.bias, ...);
``cfg`` a dict with the reference's constructor arguments (grid, extent, dt,
levels, pp, knn, in_scale, momfix, potential, kernel, split, verlet, net,
recurrent, kpad, deposit).
"""
import json
import math

import numpy as np
import torch
import torch.nn.functional as F

F64 = torch.float64


def load_web_weights(json_path, bin_path):
    """The reference's exported weights (web/scatter/weights.json manifest with
    float32 offsets into weights.bin) -> (params float64 dict, cfg dict)."""
    man = json.load(open(json_path))
    raw = np.fromfile(bin_path, dtype="<f4")
    params = {}
    for k, v in man["tensors"].items():
        n = int(np.prod(v["shape"]))
        params[k] = torch.from_numpy(raw[v["offset"]: v["offset"] + n].astype(np.float64).reshape(v["shape"]))
    cfg = dict(man["config"])
    cfg.setdefault("net", "unet")
    return params, cfg


def default_cfg(**kw):
    """Constructor defaults of ScatterField (scripts/scatter_field.py:85-88)."""
    c = {"grid": 64, "extent": 64.0, "net": "local", "recurrent": True, "hidden_ch": 0, "dt": 0.1, "width": 32,
         "layers": 8, "levels": 5, "momfix": True, "in_scale": 1.0, "potential": False, "kernel": False, "pp": 0.0,
         "split": False, "nonet": False, "verlet": False, "knn": 16, "eps": 0.5, "kpad": "zero", "deposit": "cic"}
    c.update(kw)
    return c


# ------------------------------------------------------------- nets (:15-67)
def gelu(x):
    """nn.GELU() / F.gelu default: the erf form."""
    return F.gelu(x)


def mlp2(params, prefix, feats):
    """nn.Sequential(Linear(c, 64), GELU, Linear(64, 64), GELU, Linear(64, k))
    (scripts/scatter_field.py:120-121,128-130)."""
    h = gelu(feats @ params[prefix + ".0.weight"].T + params[prefix + ".0.bias"])
    h = gelu(h @ params[prefix + ".2.weight"].T + params[prefix + ".2.bias"])
    return h @ params[prefix + ".4.weight"].T + params[prefix + ".4.bias"]


def conv3(x, params, name):
    """nn.Conv2d(cin, cout, 3, padding=1): zero padding."""
    return F.conv2d(x, params[name + ".weight"], params[name + ".bias"], padding=1)


def unet(x, params, levels):
    """UNet.forward (scripts/scatter_field.py:58-67) without the sparse/ablate
    hooks (identity at their defaults, :43-56): blocks conv-GELU-conv-GELU
    (:29-30), avg_pool2d(2) down, nearest x2 up, skip concatenated after."""
    def block(h, name):
        return gelu(conv3(gelu(conv3(h, params, name + ".0")), params, name + ".2"))
    h = block(gelu(conv3(x, params, "net.inp")), "net.enc.0")
    skips = [h]
    for lv in range(1, levels + 1):
        h = block(F.avg_pool2d(h, 2), f"net.enc.{lv}")
        skips.append(h)
    for lv in reversed(range(levels)):
        h = F.interpolate(h, scale_factor=2, mode="nearest")
        h = block(torch.cat([h, skips[lv]], 1), f"net.dec.{lv}")
    return conv3(h, params, "net.out")


def localnet(x, params, layers):
    """LocalNet.forward (scripts/scatter_field.py:22-26)."""
    h = gelu(conv3(x, params, "net.inp"))
    for k in range(layers - 1):
        h = h + gelu(conv3(h, params, f"net.mid.{k}"))
    return conv3(h, params, "net.out")


def unet_levels(cfg):
    """min(levels, log2(grid) - 2) (scripts/scatter_field.py:133)."""
    return min(cfg["levels"], int(math.log2(cfg["grid"])) - 2)


# ------------------------------------------------------ deposit / gather
def corners(pos, cfg):
    """(flat index, weight) pairs of ScatterField._corner_iter
    (scripts/scatter_field.py:190-211): CIC's 4 corners, or the TSC ("quad")
    3x3 stencil with weights 0.5 (0.5 - d)^2, 0.75 - d^2, 0.5 (0.5 + d)^2 about
    i0 = round(u) (:184-188,193-201); off-grid corners get weight 0."""
    G, h, ext = cfg["grid"], cfg["extent"] / cfg["grid"], cfg["extent"]
    B, N = pos.shape[:2]
    bidx = torch.arange(B)[:, None].expand(B, N)
    u = (pos + ext / 2) / h - 0.5
    out = []
    if cfg.get("deposit", "cic") == "quad":
        i0 = torch.round(u)
        d = u - i0
        wq = (0.5 * (0.5 - d) ** 2, 0.75 - d ** 2, 0.5 * (0.5 + d) ** 2)
        i0 = i0.long()
        for ox in (-1, 0, 1):
            for oy in (-1, 0, 1):
                ix, iy = i0[..., 0] + ox, i0[..., 1] + oy
                wt = wq[ox + 1][..., 0] * wq[oy + 1][..., 1]
                ok = ((ix >= 0) & (ix < G) & (iy >= 0) & (iy < G)).to(pos.dtype)
                out.append((bidx * G * G + iy.clamp(0, G - 1) * G + ix.clamp(0, G - 1), wt * ok))
        return out
    i0f = torch.floor(u)
    w = u - i0f
    i0 = i0f.long()
    for ox in (0, 1):
        for oy in (0, 1):
            ix, iy = i0[..., 0] + ox, i0[..., 1] + oy
            wt = (w[..., 0] if ox else 1 - w[..., 0]) * (w[..., 1] if oy else 1 - w[..., 1])
            ok = ((ix >= 0) & (ix < G) & (iy >= 0) & (iy < G)).to(pos.dtype)
            out.append((bidx * G * G + iy.clamp(0, G - 1) * G + ix.clamp(0, G - 1), wt * ok))
    return out


def scatter(pos, vel, mass, mask, cfg):
    """ScatterField.scatter (scripts/scatter_field.py:237-250): channels
    [m, m vx, m vy] * in_scale / h^2 -> (B, 3, G, G)."""
    B = pos.shape[0]
    G, h = cfg["grid"], cfg["extent"] / cfg["grid"]
    mm = mass * mask
    val = torch.stack([mm, mm * vel[..., 0], mm * vel[..., 1]], -1) * (cfg["in_scale"] / h ** 2)
    out = torch.zeros(B * G * G, 3, dtype=pos.dtype)
    for flat, w in corners(pos, cfg):
        out = out.index_add(0, flat.reshape(-1), (val * w[..., None]).reshape(-1, 3))
    return out.view(B, G, G, 3).permute(0, 3, 1, 2)


def gather(field, pos, cfg):
    """ScatterField.gather (scripts/scatter_field.py:252-261)."""
    B, C, G, _ = field.shape
    fl = field.permute(0, 2, 3, 1).reshape(B * G * G, C)
    out = 0
    for flat, w in corners(pos, cfg):
        out = out + fl[flat] * w[..., None]
    return out


# --------------------------------------------------------- grid solvers
def neg_grad(phi, h, pad="zero"):
    """ScatterField.neg_grad (scripts/scatter_field.py:551-555)."""
    p = phi if pad == "none" else F.pad(phi, (1, 1, 1, 1))
    gx = (p[:, :, 1:-1, 2:] - p[:, :, 1:-1, :-2]) / (2 * h)
    gy = (p[:, :, 2:, 1:-1] - p[:, :, :-2, 1:-1]) / (2 * h)
    return -torch.cat([gx, gy], 1)


def kernel_grid(params, cfg):
    """The learned kernel K(r) on the (2G)^2 min-image distance grid with the
    split taper (ScatterField._kernel_fft, scripts/scatter_field.py:263-273;
    features [r, log(r + h)], :160-161)."""
    G, h = cfg["grid"], cfg["extent"] / cfg["grid"]
    idx = torch.arange(2 * G, dtype=F64)
    d = torch.minimum(idx, 2 * G - idx) * h
    r = torch.sqrt(d[:, None] ** 2 + d[None, :] ** 2)
    K = mlp2(params, "kmlp", torch.stack([r, torch.log(r + h)], -1))[..., 0]
    if cfg["split"]:
        K = K * (1 - (1 - (r / cfg["pp"]).clamp(max=1.0) ** 2) ** 2)
    return K


def kernel_acc(rho, params, cfg):
    """ScatterField.kernel_acc (scripts/scatter_field.py:275-289): zero-padded
    FFT convolution on (2G)^2 times h^2, then the FD gradient of the [:G, :G]
    crop, or with kpad="ext" of rows/columns [2G-1, 0 .. G] unpadded."""
    G, h = cfg["grid"], cfg["extent"] / cfg["grid"]
    Kf = torch.fft.rfft2(kernel_grid(params, cfg))
    rf = torch.fft.rfft2(rho, s=(2 * G, 2 * G))
    full = torch.fft.irfft2(rf * Kf, s=(2 * G, 2 * G)) * h ** 2
    if cfg.get("kpad", "zero") == "ext":
        idx = torch.tensor([2 * G - 1] + list(range(G + 1)))
        return neg_grad(full[..., idx, :][..., idx], h, "none")
    return neg_grad(full[..., :G, :G], h)


def pair_acc(pos, mass, mask, params, cfg):
    """Dense kNN pair term (ScatterField.pp_acc dense path, scripts/scatter_field.py:
    504-522, with _pair_w :291-300): the k = min(knn, N-1) nearest by
    r = sqrt(|d|^2 + 1e-8), weight ppmlp([r, log(r + 0.05)]) * (1 - min(r/pp, 1)^2)^2
    * m_j, summed along d / r."""
    B, N = mask.shape
    k = min(cfg["knn"], N - 1)
    d_all = pos[:, None, :, :] - pos[:, :, None, :]
    r_all = torch.sqrt((d_all ** 2).sum(-1) + 1e-8)
    bad = (1 - mask)[:, None, :] + torch.eye(N, dtype=pos.dtype)[None]
    r_sel, idx = torch.topk(r_all + 1e6 * (bad > 0).to(pos.dtype), k, dim=-1, largest=False)
    d = torch.gather(d_all, 2, idx[..., None].expand(B, N, k, 2))
    mj = torch.gather((mass * mask)[:, None, :].expand(B, N, N), 2, idx)
    ok = (r_sel < 1e5).to(pos.dtype) * mask[:, :, None]
    r = r_sel.clamp(max=1e5)
    g = mlp2(params, "ppmlp", torch.stack([r, torch.log(r + 0.05)], -1))[..., 0]
    win = (1 - (r / cfg["pp"]).clamp(max=1.0) ** 2) ** 2
    w = g * win * mj * ok
    return (w[..., None] * d / r[..., None]).sum(2)


# ---------------------------------------------------------------- force
def force(pos, vel, mass, mask, field, params, cfg):
    """ScatterField._force / _grid_force for the plain configuration
    (scripts/scatter_field.py:591-618, 620-666): scatter, kernel
    acceleration, [x, ak, field] through the net, -grad of the potential
    channel (or the first two channels), gather, pair term, momfix, mask."""
    h = cfg["extent"] / cfg["grid"]
    x = scatter(pos, vel, mass, mask, cfg)
    ak = kernel_acc(x[:, :1], params, cfg) if cfg["kernel"] else None
    if ak is not None:
        x = torch.cat([x, ak], 1)
    if cfg["recurrent"]:
        x = torch.cat([x, field], 1)
    if cfg.get("nonet", False):
        cf = field.shape[1]
        field = x.new_zeros(x.shape[0], cf, cfg["grid"], cfg["grid"])
    elif cfg["net"] == "unet":
        field = unet(x, params, unet_levels(cfg))
    else:
        field = localnet(x, params, cfg["layers"])
    acc = neg_grad(field[:, :1], h) if cfg["potential"] else field[:, :2]
    if ak is not None:
        acc = acc + ak
    a_tok = gather(acc, pos, cfg)
    if cfg["pp"]:
        a_tok = a_tok + pair_acc(pos, mass, mask, params, cfg)
    if cfg["momfix"]:
        mm = (mass * mask)[..., None]
        a_tok = a_tok - (mm * a_tok).sum(1, keepdim=True) / mm.sum(1, keepdim=True).clamp_min(1e-9)
    return a_tok * mask[..., None], field


def init_field(B, cfg):
    """ScatterField.init_field (scripts/scatter_field.py:557-559)."""
    cf = (1 if cfg["potential"] else 2) + cfg.get("hidden_ch", 0)
    f = torch.zeros(B, cf, cfg["grid"], cfg["grid"], dtype=F64)
    return (f, None) if cfg["verlet"] else f


def step(pos, vel, mass, mask, field, params, cfg):
    """ScatterField._step (scripts/scatter_field.py:677-688), substeps = 1."""
    dt = cfg["dt"]
    if not cfg["verlet"]:
        a_tok, field = force(pos, vel, mass, mask, field, params, cfg)
        dv = a_tok * dt
        return pos + dt * (vel + 0.5 * dv), vel + dv, field, dv
    field, a_prev = field
    if a_prev is None:
        a_prev, field = force(pos, vel, mass, mask, field, params, cfg)
    pos_new = pos + dt * vel + 0.5 * dt ** 2 * a_prev
    a_new, field = force(pos_new, vel + dt * a_prev, mass, mask, field, params, cfg)
    dv = 0.5 * dt * (a_prev + a_new)
    return pos_new, vel + dv, (field, a_new), dv
