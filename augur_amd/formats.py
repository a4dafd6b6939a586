"""The reference's exported-weights format (web/scatter/weights.json +
weights.bin): a JSON manifest {"checkpoint", "iters", "config", "tensors":
{name: {"shape", "offset"}}} with offsets counted in float32 values into a
little-endian float32 blob.  Read with numpy (nothing in the files is
executed); written in the same layout, tensors in state_dict order.

``scatter_field_from_web`` builds the drop-in ScatterField that the manifest's
config describes (the reference's own loader is the JS port,
web/scatter/js/scatter_model.js:231-459, and the exporter
scripts/export_web_weights_scatter.py)."""
import json

import numpy as np
import torch


def read_web_weights(json_path, bin_path):
    """-> (state_dict of float32 CPU tensors, config dict, manifest dict)."""
    with open(json_path) as f:
        man = json.load(f)
    raw = np.fromfile(bin_path, dtype="<f4")
    sd = {}
    for name, t in man["tensors"].items():
        shape = [int(d) for d in t["shape"]]
        n = int(np.prod(shape)) if shape else 1
        off = int(t["offset"])
        if off < 0 or off + n > raw.size:
            raise ValueError(f"{name}: offset {off} + {n} values past the end of {bin_path} ({raw.size})")
        sd[name] = torch.from_numpy(raw[off: off + n].copy().reshape(shape))
    return sd, dict(man.get("config", {})), man


def write_web_weights(state_dict, config, json_path, bin_path, checkpoint="", iters=0):
    """Writes a state_dict (float tensors) in the manifest layout."""
    tensors, chunks, off = {}, [], 0
    for name, t in state_dict.items():
        a = t.detach().cpu().to(torch.float32).contiguous().numpy().astype("<f4")
        tensors[name] = {"shape": list(a.shape), "offset": off}
        chunks.append(a.reshape(-1))
        off += a.size
    blob = np.concatenate(chunks) if chunks else np.zeros(0, "<f4")
    blob.tofile(bin_path)
    with open(json_path, "w") as f:
        json.dump({"checkpoint": checkpoint, "iters": int(iters), "config": config, "tensors": tensors}, f, indent=1)


def scatter_field_from_web(json_path, bin_path, device="cuda", net="unet"):
    """The ScatterField of a web manifest's config with its weights loaded
    (the exporter's model: UNet, recurrent, the manifest's flags)."""
    from .scatter_field import ScatterField
    sd, cfg, _ = read_web_weights(json_path, bin_path)
    keys = ("grid", "extent", "dt", "levels", "pp", "knn", "eps", "in_scale", "momfix", "potential", "kernel", "split",
            "verlet")
    model = ScatterField(net=net, **{k: cfg[k] for k in keys if k in cfg}).to(device)
    model.load_state_dict(sd)
    return model
