"""cic_density with the signature of model/tiered/energy.py:9-34, on libaugur_hip.so."""
import torch

from .. import ops


def cic_density(pos, mass, mask, n, h, origin, moments=False):
    """Cloud-in-cell mass density (B, 1, n, n) (or (B, 1, n, n, n) for 3D
    positions); mass falling off the grid is dropped (model/tiered/energy.py:9-34).
    ``origin`` is the grid's lower corner: u = (pos - origin)/h - 0.5.
    The six-channel moments variant is not on the PM path."""
    if moments:
        raise NotImplementedError("moment channels are not part of this build")
    D = pos.shape[-1]
    org = torch.as_tensor(origin, dtype=torch.float32).reshape(-1)
    if org.numel() > 1 and not bool((org == org[0]).all()):
        # per-axis origins: shift positions (exact for power-of-two cells) and use one offset
        pos = pos - org.to(pos.device)
        off = 0.0
    else:
        off = -float(org[0])
    return ops.cic_deposit(pos, mass, n, h, off, False, mask=mask, scale=1.0 / h ** D)
