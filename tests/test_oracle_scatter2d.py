"""Pins for oracle.scatter2d, the CPU FP64 restatement of the reference's 2D
scatter-field step:

* the reference's own FP64 rollouts web/scatter/tests/vectors.json (copied to
  tests/golden/scatter2d/ with the shipped weights): 20 Verlet steps of the
  model ms_kp_pot_v_g128 (UNet, learned kernel with the split taper, pair
  term, momfix) for 8, 30 and 100 bodies, and the n8 potential after step 20;
* for what those rollouts do not exercise: the kernel convolution (zero and
  "ext" padding) against an O(G^4) real-space sum, the TSC weights' moments,
  the LocalNet composition and the non-Verlet step under a constant field.
"""
import json
import math
import os

import numpy as np
import pytest
import torch

from oracle import scatter2d as S

GOLD = os.path.join(os.path.dirname(__file__), "golden", "scatter2d")
F64 = torch.float64


@pytest.fixture(scope="module")
def model():
    params, cfg = S.load_web_weights(os.path.join(GOLD, "weights.json"), os.path.join(GOLD, "weights.bin"))
    return params, S.default_cfg(**cfg)


@pytest.mark.parametrize("name", ["n8", "n30", "n100"])
def test_rollouts_match_reference_vectors(model, name):
    params, cfg = model
    vec = json.load(open(os.path.join(GOLD, "vectors.json")))
    sc = vec["scenes"][name]
    pos = torch.tensor(sc["pos"][0], dtype=F64)[None]
    vel = torch.tensor(sc["vel"][0], dtype=F64)[None]
    n = pos.shape[1]
    mass = torch.ones(1, n, dtype=F64)
    mask = torch.ones(1, n, dtype=F64)
    field = S.init_field(1, cfg)
    with torch.no_grad():
        for s in range(vec["steps"]):
            pos, vel, field, _ = S.step(pos, vel, mass, mask, field, params, cfg)
            e = max(float((pos[0] - torch.tensor(sc["pos"][s + 1])).abs().max()),
                    float((vel[0] - torch.tensor(sc["vel"][s + 1])).abs().max()))
            # the reference's verifier asks 1e-4 after one step (web/scatter/tests/verify.mjs:38)
            assert e < (1e-6 if s == 0 else 5e-6), (s, e)
    if sc["field0"] is not None:
        assert float((field[0][0, 0] - torch.tensor(sc["field0"])).abs().max()) < 1e-6


def _bruteforce_potential(rho, K, h, lo, hi):
    """phi(i) = h^2 sum_j rho(j) K(i - j) on cells i in [lo, hi)^2 (any i,
    also outside the grid), K indexed by min-image offset on the 2G domain."""
    G = rho.shape[-1]
    out = torch.zeros(hi - lo, hi - lo, dtype=F64)
    for iy in range(lo, hi):
        for ix in range(lo, hi):
            s = 0.0
            for jy in range(G):
                for jx in range(G):
                    s += float(rho[jy, jx]) * float(K[(iy - jy) % (2 * G), (ix - jx) % (2 * G)])
            out[iy - lo, ix - lo] = h * h * s
    return out


@pytest.mark.parametrize("kpad", ["zero", "ext"])
def test_kernel_acc_equals_real_space_sum(model, kpad):
    params, cfg0 = model
    G = 8
    cfg = dict(cfg0, grid=G, extent=8.0, kpad=kpad)
    h = cfg["extent"] / G
    rho = torch.from_numpy(np.random.default_rng(1).uniform(0, 1, (G, G)))
    acc = S.kernel_acc(rho[None, None], params, cfg)[0]
    K = S.kernel_grid(params, cfg)
    phi = _bruteforce_potential(rho, K, h, -1, G + 1)          # cells -1 .. G
    if kpad == "zero":
        phi[0, :] = phi[-1, :] = 0.0
        phi[:, 0] = phi[:, -1] = 0.0
    gx = -(phi[1:-1, 2:] - phi[1:-1, :-2]) / (2 * h)
    gy = -(phi[2:, 1:-1] - phi[:-2, 1:-1]) / (2 * h)
    torch.testing.assert_close(acc[0], gx, rtol=1e-9, atol=1e-12)
    torch.testing.assert_close(acc[1], gy, rtol=1e-9, atol=1e-12)


def test_tsc_weights_moments():
    # the quadratic spline: weights sum to 1, centroid preserved, variance 1/4 cell^2
    cfg = S.default_cfg(grid=16, extent=16.0, deposit="quad")
    pos = torch.from_numpy(np.random.default_rng(2).uniform(-5, 5, (1, 200, 2)))
    u = pos + 8.0 - 0.5
    tot = torch.zeros(1, 200, dtype=F64)
    m1 = torch.zeros(1, 200, 2, dtype=F64)
    m2 = torch.zeros(1, 200, 2, dtype=F64)
    for flat, w in S.corners(pos, cfg):
        c = torch.stack([flat % 16, flat // 16], -1).to(F64)
        tot += w
        m1 += w[..., None] * (c - u)
        m2 += w[..., None] * (c - u) ** 2
    torch.testing.assert_close(tot, torch.ones_like(tot))
    torch.testing.assert_close(m1, torch.zeros_like(m1), atol=1e-12, rtol=0)
    torch.testing.assert_close(m2, torch.full_like(m2, 0.25), atol=1e-12, rtol=0)


def test_localnet_composition():
    # identity input and output convs, zero middle convs: y = gelu(x) exactly
    C, G = 3, 8
    eye = torch.zeros(C, C, 3, 3, dtype=F64)
    eye[range(C), range(C), 1, 1] = 1.0
    params = {"net.inp.weight": eye, "net.inp.bias": torch.zeros(C, dtype=F64),
              "net.out.weight": eye, "net.out.bias": torch.zeros(C, dtype=F64)}
    for k in range(2):
        params[f"net.mid.{k}.weight"] = torch.zeros(C, C, 3, 3, dtype=F64)
        params[f"net.mid.{k}.bias"] = torch.full((C,), 0.3, dtype=F64)
    x = torch.randn(1, C, G, G, dtype=F64)
    y = S.localnet(x, params, 3)
    exp = torch.nn.functional.gelu(x) + 2 * torch.nn.functional.gelu(torch.tensor(0.3, dtype=F64))
    torch.testing.assert_close(y, exp)


def test_non_verlet_step_under_constant_field():
    # a LocalNet whose output is the constant field (ax, ay): the non-Verlet step
    # is x + dt (v + dt a / 2), v + dt a for particles inside the grid
    G, ax, ay = 16, 0.7, -1.3
    cfg = S.default_cfg(grid=G, extent=16.0, layers=2, momfix=False, recurrent=False, verlet=False, dt=0.05)
    params = {"net.inp.weight": torch.zeros(32, 3, 3, 3, dtype=F64), "net.inp.bias": torch.zeros(32, dtype=F64),
              "net.mid.0.weight": torch.zeros(32, 32, 3, 3, dtype=F64),
              "net.mid.0.bias": torch.zeros(32, dtype=F64),
              "net.out.weight": torch.zeros(2, 32, 3, 3, dtype=F64),
              "net.out.bias": torch.tensor([ax, ay], dtype=F64)}
    rng = np.random.default_rng(3)
    pos = torch.from_numpy(rng.uniform(-6, 6, (1, 50, 2)))
    vel = torch.from_numpy(rng.normal(size=(1, 50, 2)))
    one = torch.ones(1, 50, dtype=F64)
    p, v, _, dv = S.step(pos, vel, one, one, S.init_field(1, cfg), params, cfg)
    a = torch.tensor([ax, ay], dtype=F64)
    dt = cfg["dt"]
    torch.testing.assert_close(v, vel + dt * a)
    torch.testing.assert_close(p, pos + dt * (vel + 0.5 * dt * a))
    assert math.isclose(float(dv[0, 0, 0]), dt * ax)
