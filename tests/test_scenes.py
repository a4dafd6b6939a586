"""The seeded 3D scene generators of the workloads (augur_amd.scenes, host
numpy): pinned by the Plummer cumulative mass profile M(<r) = r^3/(r^2+a^2)^1.5
(truncated at rmax) and the layout of the clumpy mixture."""
import numpy as np

from augur_amd import scenes


def test_plummer_mass_profile():
    rng = np.random.default_rng(4)
    a = 0.05
    p = scenes.plummer(20000, a, rng)
    r = np.linalg.norm(p - 0.5, axis=1)
    assert r.max() < 0.45
    for rr in (0.02, 0.05, 0.1, 0.2):
        expect = (rr ** 3 / (rr ** 2 + a ** 2) ** 1.5) / (0.45 ** 3 / (0.45 ** 2 + a ** 2) ** 1.5)
        assert abs((r < rr).mean() - expect) < 0.012
    c = p.mean(0) - 0.5
    assert np.abs(c).max() < 0.005


def test_clumpy_layout():
    rng = np.random.default_rng(5)
    p = scenes.clumpy(8192, rng)
    assert p.shape == (8192, 3) and p.min() >= 0 and p.max() < 1
    # half uniform (per-octant counts of the first half within 5 sigma of 512)
    u = p[:4096]
    oct_ = (u[:, 0] > 0.5) * 4 + (u[:, 1] > 0.5) * 2 + (u[:, 2] > 0.5)
    assert np.abs(np.bincount(oct_, minlength=8) - 512).max() < 5 * np.sqrt(512 * 7 / 8)


def test_cfg1_scene():
    # Plummer cold collapse: N=65,536 at a=0.05 centred, zero velocities, FP32 in [0, 1)
    pos, vel = scenes.cfg1()
    assert pos.shape == (65536, 3) and pos.dtype == np.float32 and (vel == 0).all()
    assert pos.min() >= 0 and pos.max() < 1
    r = np.linalg.norm(pos.astype(np.float64) - 0.5, axis=1)
    a = 0.05
    half = np.median(r)   # half-mass radius of the truncated profile, ~1.30 a
    m = lambda x: x ** 3 / (x ** 2 + a ** 2) ** 1.5
    assert abs(m(half) / m(0.45) - 0.5) < 0.01
