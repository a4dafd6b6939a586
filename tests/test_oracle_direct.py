"""Pins for oracle.direct: reference fixtures
(web/scatter/tests/truth_*.json, the two-body answer in verify.mjs), closed
forms and invariants."""
import json
import os

import numpy as np
import pytest

from oracle import direct

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_two_body_known_answer():
    # web/scatter/tests/verify.mjs:42-47: bodies 3 apart, eps 0.5 -> +0.10663718 on body 0
    a = direct.direct_acc(np.array([[0.0, 0.0], [3.0, 0.0]]), 0.5)
    assert abs(a[0, 0] - 0.10663718) < 1e-7 and abs(a[1, 0] + 0.10663718) < 1e-7
    assert a[0, 1] == 0.0


def test_newton_third_law_and_target_subset():
    rng = np.random.default_rng(1)
    pos = rng.normal(size=(300, 3))
    m = rng.uniform(0.5, 2, 300)
    a = direct.direct_acc(pos, 0.1, m)
    assert np.abs((m[:, None] * a).sum(0)).max() < 1e-11
    t = np.array([5, 17, 299])
    np.testing.assert_allclose(direct.direct_acc(pos, 0.1, m, targets=t, chunk=2), a[t], rtol=1e-14)


def test_kdk_rollout_matches_reference_fixture():
    # web/scatter/tests/truth_ref.json (gravity_sim.rollout, n=100, dt=0.1,
    # substeps=4, eps=0.5) from truth_init.json; truth.mjs:50 uses 1e-6 at t<=100
    init = json.load(open(os.path.join(GOLD, "truth_init.json")))
    ref = json.load(open(os.path.join(GOLD, "truth_ref.json")))
    n = init["n"]
    pos = np.array(init["pos"]).reshape(n, 2)
    vel = np.array(init["vel"]).reshape(n, 2)
    ps, vs = direct.kdk_rollout(pos, vel, 100, dt=init["dt"])
    for t in (10, 100):
        assert np.abs(ps[t].reshape(-1) - np.array(ref["pos"][str(t)])).max() < 1e-6
        assert abs(direct.energy(ps[t], vs[t], 0.5) - ref["E"][str(t)]) < 1e-8
    assert abs(direct.energy(ps[0], vs[0], 0.5) - ref["E"]["0"]) < 1e-10


def test_momentum_and_kinetic():
    vel = np.array([[1.0, 2.0, 3.0], [-1.0, 0.0, 1.0]])
    m = np.array([2.0, 3.0])
    np.testing.assert_allclose(direct.momentum(vel, m), [-1.0, 4.0, 9.0])
    assert direct.kinetic_energy(vel, m) == pytest.approx(0.5 * (2 * 14 + 3 * 2))


def test_ewald_cubic_lattice_zero_force():
    g = (np.arange(3) + 0.5) / 3
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    a = direct.ewald_acc(pos, np.ones(27), np.arange(27))
    assert np.abs(a).max() < 1e-9


def test_ewald_alpha_independent_and_momentum():
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 1, (40, 3))
    m = rng.uniform(0.5, 1.5, 40)
    a1 = direct.ewald_acc(pos, m, np.arange(40), alpha_l=6.0)
    a2 = direct.ewald_acc(pos, m, np.arange(40), alpha_l=10.0)
    np.testing.assert_allclose(a1, a2, atol=1e-7 * np.abs(a1).max())
    assert np.abs((m[:, None] * a1).sum(0)).max() < 1e-8 * np.abs(a1).max()


def test_ewald_compact_cluster_tends_to_isolated():
    # a cluster of size 1e-3 L: periodic minus isolated force is the tidal +
    # background term ~ (4 pi/3) M r / L^3, tiny against M/r^2
    rng = np.random.default_rng(3)
    pos = 0.5 + 1e-3 * rng.normal(size=(30, 3))
    iso = direct.direct_acc(pos, 0.0)
    per = direct.ewald_acc(pos, np.ones(30), np.arange(30))
    assert np.abs(per - iso).max() < 1e-5 * np.abs(iso).max()


def test_ewald_softening_only_changes_near_pairs():
    pos = np.array([[0.1, 0.1, 0.1], [0.4, 0.1, 0.1]])
    a0 = direct.ewald_acc(pos, np.ones(2), [0], eps=0.0)
    a1 = direct.ewald_acc(pos, np.ones(2), [0], eps=0.05)
    d_soft = 0.3 / (0.09 + 0.0025) ** 1.5 - 0.3 / 0.09 ** 1.5
    np.testing.assert_allclose(a1[0, 0] - a0[0, 0], d_soft, rtol=1e-9)
