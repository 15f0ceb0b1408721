"""Pins of the oracle by what the mathematics fixes (independent of the
reference's files): closed forms, conservation laws, limits and brute force.
Each check is chosen so that a plausible slip (sign, dropped term, wrong
index, transposed operand) fails it."""
import itertools
import math

import numpy as np
import pytest

import oracle as O
from helpers import rel_err, smoke_cube


def _two_body(eps=0.0):
    pos = np.array([[-0.5, 0.0, 0.2], [0.5, 0.0, -0.2]])
    mass = np.array([2.0, 3.0])
    return pos, mass


@pytest.mark.parametrize("eps", [0.0, 0.1])
def test_two_body_closed_form(eps):
    pos, m = _two_body()
    G = 1.7
    for acc in (O.direct_symmetric(pos, m, G, eps), O.direct_targets(pos, m, G, eps)):
        r = pos[1] - pos[0]
        s = (r @ r + eps * eps) ** 1.5
        assert np.allclose(acc[0], G * m[1] * r / s, rtol=1e-15, atol=0)
        assert np.allclose(acc[1], -G * m[0] * r / s, rtol=1e-15, atol=0)
        # reference smoke check: attraction and momentum balance (smoke_tests.cpp:54-69)
        assert acc[0, 0] > 0 and acc[1, 0] < 0
        assert abs(m @ acc[:, 0]) <= 1e-12


def test_brute_force_small_n_and_forms_agree():
    rng = np.random.default_rng(3)
    pos, m = rng.normal(size=(37, 3)), rng.uniform(0.1, 1.0, 37)
    eps, G = 0.05, 1.0
    brute = np.zeros_like(pos)
    for i, j in itertools.product(range(37), range(37)):
        if i != j:
            r = pos[j] - pos[i]
            brute[i] += G * m[j] * r / (r @ r + eps * eps) ** 1.5
    assert rel_err(O.direct_targets(pos, m, G, eps), brute).max() < 1e-13
    sym = O.direct_symmetric(pos, m, G, eps)
    assert rel_err(sym, brute).max() < 1e-13
    # symmetric form conserves momentum to rounding
    assert np.abs(m @ sym).max() < 1e-13 * np.abs(m[:, None] * sym).sum()


def test_targets_range_and_threads():
    rng = np.random.default_rng(4)
    pos, m = rng.normal(size=(300, 3)), rng.uniform(0.1, 1.0, 300)
    full = O.direct_targets(pos, m, 1.0, 0.02)
    part = O.direct_targets(pos, m, 1.0, 0.02, begin=100, end=170)
    assert np.array_equal(part[100:170], full[100:170])
    assert not part[:100].any() and not part[170:].any()
    assert np.array_equal(O.direct_targets(pos, m, 1.0, 0.02, threads=4), full)


def test_coincident_particles_without_softening_are_skipped():
    pos = np.array([[0.0, 0, 0], [0.0, 0, 0], [1.0, 0, 0]])
    m = np.ones(3)
    for acc in (O.direct_symmetric(pos, m, 1.0, 0.0), O.direct_targets(pos, m, 1.0, 0.0)):
        assert np.isfinite(acc).all()
        assert acc[0, 0] == pytest.approx(1.0) and acc[2, 0] == pytest.approx(-2.0)


def test_kick_drift_closed_form():
    v, a, x = np.ones((4, 3)), np.full((4, 3), 2.0), np.zeros((4, 3))
    assert np.array_equal(O.kick(v, a, 0.25), np.full((4, 3), 1.5))
    assert np.array_equal(O.drift(x, v, 0.5), np.full((4, 3), 0.5))


def test_leapfrog_circular_orbit_and_reversibility():
    # equal masses on a circular orbit: period, energy and reversibility
    G, M, R = 1.0, 1.0, 1.0
    pos = np.array([[-R / 2, 0, 0], [R / 2, 0, 0]])
    vc = math.sqrt(G * M / (2 * R))
    vel = np.array([[0, -vc, 0], [0, vc, 0]])
    m = np.array([M, M])
    solver = O.make_solver(O.SOLVER_DIRECT_SYMMETRIC, G, 0.0)
    acc = O.direct_symmetric(pos, m, G, 0.0)
    period = 2 * math.pi * math.sqrt(R ** 3 / (2 * G * M))
    steps = 2000
    dt = period / steps
    e0 = O.diagnostics(pos, vel, m, G, 0.0)["total_energy"]
    x1, v1, a1 = O.leapfrog(solver, pos, vel, acc, m, dt, steps)
    assert np.abs(x1 - pos).max() < 2e-4
    assert abs(O.diagnostics(x1, v1, m, G, 0.0)["total_energy"] - e0) < 1e-9
    # running backwards returns to the start (KDK is time-reversible)
    x2, v2, _ = O.leapfrog(solver, x1, -v1, a1, m, dt, steps)
    assert np.abs(x2 - pos).max() < 1e-10 and np.abs(v2 + vel).max() < 1e-10


def test_diagnostics_closed_form():
    pos = np.array([[1.0, 0, 0], [0, 2.0, 0], [0, 0, 3.0]])
    vel = np.array([[0, 1.0, 0], [0, 0, 2.0], [3.0, 0, 0]])
    m = np.array([1.0, 2.0, 3.0])
    d = O.diagnostics(pos, vel, m, 2.0, 0.5)
    assert d["total_mass"] == 6.0
    assert d["kinetic_energy"] == pytest.approx(0.5 * (1 + 2 * 4 + 3 * 9))
    L = sum(mi * np.cross(x, v) for mi, x, v in zip(m, pos, vel))
    assert np.allclose([d[f"angular_momentum_{c}"] for c in "xyz"], L)
    assert np.allclose([d[f"center_of_mass_{c}"] for c in "xyz"], m @ pos / 6.0)
    pe = -sum(2.0 * m[i] * m[j] / math.sqrt(np.sum((pos[j] - pos[i]) ** 2) + 0.25)
              for i in range(3) for j in range(i + 1, 3))
    assert d["potential_energy"] == pytest.approx(pe, rel=1e-15)


# ---------------------------------------------------------------- multipoles
def _brute_moments(offsets, masses):
    e = O.exponents()
    mom = np.zeros(35)
    for k, (a, b, c) in enumerate(e):
        if a + b + c >= 2:
            mom[k] = np.sum(masses * offsets[:, 0] ** a * offsets[:, 1] ** b * offsets[:, 2] ** c)
    return mom


def test_exponent_order_graded():
    e = O.exponents()
    deg = e.sum(axis=1)
    assert list(deg) == sorted(deg) and len({tuple(r) for r in e}) == 35
    assert tuple(e[4]) == (2, 0, 0) and tuple(e[5]) == (1, 1, 0) and tuple(e[34]) == (0, 0, 4)


def test_p2m_equals_brute_force_and_m2m_shift():
    rng = np.random.default_rng(11)
    pts, ms = rng.normal(size=(40, 3)) * 0.3, rng.uniform(0.5, 1.5, 40)
    groups = [slice(0, 13), slice(13, 29), slice(29, 40)]
    com = ms @ pts / ms.sum()
    direct = np.zeros(35)
    for p, mi in zip(pts, ms):
        direct = O.p2m(direct, p - com, mi)
    assert np.allclose(direct, _brute_moments(pts - com, ms), rtol=1e-13, atol=1e-15)
    shifted = np.zeros(35)
    for g in groups:
        cm = ms[g] @ pts[g] / ms[g].sum()
        child = np.zeros(35)
        for p, mi in zip(pts[g], ms[g]):
            child = O.p2m(child, p - cm, mi)
        shifted = O.m2m(shifted, child, cm - com, ms[g].sum())
    assert np.allclose(shifted, direct, rtol=1e-12, atol=1e-14)


def test_m2p_order0_is_point_mass_force():
    rng = np.random.default_rng(12)
    for _ in range(20):
        t, c = rng.normal(size=3), rng.normal(size=3) + 3.0
        mom = rng.normal(size=35)
        got = O.m2p(t, c, 2.5, mom, G=1.3, eps=0.05, order=0)
        want = O.softened_acceleration(t, c, 2.5, G=1.3, eps=0.05)
        assert np.allclose(got, want, rtol=1e-14, atol=0)


def test_m2p_converges_with_order_to_direct_sum():
    rng = np.random.default_rng(13)
    pts, ms = rng.uniform(-0.25, 0.25, size=(30, 3)), rng.uniform(0.5, 1.5, 30)
    com = ms @ pts / ms.sum()
    mom = np.zeros(35)
    for p, mi in zip(pts, ms):
        mom = O.p2m(mom, p - com, mi)
    target = np.array([2.0, 0.7, -0.4])
    exact = sum(O.softened_acceleration(target, p, mi, 1.0, 0.02) for p, mi in zip(pts, ms))
    errs = [np.linalg.norm(O.m2p(target, com, ms.sum(), mom, 1.0, 0.02, p) - exact)
            for p in (0, 2, 4)]
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] / np.linalg.norm(exact) < 1e-4
    # the p=4 remainder is O((size/distance)^5): halving the cell divides it by ~32
    mom_half = np.zeros(35)
    for p, mi in zip(pts, ms):
        mom_half = O.p2m(mom_half, (p - com) * 0.5, mi)
    exact_half = sum(O.softened_acceleration(target, com + (p - com) * 0.5, mi, 1.0, 0.02)
                     for p, mi in zip(pts, ms))
    err_half = np.linalg.norm(O.m2p(target, com, ms.sum(), mom_half, 1.0, 0.02, 4) - exact_half)
    assert 16 < errs[2] / err_half < 64
    # orders 1 and 3 are raised to 2 and 4 (multipole.cpp:275-278)
    assert np.array_equal(O.m2p(target, com, ms.sum(), mom, 1.0, 0.02, 1),
                          O.m2p(target, com, ms.sum(), mom, 1.0, 0.02, 2))
    # equal masses at +-x about their centre of mass (no dipole): the p=4
    # Taylor polynomial leaves only degree >= 6 terms, O(|x|^6 / |d|^8)
    x = np.array([0.01, -0.02, 0.015])
    two = O.p2m(O.p2m(np.zeros(35), x, 1.0), -x, 1.0)
    got = O.m2p(target, np.zeros(3), 2.0, two, 1.0, 0.02, 4)
    want = (O.softened_acceleration(target, x, 1.0, 1.0, 0.02)
            + O.softened_acceleration(target, -x, 1.0, 1.0, 0.02))
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-9


# ---------------------------------------------------------------- FMM
def test_fmm_reference_smoke_threshold():
    # reference smoke test: 80 particles, theta 0.35, leaf 4, p 4: mean error < 0.25
    pos, m = smoke_cube()
    direct = O.direct_symmetric(pos, m, 1.0, 0.03)
    acc, stats = O.fmm(pos, m, 1.0, 0.03, theta=0.35, leaf_capacity=4, expansion_order=4)
    assert rel_err(acc, direct).mean() < 0.25
    assert stats["node_count"] % 8 == 1 and stats["leaf_count"] >= 80 // 4


def test_fmm_without_far_field_is_direct():
    rng = np.random.default_rng(21)
    pos, m = rng.normal(size=(500, 3)), rng.uniform(0.5, 1.0, 500)
    acc, stats = O.fmm(pos, m, 1.0, 0.02, theta=1e-9, leaf_capacity=8)
    assert stats["far_interactions"] == 0
    assert rel_err(acc, O.direct_targets(pos, m, 1.0, 0.02)).max() < 1e-13


def test_fmm_accuracy_improves_with_theta_and_order():
    rng = np.random.default_rng(22)
    pos, m = rng.normal(size=(2000, 3)), np.full(2000, 1 / 2000)
    ref = O.direct_targets(pos, m, 1.0, 0.01)
    e = {}
    for th, p in [(0.8, 0), (0.8, 2), (0.8, 4), (0.4, 4)]:
        acc, _ = O.fmm(pos, m, 1.0, 0.01, theta=th, leaf_capacity=16, expansion_order=p)
        e[(th, p)] = np.median(rel_err(acc, ref))
    assert e[(0.8, 0)] > e[(0.8, 2)] > e[(0.8, 4)] > e[(0.4, 4)]
    assert e[(0.4, 4)] < 1e-5


def test_fmm_target_range_and_empty():
    rng = np.random.default_rng(23)
    pos, m = rng.normal(size=(300, 3)), rng.uniform(0.5, 1.0, 300)
    full, _ = O.fmm(pos, m, 1.0, 0.02, theta=0.5, leaf_capacity=8)
    part, _ = O.fmm(pos, m, 1.0, 0.02, theta=0.5, leaf_capacity=8, begin=50, end=120)
    assert np.array_equal(part[50:120], full[50:120]) and not part[:50].any()
    acc, stats = O.fmm(np.zeros((0, 3)), np.zeros(0))
    assert acc.shape == (0, 3) and stats["node_count"] == 0
