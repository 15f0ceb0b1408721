"""The oracle against the reference's own solvers compiled from their
sources (oracle/_ref/libgcref.so, built by oracle/Makefile when the
reference is present), bit for bit on inputs the goldens do not cover."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def _cloud(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3)) * scale
    pos[: n // 4] *= 0.05  # a dense core: deep leaves
    return pos, rng.uniform(0.1, 1.0, n)


@pytest.mark.parametrize("eps", [0.0, 0.02])
def test_direct_forms(eps):
    pos, m = _cloud(700, 1)
    pos[5] = pos[6]  # coincident pair (skipped when eps == 0)
    assert np.array_equal(O.direct_symmetric(pos, m, 1.3, eps), O.ref_direct_symmetric(pos, m, 1.3, eps))
    assert np.array_equal(O.direct_targets(pos, m, 1.3, eps, 100, 400),
                          O.ref_direct_targets(pos, m, 1.3, eps, 100, 400))


def test_leapfrog_and_diagnostics():
    pos, m = _cloud(300, 2)
    vel = np.random.default_rng(3).normal(size=(300, 3)) * 0.1
    acc = O.direct_symmetric(pos, m, 1.0, 0.05)
    solver = O.make_solver(O.SOLVER_DIRECT_SYMMETRIC, 1.0, 0.05)
    got = O.leapfrog(solver, pos, vel, acc, m, 0.003, 7)
    want = O.ref_leapfrog(0, pos, vel, acc, m, 1.0, 0.05, 0.003, 7)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    assert O.diagnostics(*got[:2], m, 1.0, 0.05) == O.ref_diagnostics(*want[:2], m, 1.0, 0.05)


def test_multipole_kernels():
    rng = np.random.default_rng(4)
    for _ in range(50):
        mom = rng.normal(size=35)
        off = rng.normal(size=3) * 0.3
        assert np.array_equal(O.p2m(mom, off, 0.7), O.ref_p2m(mom, off, 0.7))
        child = rng.normal(size=35)
        assert np.array_equal(O.m2m(mom, child, off, 1.9), O.ref_m2m(mom, child, off, 1.9))
        off[rng.integers(3)] = 0.0  # zero shift components exercise the skip rule
        assert np.array_equal(O.m2m(mom, child, off, 1.9), O.ref_m2m(mom, child, off, 1.9))
        t, c = rng.normal(size=3), rng.normal(size=3) * 2
        for p in range(-1, 6):
            assert np.array_equal(O.m2p(t, c, 1.5, mom, 1.1, 0.03, p),
                                  O.ref_m2p(t, c, 1.5, mom, 1.1, 0.03, p))


@pytest.mark.parametrize("theta,leaf,order,eps", [(0.5, 16, 4, 0.01), (0.35, 4, 4, 0.03),
                                                  (0.8, 32, 2, 0.0), (0.6, 1, 0, 0.02),
                                                  (1.2, 8, 3, 0.02)])
def test_fmm(theta, leaf, order, eps):
    pos, m = _cloud(1500, 5)
    m[10:20] = 0.0  # massless particles: massless leaves are never evaluated
    got, st = O.fmm(pos, m, 1.0, eps, theta, leaf, 32, order)
    want, rst = O.ref_fmm(pos, m, 1.0, eps, theta, leaf, 32, order)
    assert np.array_equal(got, want)
    for k in ("node_count", "leaf_count", "far_interactions", "near_interactions"):
        assert st[k] == rst[k]


def test_fmm_depth_cap_with_duplicates():
    pos, m = _cloud(200, 6)
    pos[:40] = pos[0]  # 40 coincident particles force depth-capped leaves
    got, st = O.fmm(pos, m, 1.0, 0.02, 0.5, 8, 6, 4)
    want, rst = O.ref_fmm(pos, m, 1.0, 0.02, 0.5, 8, 6, 4)
    assert np.array_equal(got, want) and st["node_count"] == rst["node_count"]


@pytest.mark.parametrize("theta,leaf,order,eps", [(0.6, 16, 4, 0.02), (0.35, 4, 4, 0.01),
                                                  (0.8, 32, 2, 0.0), (0.6, 1, 0, 0.02),
                                                  (1.2, 8, 3, 0.05), (0.5, 8, 1, 0.02)])
def test_tree(theta, leaf, order, eps):
    # the Barnes-Hut walk per particle (cpp/fmm/quadtree.cpp): massless
    # particles in massless leaves still get accelerations; p = 1 and 3 run
    # as 2 and 4 (multipole.cpp:275-278)
    pos, m = _cloud(1200, 7)
    m[10:30] = 0.0
    got, _ = O.tree(pos, m, 1.0, eps, theta, leaf, 32, order)
    assert np.array_equal(got, O.ref_tree(pos, m, 1.0, eps, theta, leaf, order))


def test_tree_duplicates_at_depth_cap_and_leapfrog():
    # 40 coincident particles with leaves of 4 reach the depth cap of 32;
    # with eps = 0 their pairs are skipped (s2 == 0) but every other pair
    # counts; then 5 KDK steps through each solver's own loop
    pos, m = _cloud(300, 8)
    pos[:40] = pos[0]
    got, st = O.tree(pos, m, 1.0, 0.0, 0.6, 4, 32, 4)
    assert np.array_equal(got, O.ref_tree(pos, m, 1.0, 0.0, 0.6, 4, 4))
    assert st["max_depth_reached"] == 32
    vel = np.random.default_rng(9).normal(size=pos.shape) * 0.1
    pos[:40] += np.random.default_rng(10).normal(size=(40, 3)) * 1e-3
    acc, _ = O.tree(pos, m, 1.0, 0.02, 0.6, 8, 32, 4)
    solver = O.make_solver(O.SOLVER_TREE, 1.0, 0.02, 0.6, 8, 32, 4)
    mine = O.leapfrog(solver, pos, vel, acc, m, 0.002, 5)
    ref = O.ref_leapfrog(3, pos, vel, acc, m, 1.0, 0.02, 0.002, 5, 0.6, 8, 4)
    for a, b in zip(mine, ref):
        assert np.array_equal(a, b)
