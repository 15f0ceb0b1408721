"""hip-tree on the GPU against the oracle's Barnes-Hut solver (the
reference's cpp/fmm/quadtree.cpp, restated and pinned bit for bit against
the compiled reference in test_oracle_vs_ref.py).

Every target takes the reference's own acceptance decisions (in FP64), so
the walk's work counts (M2P evaluations, directed pairs) must equal the
oracle's exactly; the FP64 evaluation agrees with the oracle to rounding
(different summation order and an exact reformulation of the M2P), the FP32
("mixed") evaluation to FP32 rounding."""
import numpy as np
import pytest

import oracle as O
from conftest import load_snapshot
from helpers import rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11
MIXED_TOL = 1e-5


def _tree(gpu, precision, pos, mass, eps, theta, leaf, order=4, begin=None, end=None):
    with gpu.Engine("hip-tree", precision, softening=eps, theta=theta, leaf_capacity=leaf,
                    expansion_order=order) as eng:
        eng.upload(pos, None, mass)
        eng.compute_accelerations(begin, end)
        return eng.download()[2], eng.fmm_stats()


def _check(gpu, pos, m, eps, theta, leaf, order=4):
    want, st = O.tree(pos, m, 1.0, eps, theta, leaf, 32, order)
    a64, s64 = _tree(gpu, "fp64", pos, m, eps, theta, leaf, order)
    for k in ("node_count", "leaf_count", "m2p_evaluations", "p2p_pairs"):
        assert s64[k] == st[k], (k, s64[k], st[k])
    assert rel_err(a64, want).max() < FP64_TOL
    a32, s32 = _tree(gpu, "mixed", pos, m, eps, theta, leaf, order)
    assert s32["m2p_evaluations"] == st["m2p_evaluations"]
    assert rel_err(a32, want).max() < MIXED_TOL
    return want, a64, a32


@pytest.mark.parametrize("theta,leaf,order,eps", [(0.6, 16, 4, 0.02), (0.35, 4, 4, 0.01),
                                                  (0.8, 32, 2, 0.02), (0.6, 1, 0, 0.02),
                                                  (0.7, 8, 3, 0.0), (1.2, 16, 1, 0.05)])
def test_plummer_against_oracle(gpu, theta, leaf, order, eps):
    pos, vel, m = gpu.generators.plummer(3000, seed=21)
    _check(gpu, pos, m, eps, theta, leaf, order)


def test_two_disk_readme_snapshot(gpu):
    m, x, v, a = load_snapshot("readme1000_snapshot_000000.csv.gz")
    _check(gpu, x, m, 0.025, 0.58, 16)


def test_massless_particles_still_feel_forces(gpu):
    # the reference walks for every particle: massless ones in massless
    # leaves get accelerations too (unlike the FMM, which skips such leaves)
    pos, vel, m = gpu.generators.plummer(2000, seed=22)
    m[:300] = 0.0
    want, a64, a32 = _check(gpu, pos, m, 0.02, 0.6, 16)
    assert np.abs(want[:300]).max() > 0 and np.abs(a32[:300]).max() > 0


def test_target_range_and_kdk_steps(gpu):
    pos, vel, m = gpu.generators.plummer(2500, seed=23)
    want, _ = O.tree(pos, m, 1.0, 0.02, 0.6, 16, 32, 4)
    got, st = _tree(gpu, "fp64", pos, m, 0.02, 0.6, 16, 4, 500, 1800)
    assert rel_err(got[500:1800], want[500:1800]).max() < FP64_TOL
    # only the range's targets walk: the counts are the range's own
    _, ost = O.tree(pos, m, 1.0, 0.02, 0.6, 16, 32, 4, begin=500, end=1800)
    assert st["m2p_evaluations"] == ost["m2p_evaluations"]
    assert st["p2p_pairs"] == ost["p2p_pairs"]
    solver = O.make_solver(O.SOLVER_TREE, 1.0, 0.02, 0.6, 16, 32, 4)
    rp, rv, ra = O.leapfrog(solver, pos, vel, want, m, 1e-3, 10)
    with gpu.Engine("hip-tree", "fp64", softening=0.02, theta=0.6, leaf_capacity=16) as eng:
        eng.upload(pos, vel, m)
        eng.compute_accelerations()
        eng.step(1e-3, 10)
        gp, gv, ga = eng.download()
    assert np.abs(gp - rp).max() < 1e-11
    assert rel_err(ga, ra).max() < 1e-9


def test_plummer_full_size_sampled(gpu):
    # N = 2^20 Plummer sphere at the tree benchmark's settings (theta 0.6,
    # leaves <= 16, p = 4, eps 0.01).  The FP32 walk over every particle, on
    # its first 2048 targets (i.i.d. draws, so a random sample): against the
    # oracle tree, and against FP64 direct sums for the method's own error.
    # The same targets as a range give the oracle's exact walk counts.
    pos, vel, m = gpu.generators.plummer(1 << 20, seed=1)
    k = 2048
    got, st = _tree(gpu, "mixed", pos, m, 0.01, 0.6, 16)
    want, ost = O.tree(pos, m, 1.0, 0.01, 0.6, 16, 32, 4, begin=0, end=k)
    assert st["node_count"] == ost["node_count"] and st["leaf_count"] == ost["leaf_count"]
    assert rel_err(got[:k], want[:k]).max() < MIXED_TOL
    part, pst = _tree(gpu, "mixed", pos, m, 0.01, 0.6, 16, 4, 0, k)
    assert pst["m2p_evaluations"] == ost["m2p_evaluations"]
    assert pst["p2p_pairs"] == ost["p2p_pairs"]
    assert np.array_equal(part[:k], got[:k])
    direct = O.direct_targets(pos, m, 1.0, 0.01, 0, k, threads=0)
    err = rel_err(got[:k], direct[:k])
    assert np.median(err) < 2e-4 and err.max() < 1e-2


@pytest.mark.parametrize("leaf,eps", [(4, 0.0), (8, 0.02)])
def test_duplicates_at_depth_cap(gpu, leaf, eps):
    # the oracle's depth-cap case (pinned against the compiled reference):
    # 40 coincident particles reach depth 32; with eps = 0 their pairs are
    # skipped (s2 == 0) and every other pair counts
    rng = np.random.default_rng(8)
    pos = rng.normal(size=(300, 3))
    pos[:75] *= 0.05
    pos[:40] = pos[0]
    m = rng.uniform(0.1, 1.0, 300)
    want, st = O.tree(pos, m, 1.0, eps, 0.6, leaf, 32, 4)
    assert st["max_depth_reached"] == 32
    a64, s64 = _tree(gpu, "fp64", pos, m, eps, 0.6, leaf)
    for k in ("node_count", "leaf_count", "m2p_evaluations", "p2p_pairs"):
        assert s64[k] == st[k], (k, s64[k], st[k])
    assert s64["max_depth"] == 32
    assert rel_err(a64, want).max() < FP64_TOL
    a32, s32 = _tree(gpu, "mixed", pos, m, eps, 0.6, leaf)
    assert s32["p2p_pairs"] == st["p2p_pairs"]
    assert rel_err(a32, want).max() < MIXED_TOL
