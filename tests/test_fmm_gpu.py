"""hip-fmm on the GPU against the oracle's FMM (the reference's algorithm).

The device tree, moments and lists follow the reference's geometric rules,
so the FmmStats counts (node, leaf, far and near list totals) must equal the
oracle's exactly.  The FP64 evaluation then agrees with the oracle to
rounding (the same M2P through a different but exact formula: the radial
structure of the softened kernel instead of explicit derivative tensors),
the FP32 ("mixed") evaluation to FP32 rounding.
"""
import numpy as np
import pytest

import oracle as O
from conftest import load_snapshot
from helpers import rel_err, smoke_cube

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11     # max relative difference, FP64 device FMM vs oracle FMM
MIXED_TOL = 1e-5     # max relative difference, FP32 device FMM vs oracle FMM
STAT_KEYS = ("node_count", "leaf_count", "far_interactions", "near_interactions")


def _fmm(gpu, precision, pos, mass, eps, theta, leaf, order=4, vel=None, begin=None, end=None):
    with gpu.Engine("hip-fmm", precision, softening=eps, theta=theta, leaf_capacity=leaf,
                    expansion_order=order) as eng:
        eng.upload(pos, vel, mass)
        eng.compute_accelerations(begin, end)
        acc = eng.download()[2]
        return acc, eng.fmm_stats()


def _check(gpu, pos, m, eps, theta, leaf, order=4):
    want, st = O.fmm(pos, m, 1.0, eps, theta, leaf, 32, order)
    a64, s64 = _fmm(gpu, "fp64", pos, m, eps, theta, leaf, order)
    for k in STAT_KEYS:
        assert s64[k] == st[k], k
    assert rel_err(a64, want).max() < FP64_TOL
    a32, s32 = _fmm(gpu, "mixed", pos, m, eps, theta, leaf, order)
    assert rel_err(a32, want).max() < MIXED_TOL
    return want, a64, a32, s64


def test_readme_golden_step0(gpu):
    m, x, v, a = load_snapshot("readme1000_snapshot_000000.csv.gz")
    want, a64, a32, st = _check(gpu, x, m, 0.025, 0.58, 16)
    assert np.array_equal(want, a)  # the oracle is the golden here
    assert st["max_depth"] >= 1 and st["eval_ms"] > 0


@pytest.mark.parametrize("theta,leaf,order,eps", [(0.5, 64, 4, 0.01), (0.35, 4, 4, 0.03),
                                                  (0.8, 32, 2, 0.02), (0.6, 8, 0, 0.02),
                                                  (0.7, 16, 3, 0.0)])
def test_plummer_against_oracle(gpu, theta, leaf, order, eps):
    pos, vel, m = gpu.generators.plummer(3000, seed=11)
    _check(gpu, pos, m, eps, theta, leaf, order)


def test_large_leaves(gpu):
    # leaves of up to 300 particles: target chunks of the near kernel (24)
    # and far kernel (64) repeat within a leaf, and pieces of near leaves
    # can exceed 64 x 64 sources (the near kernel then finds a source's
    # leaf by binary search instead of the bitmap)
    pos, vel, m = gpu.generators.plummer(6000, seed=17)
    _check(gpu, pos, m, 0.02, 0.5, 300)


@pytest.mark.parametrize("G", [1.0, 3e-4])
def test_equal_masses_leave_out_the_pair_mass(gpu, G):
    # every particle of one mass: the FP32 near field sums d r^-3 and
    # multiplies by G m once per target (its padding sources sit where r^3
    # underflows to zero); one particle 1 ulp heavier takes the per-pair
    # G m form.  Both match the oracle's FMM to FP32 rounding.
    pos, vel, _ = gpu.generators.plummer(4000, seed=23)
    m = np.full(len(pos), 2.5e-4)
    m2 = m.copy()
    m2[1234] = np.nextafter(m2[1234], 1.0)
    for mass in (m, m2):
        want, st = O.fmm(pos, mass, G, 0.01, 0.5, 64, 32, 4)
        with gpu.Engine("hip-fmm", "mixed", softening=0.01, theta=0.5, leaf_capacity=64,
                        expansion_order=4, G=G) as eng:
            eng.upload(pos, vel, mass)
            eng.compute_accelerations()
            got = eng.download()[2]
        assert rel_err(got, want).max() < MIXED_TOL


def test_pieces_of_sixty_four_full_leaves(gpu):
    # a cell-centred 32^3 lattice with leaves of up to 64 particles: 512
    # leaves of exactly 64 and near lists of about 137 leaves, so the near
    # kernel's pieces of 64 near leaves hold exactly 64 x 64 sources, the
    # largest piece its bitmap covers (the sentinel leaf then starts in the
    # bitmap's last word)
    g = (np.arange(32) + 0.5) / 32
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    m = np.random.default_rng(5).uniform(0.5, 1.5, len(pos))
    want, a64, a32, st = _check(gpu, pos, m, 0.01, 0.5, 64)
    assert st["leaf_count"] == 512
    assert st["near_interactions"] >= 2 * 64 * 512


def test_tiny_softening_checks_the_self_pair(gpu):
    # eps = 1e-13: the self pair's factor G m eps^-3 (1e39) would overflow
    # FP32, so the mixed near field drops the self pair by comparing sorted
    # indices instead of adding it as an exact zero; both must match the oracle
    pos, vel, m = gpu.generators.plummer(3000, seed=13)
    for eps in (1e-13, 1e-3):
        want, _ = O.fmm(pos, m, 1.0, eps, 0.5, 16, 32, 4)
        got, _ = _fmm(gpu, "mixed", pos, m, eps, 0.5, 16)
        assert np.isfinite(got).all()
        assert rel_err(got, want).max() < MIXED_TOL, eps


def test_two_disk_and_reference_smoke_threshold(gpu):
    pos, vel, m = gpu.generators.config0()
    _check(gpu, pos, m, 0.05, 0.5, 16)
    # cpp/tests/smoke_tests.cpp:96-123: p=4 FMM (theta 0.35, leaf 4) vs direct
    cube, cm = smoke_cube()
    acc, _ = _fmm(gpu, "mixed", cube, cm, 0.03, 0.35, 4)
    assert rel_err(acc, O.direct_symmetric(cube, cm, 1.0, 0.03)).mean() < 0.25


def test_massless_particles_and_target_range(gpu):
    rng = np.random.default_rng(3)
    pos, m = rng.normal(size=(2000, 3)), rng.uniform(0.5, 1.0, 2000)
    m[:300] = 0.0
    want, _ = O.fmm(pos, m, 1.0, 0.02, 0.5, 8, 32, 4, begin=500, end=1500)
    got, _ = _fmm(gpu, "fp64", pos, m, 0.02, 0.5, 8, begin=500, end=1500)
    assert not got[:500].any() and not got[1500:].any()
    assert rel_err(got[500:1500], want[500:1500]).max() < FP64_TOL


def test_fmm_kdk_steps_track_golden(gpu):
    m, x, v, a = load_snapshot("fmm_1000_snapshot_000000.csv.gz")
    m20, x20, v20, a20 = load_snapshot("fmm_1000_snapshot_000020.csv.gz")
    with gpu.Engine("hip-fmm", "fp64", softening=0.025, theta=0.58, leaf_capacity=16) as eng:
        eng.upload(x, v, m)
        eng.compute_accelerations()
        eng.step(0.01, 20)
        p, vv, acc = eng.download()
    assert np.abs(p - x20).max() < 1e-9
    assert rel_err(acc, a20).max() < 1e-6


def _deep_cases():
    """Inputs whose trees go below 21 levels, up to the depth cap of 32."""
    rng = np.random.default_rng(6)
    # the oracle's own depth-cap case: 40 coincident particles in a cloud
    # with a dense core (cpp/fmm/fmm_solver.cpp:136-155 stops at depth 32)
    pos = rng.normal(size=(200, 3))
    pos[:50] *= 0.05
    pos[:40] = pos[0]
    yield "coincident40", pos, rng.uniform(0.1, 1.0, 200), 8
    # two clumps of 50 coincident particles
    pos = np.zeros((100, 3))
    pos[50:] = 1.0
    yield "coincident50x2", pos, np.ones(100), 8
    # a cluster resolved only by levels 22-31 (spread 1e-8 in a unit cube):
    # the second key word's digits decide the splits
    pos = rng.normal(size=(600, 3))
    pos[:120] = pos[0] + rng.normal(size=(120, 3)) * 1e-8
    yield "cluster1e-8", pos, rng.uniform(0.1, 1.0, 600), 4


@pytest.mark.parametrize("case", ["coincident40", "coincident50x2", "cluster1e-8"])
def test_depth_up_to_32_levels(gpu, case):
    name, pos, m, leaf = next(c for c in _deep_cases() if c[0] == case)
    want, st = O.fmm(pos, m, 1.0, 0.02, 0.5, leaf, 32, 4)
    assert st["max_depth_reached"] > 21
    a64, s64 = _fmm(gpu, "fp64", pos, m, 0.02, 0.5, leaf)
    for k in STAT_KEYS:
        assert s64[k] == st[k], (name, k, s64[k], st[k])
    assert s64["max_depth"] == st["max_depth_reached"]
    assert rel_err(a64, want).max() < FP64_TOL
    a32, s32 = _fmm(gpu, "mixed", pos, m, 0.02, 0.5, leaf)
    assert s32["node_count"] == st["node_count"]
    assert rel_err(a32, want).max() < MIXED_TOL
    # a shallower cap keeps the clumps in one leaf, as the reference does
    want6, st6 = O.fmm(pos, m, 1.0, 0.02, 0.5, leaf, 6, 4)
    with gpu.Engine("hip-fmm", "fp64", softening=0.02, theta=0.5, leaf_capacity=leaf,
                    max_depth=6) as eng:
        eng.upload(pos, None, m)
        eng.compute_accelerations()
        got = eng.download()[2]
        s6 = eng.fmm_stats()
    assert s6["node_count"] == st6["node_count"] and s6["max_depth"] == 6
    assert rel_err(got, want6).max() < FP64_TOL


def test_depth_cap_above_32_is_rejected(gpu):
    pos, _, m = gpu.generators.plummer(500, seed=3)
    with pytest.raises(gpu.GcError) as e:
        with gpu.Engine("hip-fmm", "fp64", softening=0.02, max_depth=33) as eng:
            eng.upload(pos, None, m)
            eng.compute_accelerations()
    assert "max_depth above 32" in str(e.value)


def test_plummer_full_size_sampled(gpu):
    # N = 2^20 Plummer sphere, leaves <= 64, p = 4, theta 0.5, eps 0.01: the
    # FP32 device FMM against the oracle FMM and FP64 direct sums on 4096
    # targets (indices are i.i.d. draws, so a contiguous range is a random
    # sample)
    pos, vel, m = gpu.generators.plummer(1 << 20, seed=1)
    got, st = _fmm(gpu, "mixed", pos, m, 0.01, 0.5, 64)
    want, ost = O.fmm(pos, m, 1.0, 0.01, 0.5, 64, 32, 4, begin=0, end=4096)
    for k in STAT_KEYS:
        assert st[k] == ost[k], k
    assert rel_err(got[:4096], want[:4096]).max() < MIXED_TOL
    direct = O.direct_targets(pos, m, 1.0, 0.01, 0, 4096, threads=0)
    err = rel_err(got[:4096], direct[:4096])
    assert np.median(err) < 1e-4 and err.max() < 1e-2


def test_one_traversal_lists_match_two_pass_and_regrow(gpu):
    # An engine's first build counts, scans and then writes the interaction
    # lists; later builds count and stash them per wave in one traversal
    # (regions sized by the last two-pass build) and move them into place.
    # On the same state both give bit-identical forces.  A problem 60x
    # larger on the same engine outgrows the regions (each wave then holds
    # several leaves' lists), so its first build falls back to writing; it
    # must still match the FP64 evaluation, and its next build (regions
    # resized) must repeat it bit for bit.
    pos, vel, m = gpu.generators.plummer(3000, seed=11)
    big = gpu.generators.plummer(200000, seed=12)
    with gpu.Engine("hip-fmm", "mixed", softening=0.01, theta=0.5, leaf_capacity=16) as eng:
        eng.upload(pos, vel, m)
        eng.compute_accelerations()
        a1 = eng.download()[2]
        eng.compute_accelerations()
        assert np.array_equal(eng.download()[2], a1)
        eng.upload(*big)
        eng.compute_accelerations()
        b1 = eng.download()[2]
        eng.compute_accelerations()
        assert np.array_equal(eng.download()[2], b1)
    b64, _ = _fmm(gpu, "fp64", big[0], big[2], 0.01, 0.5, 16)
    assert rel_err(b1, b64).max() < MIXED_TOL
    want, _ = O.fmm(pos, m, 1.0, 0.01, 0.5, 16, 32, 4)
    assert rel_err(a1, want).max() < MIXED_TOL


@pytest.mark.parametrize("begin,end", [(0, 375), (1000, 1001), (2900, 3000), (700, 2300)])
def test_target_ranges_evaluate_only_their_targets(gpu, begin, end):
    # compute_targets (cpp/fmm/fmm_solver.cpp:24-46): the whole tree and
    # lists, accelerations written for [begin, end) only; the device walks
    # only the range's targets, and the statistics are the whole tree's
    pos, vel, m = gpu.generators.plummer(3000, seed=31)
    want, st = O.fmm(pos, m, 1.0, 0.02, 0.5, 16, 32, 4, begin=begin, end=end)
    for precision, tol in (("fp64", FP64_TOL), ("mixed", MIXED_TOL)):
        got, gst = _fmm(gpu, precision, pos, m, 0.02, 0.5, 16, begin=begin, end=end)
        assert not got[:begin].any() and not got[end:].any()
        assert rel_err(got[begin:end], want[begin:end]).max() < tol, precision
        for k in STAT_KEYS:
            assert gst[k] == st[k], (precision, k)


def test_target_range_steps_match_full_range(gpu):
    # one engine stepping the whole set, another evaluating two halves as
    # ranges: the same accelerations to rounding
    pos, vel, m = gpu.generators.plummer(5000, seed=32)
    with gpu.Engine("hip-fmm", "fp64", softening=0.02, theta=0.5, leaf_capacity=16) as eng:
        eng.upload(pos, vel, m)
        eng.compute_accelerations()
        full = eng.download()[2]
        eng.compute_accelerations(0, 2500)
        eng.compute_accelerations(2500, 5000)
        halves = eng.download()[2]
    assert rel_err(halves, full).max() < FP64_TOL
