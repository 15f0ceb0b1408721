"""hip-direct on the GPU, through the C ABI, against the oracle.

FP64 path: bit-identical to the reference's FP64 targets form
(cpp/direct/direct_solver.cpp:59-83), per force evaluation and over KDK
steps (cpp/core/integrator.cpp:33-42).
Mixed path: max relative force error <= 1e-5 (the bound the mixed-precision
metric is quoted at) against the FP64 oracle.
Diagnostics: <= 1e-12 relative against the oracle's FP64 restatement.
"""
import numpy as np
import pytest

import oracle as O
from conftest import load_diagnostics, load_snapshot
from helpers import pair_magnitude_sums, rel_err, smoke_cube

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-5  # max relative force error, mixed path vs FP64


def _engine(gpu, precision, eps, **kw):
    return gpu.Engine("hip-direct", precision, softening=eps, **kw)


def _accel(gpu, precision, pos, vel, mass, eps, G=1.0, shape=None, sources=None, pairs=None):
    with _engine(gpu, precision, eps, G=G) as eng:
        if shape:
            eng.set_launch_shape(*shape)
        if sources:
            eng.set_sources(sources)
            assert eng.sources() == sources
        if pairs:
            eng.set_pairs(pairs)
        eng.upload(pos, vel, mass)
        eng.compute_accelerations()
        return eng.download()[2]


@pytest.mark.parametrize("n", [250, 1000])
def test_fp64_bit_exact_on_golden_inputs(gpu, n):
    m, x, v, a = load_snapshot(f"direct_{n}_snapshot_000000.csv.gz")
    got = _accel(gpu, "fp64", x, v, m, 0.025)
    assert np.array_equal(got, O.direct_targets(x, m, 1.0, 0.025))
    # the golden itself is the symmetric form: same sums, other order
    assert rel_err(got, a).max() < 1e-13


@pytest.mark.parametrize("n,eps,G", [(1, 0.02, 1.0), (2, 0.0, 1.0), (257, 0.01, 2.5),
                                     (3001, 0.05, 1.0)])
def test_fp64_bit_exact_random(gpu, n, eps, G):
    rng = np.random.default_rng(n)
    pos, m = rng.normal(size=(n, 3)), rng.uniform(0.1, 1.0, n)
    got = _accel(gpu, "fp64", pos, None, m, eps, G)
    assert np.array_equal(got, O.direct_targets(pos, m, G, eps))


def test_fp64_coincident_particles_without_softening(gpu):
    pos = np.array([[0.0, 0, 0], [0.0, 0, 0], [1.0, 0, 0], [0.5, 0.5, 0.5]])
    m = np.array([1.0, 2.0, 3.0, 4.0])
    got = _accel(gpu, "fp64", pos, None, m, 0.0)
    assert np.isfinite(got).all()
    assert np.array_equal(got, O.direct_targets(pos, m, 1.0, 0.0))
    mixed = _accel(gpu, "mixed", pos, None, m, 0.0)
    assert np.isfinite(mixed).all() and rel_err(mixed, got).max() < MIXED_TOL


def test_fp64_kdk_steps_bit_exact(gpu):
    m, x, v, a = load_snapshot("direct_1000_snapshot_000000.csv.gz")
    solver = O.make_solver(O.SOLVER_DIRECT_TARGETS, 1.0, 0.025)
    a0 = O.direct_targets(x, m, 1.0, 0.025)
    want = O.leapfrog(solver, x, v, a0, m, 0.01, 20)
    with _engine(gpu, "fp64", 0.025) as eng:
        eng.upload(x, v, m)
        eng.compute_accelerations()
        eng.step(0.01, 20)
        got = eng.download()
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
    # and the reference's golden (symmetric-form sums) to rounding
    m20, x20, v20, a20 = load_snapshot("direct_1000_snapshot_000020.csv.gz")
    assert np.abs(got[0] - x20).max() < 1e-12
    assert rel_err(got[2], a20).max() < 1e-10


def test_fp64_step_from_caller_accelerations(gpu):
    # cuda_direct_leapfrog_step semantics: the step starts from the
    # accelerations the particles carry, not from a fresh a(x)
    m, x, v, a = load_snapshot("direct_250_snapshot_000000.csv.gz")
    fake = a * 0.5
    solver = O.make_solver(O.SOLVER_DIRECT_TARGETS, 1.0, 0.025)
    want = O.leapfrog(solver, x, v, fake, m, 0.01, 1)
    with _engine(gpu, "fp64", 0.025) as eng:
        eng.upload(x, v, m, acc=fake)
        eng.step(0.01, 1)
        got = eng.download()
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_targets_range_leaves_others(gpu):
    rng = np.random.default_rng(5)
    pos, m = rng.normal(size=(900, 3)), rng.uniform(0.1, 1.0, 900)
    for prec in ("fp64", "mixed"):
        with _engine(gpu, prec, 0.02) as eng:
            eng.upload(pos, None, m)
            eng.compute_accelerations(300, 650)
            acc = eng.download()[2]
        ref = O.direct_targets(pos, m, 1.0, 0.02)
        assert not acc[:300].any() and not acc[650:].any()
        tol = 0 if prec == "fp64" else MIXED_TOL
        assert rel_err(acc[300:650], ref[300:650]).max() <= tol


def test_reference_smoke_cube_thresholds(gpu):
    # cpp/tests/smoke_tests.cpp:83-123: the device solver matches direct < 1e-10
    pos, m = smoke_cube()
    ref = O.direct_symmetric(pos, m, 1.0, 0.03)
    assert rel_err(_accel(gpu, "fp64", pos, None, m, 0.03), ref).mean() < 1e-10
    assert rel_err(_accel(gpu, "mixed", pos, None, m, 0.03), ref).max() < MIXED_TOL


def test_mixed_config0_error(gpu):
    pos, vel, m = gpu.generators.config0()
    ref = O.direct_targets(pos, m, 1.0, 0.05)
    err = rel_err(_accel(gpu, "mixed", pos, vel, m, 0.05), ref)
    assert err.max() <= MIXED_TOL and np.median(err) < 1e-6


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (4, 8), (1, 64), (2, 0), (0, 0)])
def test_mixed_launch_shapes(gpu, shape):
    pos, vel, m = gpu.generators.config1(n=6000)
    ref = O.direct_targets(pos, m, 1.0, 0.02, threads=0)
    got = {src: _accel(gpu, "mixed", pos, vel, m, 0.02, shape=shape, sources=src)
           for src in ("lds", "scalar")}
    assert rel_err(got["lds"], ref).max() <= MIXED_TOL
    # the two source paths run the same arithmetic in the same order
    assert np.array_equal(got["lds"], got["scalar"])


@pytest.mark.parametrize("sources", ["lds", "scalar"])
def test_mixed_sources_edge_cases(gpu, sources):
    rng = np.random.default_rng(17)
    pos, m = rng.normal(size=(1100, 3)), rng.uniform(0.1, 1.0, 1100)
    ref = O.direct_targets(pos, m, 1.0, 0.02)
    for lo, hi in ((0, 1100), (256, 768), (300, 650), (1099, 1100)):
        with _engine(gpu, "mixed", 0.02) as eng:
            eng.set_sources(sources)
            eng.upload(pos, None, m)
            eng.compute_accelerations(lo, hi)
            acc = eng.download()[2]
        assert not acc[:lo].any() and not acc[hi:].any()
        assert rel_err(acc[lo:hi], ref[lo:hi]).max() <= MIXED_TOL
    # coincident particles without softening: finite, the pair adds nothing
    pos2, m2 = np.concatenate([pos[:300], pos[:1]]), m[:301]
    got = _accel(gpu, "mixed", pos2, None, m2, 0.0, sources=sources)
    assert np.isfinite(got).all()
    assert rel_err(got, O.direct_targets(pos2, m2, 1.0, 0.0)).max() <= MIXED_TOL
    assert not _accel(gpu, "mixed", pos[:1], None, m[:1], 0.02, sources=sources).any()


@pytest.mark.parametrize("ipt", [1, 2, 4])
@pytest.mark.parametrize("n", [1, 2, 255, 256, 257, 700, 1000, 3001])
def test_mixed_symmetric_pairs_random(gpu, n, ipt):
    # each unordered pair once (the reference's symmetric form,
    # cpp/direct/direct_solver.cpp:37-57), against the FP64 targets form
    rng = np.random.default_rng(100 + n)
    pos, m = rng.normal(size=(n, 3)), rng.uniform(0.1, 1.0, n)
    ref = O.direct_targets(pos, m, 1.5, 0.02)
    with _engine(gpu, "mixed", 0.02, G=1.5) as eng:
        eng.set_pairs("symmetric")
        eng.set_launch_shape(ipt, 0)
        eng.upload(pos, None, m)
        assert eng.pairs() == "symmetric"
        eng.compute_accelerations()
        got = eng.download()[2]
    # a random cloud at eps = 0.02 has particles whose force cancels to a
    # small fraction of its terms: the bound is on the FP32 rounding of the
    # terms (64 ulp of sum |term|), and 1e-6 relative in median
    err = np.linalg.norm(got - ref, axis=1)
    assert (err <= 64 * 2.0**-24 * pair_magnitude_sums(pos, m, 1.5, 0.02) + 1e-300).all()
    assert np.median(rel_err(got, ref)) < 1e-6
    # momentum: the pair forces cancel, sum m a = 0 up to FP32 rounding
    scale = np.abs(m[:, None] * ref).sum(axis=0).max()
    assert np.abs((m[:, None] * got).sum(axis=0)).max() <= 1e-6 * max(scale, 1e-300)


def test_mixed_symmetric_pairs_two_disks(gpu):
    pos, vel, m = gpu.generators.config1(n=6000)
    ref = O.direct_targets(pos, m, 1.0, 0.02, threads=0)
    sym = _accel(gpu, "mixed", pos, vel, m, 0.02, pairs="symmetric")
    tgt = _accel(gpu, "mixed", pos, vel, m, 0.02, pairs="targets")
    for got in (sym, tgt):
        err = rel_err(got, ref)
        assert err.max() <= MIXED_TOL and np.median(err) < 1e-6
    # deterministic: a second evaluation gives the same bits
    assert np.array_equal(sym, _accel(gpu, "mixed", pos, vel, m, 0.02, pairs="symmetric"))


def test_mixed_symmetric_edge_cases(gpu):
    rng = np.random.default_rng(23)
    pos, m = rng.normal(size=(700, 3)), rng.uniform(0.1, 1.0, 700)
    # target ranges run the per-target form and leave the others untouched
    with _engine(gpu, "mixed", 0.02) as eng:
        eng.set_pairs("symmetric")
        eng.upload(pos, None, m)
        eng.compute_accelerations(100, 600)
        acc = eng.download()[2]
    ref = O.direct_targets(pos, m, 1.0, 0.02)
    assert not acc[:100].any() and not acc[600:].any()
    assert rel_err(acc[100:600], ref[100:600]).max() <= MIXED_TOL
    # coincident particles in different tiles without softening
    pos2, m2 = np.concatenate([pos[:300], pos[:1]]), m[:301]
    got = _accel(gpu, "mixed", pos2, None, m2, 0.0, pairs="symmetric")
    assert np.isfinite(got).all()
    assert rel_err(got, O.direct_targets(pos2, m2, 1.0, 0.0)).max() <= MIXED_TOL
    # re-upload with fewer particles reuses the context: stale state past n
    # must not leak into the sums
    with _engine(gpu, "mixed", 0.02) as eng:
        eng.set_pairs("symmetric")
        eng.upload(pos, None, m)
        eng.compute_accelerations()
        eng.upload(pos[:333], None, m[:333])
        eng.compute_accelerations()
        got = eng.download()[2]
    assert rel_err(got, O.direct_targets(pos[:333], m[:333], 1.0, 0.02)).max() <= MIXED_TOL


def _equal_mass_cloud(n, seed):
    """A random cloud whose 256-particle tiles mostly hold equal masses:
    tiles 0-1 one mass, tile 2 another, tile 3 equal masses but one particle
    1 ulp heavier, the rest a third mass; a partial last tile when n is not
    a multiple of 256."""
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3))
    m = np.full(n, 0.37)
    m[512:768] = 2.9e-3
    m[768:1024] = 0.11
    m[800] = np.nextafter(np.float32(0.11), np.float32(1.0))
    m[1024:] = 1.3e-5
    return pos, m


@pytest.mark.parametrize("n", [2048, 3001])
def test_mixed_symmetric_equal_mass_tiles(gpu, n):
    # tile pairs whose particles share one mass per tile are evaluated in
    # coordinates scaled by (G m_J)^-1/2 with no per-pair mass product; tiles
    # of unequal masses, partial tiles and the diagonal tiles take the
    # general pair step.  Both are held to the FP32 rounding bound of the
    # random-cloud test against the FP64 oracle.
    pos, m = _equal_mass_cloud(n, 31 + n)
    for G, eps in ((1.0, 0.02), (0.7, 0.3)):
        ref = O.direct_targets(pos, m, G, eps)
        got = _accel(gpu, "mixed", pos, None, m, eps, G=G, pairs="symmetric")
        err = np.linalg.norm(got - ref, axis=1)
        assert (err <= 64 * 2.0**-24 * pair_magnitude_sums(pos, m, G, eps)).all(), (G, eps)
        assert np.median(rel_err(got, ref)) < 1e-6
        scale = np.abs(m[:, None] * ref).sum(axis=0).max()
        assert np.abs((m[:, None] * got).sum(axis=0)).max() <= 1e-6 * scale
        # the per-target form (no tile-mass factoring) agrees as closely
        tgt = _accel(gpu, "mixed", pos, None, m, eps, G=G, pairs="targets")
        assert (np.linalg.norm(got - tgt, axis=1)
                <= 128 * 2.0**-24 * pair_magnitude_sums(pos, m, G, eps)).all()


@pytest.mark.parametrize("gm, eps", [(1e-13, 0.02), (2e12, 0.02), (1.0, 1e-11), (4e-6, 1e-13)])
def test_mixed_symmetric_equal_mass_outside_scaled_range(gpu, gm, eps):
    # G m outside [2^-40, 2^40], or eps^2 / (G m) below 1e-20: the scaled
    # form is not used and the general step gives the oracle's forces
    rng = np.random.default_rng(5)
    n = 1024
    pos, m = rng.normal(size=(n, 3)), np.full(n, gm)
    ref = O.direct_targets(pos, m, 1.0, eps)
    got = _accel(gpu, "mixed", pos, None, m, eps, pairs="symmetric")
    err = np.linalg.norm(got - ref, axis=1)
    assert np.isfinite(got).all()
    assert (err <= 64 * 2.0**-24 * pair_magnitude_sums(pos, m, 1.0, eps)).all()


def test_mixed_symmetric_kdk_tracks_fp64(gpu):
    pos, vel, m = gpu.generators.config0()
    with _engine(gpu, "mixed", 0.05) as mx, _engine(gpu, "fp64", 0.05) as fp:
        mx.set_pairs("symmetric")
        for eng in (mx, fp):
            eng.upload(pos, vel, m)
            eng.compute_accelerations()
            eng.step(1e-3, 50)
        pm = mx.download()[0]
        pf = fp.download()[0]
        e_m, e_f = mx.diagnostics(), fp.diagnostics()
    assert np.abs(pm - pf).max() < 1e-6
    assert abs(e_m["total_energy"] - e_f["total_energy"]) < 1e-7 * abs(e_f["total_energy"])


def test_mixed_full_size_against_fp64_gpu(gpu):
    # N = 65,536 (the precision workload): the FP64 GPU path is the oracle
    # here (bit-identical to the CPU FP64 form above); spot-check it on a
    # CPU sample too.
    pos, vel, m = gpu.generators.config1()
    a64 = _accel(gpu, "fp64", pos, vel, m, 0.02)
    sample = np.arange(0, 65536, 257)
    cpu = O.direct_targets(pos, m, 1.0, 0.02, threads=0)
    assert np.array_equal(a64[sample], cpu[sample])
    for pairs in ("targets", "symmetric"):
        err = rel_err(_accel(gpu, "mixed", pos, vel, m, 0.02, pairs=pairs), a64)
        assert err.max() <= MIXED_TOL, pairs
        assert np.median(err) < 1e-6, pairs


def test_mixed_kdk_tracks_fp64(gpu):
    pos, vel, m = gpu.generators.config0()
    with _engine(gpu, "mixed", 0.05) as mx, _engine(gpu, "fp64", 0.05) as fp:
        for eng in (mx, fp):
            eng.upload(pos, vel, m)
            eng.compute_accelerations()
            eng.step(1e-3, 50)
        pm, vm, am = mx.download()
        pf, vf, af = fp.download()
        e_m, e_f = mx.diagnostics(), fp.diagnostics()
    assert np.abs(pm - pf).max() < 1e-6
    assert abs(e_m["total_energy"] - e_f["total_energy"]) < 1e-7 * abs(e_f["total_energy"])


@pytest.mark.parametrize("n", [250, 1000])
def test_diagnostics_against_oracle_and_golden(gpu, n):
    m, x, v, a = load_snapshot(f"direct_{n}_snapshot_000000.csv.gz")
    rows = load_diagnostics(f"direct_{n}_diagnostics.csv")
    with _engine(gpu, "fp64", 0.025) as eng:
        eng.upload(x, v, m)
        d = eng.diagnostics()
    want = O.diagnostics(x, v, m, 1.0, 0.025)
    for k, w in want.items():
        assert abs(d[k] - w) <= 1e-12 * max(abs(w), 1e-3), k
    got_row = np.array([d[k] for k in O.DIAG_FIELDS])
    assert np.allclose(got_row, rows[0, 3:], rtol=1e-12, atol=1e-15)


def _diag_scales(pos, vel, m, d):
    """Magnitude of the summed terms per diagnostic: sums with cancellation
    (momentum, centre of mass, angular momentum) are compared relative to
    it, since summation order alone moves them by ~1e-16 of it."""
    M = m.sum()
    mom = np.sum(m * np.abs(vel).sum(1))
    com = np.sum(m * np.abs(pos).sum(1)) / M
    ang = np.sum(m * np.linalg.norm(pos, axis=1) * np.linalg.norm(vel, axis=1))
    out = {}
    for k in d:
        if k.startswith("momentum"):
            out[k] = mom
        elif k.startswith("center_of_mass"):
            out[k] = com
        elif k.startswith("angular"):
            out[k] = ang
        else:
            out[k] = abs(d["potential_energy"]) if "energy" in k else abs(d[k])
    return out


def test_diagnostics_large_n(gpu):
    pos, vel, m = gpu.generators.config1(n=20000)
    with _engine(gpu, "fp64", 0.02) as eng:
        eng.upload(pos, vel, m)
        d = eng.diagnostics()
    want = O.diagnostics(pos, vel, m, 1.0, 0.02)
    scale = _diag_scales(pos, vel, m, want)
    for k, w in want.items():
        assert abs(d[k] - w) <= 1e-12 * max(abs(w), scale[k]), k


@pytest.mark.parametrize("case", ["coincident_eps0", "subnormal_r2", "overflow_r2", "ragged"])
def test_diagnostics_edge_cases(gpu, case):
    # the potential's 1/d (refined hardware reciprocal square root) at the
    # ends of the FP64 range, against the oracle's sqrt-and-divide
    rng = np.random.default_rng(3)
    n, eps = 300, 0.0
    pos, vel, m = rng.normal(size=(n, 3)), rng.normal(size=(n, 3)) * 0.1, rng.uniform(0.5, 1, n)
    if case == "coincident_eps0":
        pos[7] = pos[200]  # d = 0: skipped
        pos[8] = pos[9]
    elif case == "subnormal_r2":
        pos[200] = 0.0
        pos[7] = [1e-160, 0.0, 0.0]  # r^2 = 1e-320, subnormal; d = 1e-160
    elif case == "overflow_r2":
        pos[7] = [1e200, 0.0, 0.0]  # r^2 = inf: those pairs contribute 0
    else:
        n, eps = 257, 0.01
        pos, vel, m = pos[:n], vel[:n], m[:n]
    with _engine(gpu, "fp64", eps) as eng:
        eng.upload(pos, vel, m)
        d = eng.diagnostics()
    want = O.diagnostics(pos, vel, m, 1.0, eps)
    assert np.isfinite(want["potential_energy"]) and np.isfinite(d["potential_energy"])
    for k in ("potential_energy", "kinetic_energy", "total_energy"):
        assert abs(d[k] - want[k]) <= 1e-12 * abs(want[k]), (k, d[k], want[k])


def test_empty_and_reupload(gpu):
    with _engine(gpu, "mixed", 0.02) as eng:
        eng.upload(np.zeros((0, 3)), None, np.zeros(0))
        eng.compute_accelerations()
        eng.step(0.01, 2)
        assert eng.n == 0
        rng = np.random.default_rng(9)
        for n in (700, 100, 1500):
            pos, m = rng.normal(size=(n, 3)), rng.uniform(0.1, 1, n)
            eng.upload(pos, None, m)
            eng.compute_accelerations()
            acc = eng.download()[2]
            assert acc.shape == (n, 3)
            assert rel_err(acc, O.direct_targets(pos, m, 1.0, 0.02)).max() <= MIXED_TOL


def test_cpp_interface(gpu):
    import subprocess
    from pathlib import Path

    exe = Path(__file__).resolve().parent / "cpp" / "hip_solver_test"
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "hip_solver_test passed" in res.stdout and "no device" not in res.stdout


@pytest.mark.parametrize("n", [65536, 1000, 257])
def test_multi_step_fusion_is_bit_identical(gpu, n):
    # gc_step over several steps fuses each merge with the next opening kick,
    # drift and source packing; one step at a time runs them as separate
    # kernels.  The trajectories must agree bit for bit.
    pos, vel, m = gpu.generators.config1(n=65536)
    pos, vel, m = pos[:n], vel[:n], m[:n]
    out = []
    for fused in (True, False):
        with _engine(gpu, "mixed", 0.02) as eng:
            eng.upload(pos, vel, m)
            eng.compute_accelerations()
            if fused:
                eng.step(5e-4, 6)
            else:
                for _ in range(6):
                    eng.step(5e-4, 1)
            out.append(eng.download())
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision,pairs", [("mixed", "symmetric"), ("mixed", "targets"),
                                             ("fp64", None), ("fp64-fast", None)])
def test_graph_chunks_are_bit_identical(gpu, precision, pairs):
    # gc_step replays whole chunks of graph_steps steps from a captured
    # hipGraph (on the context's own stream and on the null stream).  Every
    # replay, with eager work (downloads, diagnostics, a one-step call) and
    # synchronisations between replays, and a recapture after a change of
    # dt, must repeat the eagerly launched trajectory bit for bit.
    pos, vel, m = gpu.generators.config1(n=65536)
    pos, vel, m = pos[:3000], vel[:3000], m[:3000]
    plan = [(5e-4, 25), (5e-4, 10), (5e-4, 1), (5e-4, 11), (2.5e-4, 20), (5e-4, 30)]
    runs = []
    for chunk, null_stream in ((0, False), (10, False), (7, False), (10, True)):
        with _engine(gpu, precision, 0.02) as eng:
            if null_stream:
                eng.set_stream(0)
            if pairs:
                eng.set_pairs(pairs)
            eng.set_graph_steps(chunk)
            assert eng.graph_steps() == chunk
            eng.upload(pos, vel, m)
            eng.compute_accelerations()
            states = []
            for dt, k in plan:
                eng.step(dt, k)
                states.append(eng.download())
                eng.diagnostics()
            runs.append(states)
    for got in runs[1:]:
        for a_state, b_state in zip(runs[0], got):
            for a, b in zip(a_state, b_state):
                assert np.array_equal(a, b)


@pytest.mark.parametrize("n,form", [(262144, "symmetric"), (262145, "targets")])
def test_mixed_at_the_symmetric_size_limit(gpu, n, form):
    # The symmetric form covers up to 1,024 tiles (262,144 particles); one
    # particle more runs the per-target form.  Both against the FP64
    # throughput path (<= 1e-13 of the oracle) on two-disk states.
    pos, vel, m = gpu.generators.two_disk((n + 1) // 2, softening=0.02, seed=5)
    pos, vel, m = pos[:n], vel[:n], m[:n]
    with _engine(gpu, "mixed", 0.02) as eng:
        eng.upload(pos, vel, m)
        assert eng.pairs() == form
        eng.compute_accelerations()
        got = eng.download()[2]
    with gpu.Engine("hip-direct", "fp64-fast", softening=0.02) as e64:
        e64.upload(pos, vel, m)
        e64.compute_accelerations()
        want = e64.download()[2]
    err = rel_err(got, want)
    print(f"n={n} {form}: max rel err {err.max():.2e}", flush=True)
    assert err.max() <= MIXED_TOL
    assert np.median(err) < 1e-6


def test_contexts_on_threads_step_concurrently(gpu):
    # One context per host thread: several threads stepping their own
    # contexts at once (each capturing and replaying its step graphs, one of
    # them on the FMM solver with its second stream) give the results of the
    # same runs stepped one after another.
    import threading

    pos, vel, m = gpu.generators.config1(n=65536)
    pos, vel, m = pos[:3000], vel[:3000], m[:3000]
    cases = [("hip-direct", "mixed"), ("hip-direct", "fp64-fast"), ("hip-direct", "mixed"),
             ("hip-fmm", "mixed")]

    def run(solver, precision, out, k):
        with gpu.Engine(solver, precision, softening=0.02) as eng:
            eng.upload(pos, vel, m)
            eng.compute_accelerations()
            for _ in range(3):
                eng.step(5e-4, 25)
                eng.diagnostics()
            out[k] = eng.download()

    alone = {}
    for k, (solver, precision) in enumerate(cases):
        run(solver, precision, alone, k)
    together, errors = {}, []

    def guarded(*a):
        try:
            run(*a)
        except Exception as e:  # reported below, in the main thread
            errors.append(e)

    threads = [threading.Thread(target=guarded, args=(sv, pr, together, k))
               for k, (sv, pr) in enumerate(cases)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for k in range(len(cases)):
        for a, b in zip(alone[k], together[k]):
            assert np.array_equal(a, b), cases[k]


def test_graph_steps_limits(gpu):
    with _engine(gpu, "mixed", 0.02) as eng:
        assert eng.graph_steps() == 10
        with pytest.raises(gpu.GcError):
            eng.set_graph_steps(-1)
        with pytest.raises(gpu.GcError):
            eng.set_graph_steps(1001)
        eng.set_graph_steps(1000)
        assert eng.graph_steps() == 1000


# ---- FP64 throughput form (precision "fp64-fast") ----------------------------
FAST64_TOL = 1e-13  # max relative force difference from the oracle's targets form


def _fast64(gpu, pos, mass, eps, G=1.0, shape=None, begin=None, end=None):
    with gpu.Engine("hip-direct", "fp64-fast", softening=eps, G=G) as eng:
        if shape:
            eng.set_launch_shape(*shape)
        eng.upload(pos, None, mass)
        eng.compute_accelerations(begin, end)
        return eng.download()[2]


@pytest.mark.parametrize("n,eps,G", [(1, 0.02, 1.0), (2, 0.0, 1.0), (257, 0.01, 2.5),
                                     (3001, 0.05, 1.0)])
def test_fp64_fast_against_oracle(gpu, n, eps, G):
    rng = np.random.default_rng(n + 7)
    pos, m = rng.normal(size=(n, 3)), rng.uniform(0.1, 1.0, n)
    want = O.direct_targets(pos, m, G, eps)
    got = _fast64(gpu, pos, m, eps, G)
    assert rel_err(got, want).max() < FAST64_TOL


def test_fp64_fast_coincident_without_softening(gpu):
    # s2 == 0 pairs are skipped, as the reference's softened_acceleration does
    pos = np.array([[0.0, 0, 0], [0.0, 0, 0], [1.0, 0, 0], [0.5, 0.5, 0.5], [1.0, 0, 0]])
    m = np.array([1.0, 2.0, 3.0, 4.0, 0.5])
    got = _fast64(gpu, pos, m, 0.0)
    assert np.isfinite(got).all()
    assert rel_err(got, O.direct_targets(pos, m, 1.0, 0.0)).max() < FAST64_TOL


@pytest.mark.parametrize("shape", [(1, 1), (2, 7), (4, 64), (2, 0), (1, 200)])
def test_fp64_fast_launch_shapes_and_ranges(gpu, shape):
    pos, vel, m = gpu.generators.config0()
    want = O.direct_targets(pos, m, 1.0, 0.05)
    got = _fast64(gpu, pos, m, 0.05, shape=shape)
    assert rel_err(got, want).max() < FAST64_TOL
    assert np.array_equal(got, _fast64(gpu, pos, m, 0.05, shape=shape))  # deterministic
    part = _fast64(gpu, pos, m, 0.05, shape=shape, begin=123, end=777)
    assert not part[:123].any() and not part[777:].any()
    assert rel_err(part[123:777], want[123:777]).max() < FAST64_TOL


def test_fp64_fast_full_size_and_steps(gpu):
    pos, vel, m = gpu.generators.config1()
    with gpu.Engine("hip-direct", "fp64", softening=0.02) as e64, \
            gpu.Engine("hip-direct", "fp64-fast", softening=0.02) as ef:
        for e in (e64, ef):
            e.upload(pos, vel, m)
            e.compute_accelerations()
        a64, af = e64.download()[2], ef.download()[2]
        assert rel_err(af, a64).max() < FAST64_TOL
        for e in (e64, ef):
            e.step(5e-4, 20)
        (p64, v64, a64), (pf, vf, af) = e64.download(), ef.download()
    assert np.abs(pf - p64).max() < 1e-12 and np.abs(vf - v64).max() < 1e-11
    assert rel_err(af, a64).max() < 1e-12
