"""Run output from device-resident state (gc_writer_snapshot,
gc_writer_diagnostics) through the C ABI.

A snapshot of uploaded golden state must equal the reference's file byte
for byte; snapshots taken between queued KDK steps, without synchronising,
must hold the state at the point they were enqueued (the oracle's FP64
leapfrog, bit-identical to the hip-direct FP64 path); device-path files
must equal the host-path files of the downloaded state."""
import gzip

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, load_diagnostics, load_snapshot
from gcsim.writer import DIAG_KEYS

pytestmark = pytest.mark.gpu

EPS = 0.025


def _golden_bytes(name):
    with gzip.open(GOLDEN / name, "rb") as fh:
        return fh.read().replace(b"\r\n", b"\n")


def _groups(name):
    with gzip.open(GOLDEN / name, "rt") as fh:
        return np.loadtxt(fh, delimiter=",", skiprows=2, usecols=1).astype(np.uint32)


def _parse(path):
    lines = path.read_text().splitlines()
    return float(lines[0].split("=", 1)[1]), np.loadtxt(lines[2:], delimiter=",", ndmin=2)


@pytest.mark.parametrize("n", [250, 1000])
def test_device_snapshot_of_golden_state_is_byte_identical(gpu, tmp_path, n):
    m, x, v, a = load_snapshot(f"direct_{n}_snapshot_000000.csv.gz")
    with gpu.Engine("hip-direct", "fp64", softening=EPS) as eng, \
            gpu.SnapshotWriter(tmp_path) as out:
        eng.upload(x, v, m, acc=a)
        out.write_snapshot(0, 0.0, eng,
                           group_ids=_groups(f"direct_{n}_snapshot_000000.csv.gz"))
        out.flush()
    assert (tmp_path / "snapshot_000000.csv").read_bytes() == \
        _golden_bytes(f"direct_{n}_snapshot_000000.csv.gz")


def test_snapshots_between_queued_steps(gpu, tmp_path):
    m, x, v, a = load_snapshot("direct_500_snapshot_000000.csv.gz")
    groups = (np.arange(500) >= 250).astype(np.uint32)
    solver = O.make_solver(O.SOLVER_DIRECT_TARGETS, 1.0, EPS)
    a0 = O.direct_targets(x, m, 1.0, EPS)
    dt, nsnap = 0.01, 6  # more snapshots than staging buffers
    with gpu.Engine("hip-direct", "fp64", softening=EPS) as eng, \
            gpu.SnapshotWriter(tmp_path) as out:
        eng.upload(x, v, m)
        eng.compute_accelerations()
        for k in range(nsnap):
            out.write_snapshot(k, k * dt, eng, group_ids=groups)
            eng.step(dt, 1)  # queued behind the staging copy, no host wait
        out.flush()
        final = eng.download()
    want = (x, v, a0)
    for k in range(nsnap):
        time, d = _parse(tmp_path / f"snapshot_{k:06d}.csv")
        assert time == k * dt
        assert np.array_equal(d[:, 1], groups) and np.array_equal(d[:, 2], m)
        for got, w in zip((d[:, 3:6], d[:, 6:9], d[:, 9:12]), want):
            assert np.array_equal(got, w), k
        want = O.leapfrog(solver, *want, m, dt, 1)
    for got, w in zip(final, want):
        assert np.array_equal(got, w)


def test_device_and_host_paths_write_the_same_bytes(gpu, tmp_path):
    pos, vel, m = gpu.generators.config0()
    with gpu.Engine("hip-direct", "mixed", softening=0.05) as eng, \
            gpu.SnapshotWriter(tmp_path / "dev") as dev, \
            gpu.SnapshotWriter(tmp_path / "host") as host:
        eng.upload(pos, vel, m)
        eng.compute_accelerations()
        eng.step(1e-3, 3)
        dev.write_snapshot(3, 3e-3, eng)
        dev.write_diagnostics(3, 3e-3, eng)
        p, v, acc = eng.download()
        host.write_snapshot(3, 3e-3, p, v, acc, m)
        host.write_diagnostics(3, 3e-3, eng.diagnostics(), particle_count=len(m))
    for f in ("snapshot_000003.csv", "diagnostics.csv"):
        assert (tmp_path / "dev" / f).read_bytes() == (tmp_path / "host" / f).read_bytes(), f


def test_staging_regrows_for_a_larger_context(gpu, tmp_path):
    small = load_snapshot("direct_250_snapshot_000000.csv.gz")
    large = load_snapshot("fmm_1000_snapshot_000000.csv.gz")
    with gpu.SnapshotWriter(tmp_path) as out:
        for step, (m, x, v, a) in ((1, small), (2, large), (3, small)):
            with gpu.Engine("hip-direct", "fp64", softening=EPS) as eng:
                eng.upload(x, v, m, acc=a)
                out.write_snapshot(step, 0.0, eng)
    for step, (m, x, v, a) in ((1, small), (2, large), (3, small)):
        _, d = _parse(tmp_path / f"snapshot_{step:06d}.csv")
        assert np.array_equal(d[:, 3:6], x) and np.array_equal(d[:, 9:12], a)


@pytest.mark.parametrize("n", [250, 1000])
def test_device_diagnostics_rows(gpu, tmp_path, n):
    m, x, v, a = load_snapshot(f"direct_{n}_snapshot_000000.csv.gz")
    rows = load_diagnostics(f"direct_{n}_diagnostics.csv")
    with gpu.Engine("hip-direct", "fp64", softening=EPS) as eng, \
            gpu.SnapshotWriter(tmp_path) as out:
        eng.upload(x, v, m)
        out.write_diagnostics(0, 0.0, eng)
        d = eng.diagnostics()
    got = np.loadtxt(tmp_path / "diagnostics.csv", delimiter=",", skiprows=1, ndmin=2)
    assert got.shape == (1, 16) and got[0, 2] == n
    assert np.array_equal(got[0, 3:], [d[k] for k in DIAG_KEYS])
    assert np.allclose(got[0, 3:], rows[0, 3:], rtol=1e-12, atol=1e-15)


def test_cpp_mirror_device_snapshot(gpu, tmp_path):
    """HipSnapshotWriter::write_snapshot of a HipForceSolver's resident
    copy of golden state equals the reference's file."""
    import subprocess
    from pathlib import Path

    exe = Path(__file__).resolve().parent / "cpp" / "writer_check"
    src = tmp_path / "in.csv"
    want = _golden_bytes("direct_250_snapshot_000020.csv.gz")
    src.write_bytes(want)
    out = tmp_path / "out"
    print("writer_check direct_250 step 20", flush=True)
    res = subprocess.run([str(exe), str(src), "20", str(out)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "device snapshot written" in res.stdout
    assert (out / "snapshot_000020.csv").read_bytes() == want
    assert (out / "snapshot_000021.csv").read_bytes() == want
    rows = load_diagnostics("direct_250_diagnostics.csv")
    got = np.loadtxt(out / "diagnostics.csv", delimiter=",", skiprows=1, ndmin=2)
    assert got.shape == (2, 16) and got[1, 0] == 21
    assert np.allclose(got[1, 3:], rows[1, 3:], rtol=1e-12, atol=1e-15)


def test_run_loop_reproduces_reference_run_directory(gpu, tmp_path):
    """The direct_250 benchmark run (20 KDK steps, dt 0.01, snapshot_every
    20) through run_with_output: the same files as the reference's run
    directory; metadata and the step-0 snapshot byte for byte, step 20 equal
    to the oracle's FP64 leapfrog bit for bit and to the reference's
    (symmetric-form sums) to rounding."""
    from gcsim.writer import run_with_output

    m, x, v, a = load_snapshot("direct_250_snapshot_000000.csv.gz")
    groups = _groups("direct_250_snapshot_000000.csv.gz")
    meta = dict(name="benchmark_direct_250", solver="direct", particle_count=250, steps=20,
                dt=0.01, snapshot_every=20, dim=3, gravitational_constant=1,
                softening=EPS, tree_theta=0.58, tree_leaf_capacity=16, fmm_expansion_order=4)
    with gpu.Engine("hip-direct", "fp64", softening=EPS) as eng, \
            gpu.SnapshotWriter(tmp_path) as out:
        out.write_metadata(**meta)
        eng.upload(x, v, m, acc=a)
        run_with_output(eng, out, 20, 0.01, 20, group_ids=groups)
    assert sorted(p.name for p in tmp_path.iterdir()) == [
        "diagnostics.csv", "metadata.json", "snapshot_000000.csv", "snapshot_000020.csv"]
    with open(GOLDEN / "direct_250_metadata.json", "rb") as fh:
        assert (tmp_path / "metadata.json").read_bytes() == fh.read().replace(b"\r\n", b"\n")
    assert (tmp_path / "snapshot_000000.csv").read_bytes() == \
        _golden_bytes("direct_250_snapshot_000000.csv.gz")
    time, d = _parse(tmp_path / "snapshot_000020.csv")
    assert time == 20 * 0.01
    solver = O.make_solver(O.SOLVER_DIRECT_TARGETS, 1.0, EPS)
    want = O.leapfrog(solver, x, v, a, m, 0.01, 20)
    for got, w in zip((d[:, 3:6], d[:, 6:9], d[:, 9:12]), want):
        assert np.array_equal(got, w)
    m20, x20, v20, a20 = load_snapshot("direct_250_snapshot_000020.csv.gz")
    assert np.abs(d[:, 3:6] - x20).max() < 1e-12
    rows = load_diagnostics("direct_250_diagnostics.csv")
    got = np.loadtxt(tmp_path / "diagnostics.csv", delimiter=",", skiprows=1, ndmin=2)
    assert np.array_equal(got[:, :3], rows[:, :3])
    assert np.allclose(got[:, 3:], rows[:, 3:], rtol=1e-11, atol=1e-14)


def test_queued_diagnostics_rows_keep_order(gpu, tmp_path):
    # more rows than staging slots (16), each queued between KDK steps with
    # no host wait, interleaved with host-value rows and snapshots; every
    # row equals the synchronous diagnostics of the same state
    m, x, v, a = load_snapshot("direct_250_snapshot_000000.csv.gz")
    nrows, dt = 40, 0.01
    with gpu.Engine("hip-direct", "fp64", softening=EPS) as eng, \
            gpu.SnapshotWriter(tmp_path) as out:
        eng.upload(x, v, m, acc=a)
        for k in range(nrows):
            if k % 10 == 5:
                out.write_diagnostics(-k, 0.5, dict.fromkeys(DIAG_KEYS, float(k)),
                                      particle_count=7)
                out.write_snapshot(k, k * dt, eng)
            out.write_diagnostics(k, k * dt, eng)
            eng.step(dt, 1)
    want = []
    with gpu.Engine("hip-direct", "fp64", softening=EPS) as eng:
        eng.upload(x, v, m, acc=a)
        for k in range(nrows):
            if k % 10 == 5:
                want.append([-k, 0.5, 7] + [float(k)] * 13)
            d = eng.diagnostics()
            want.append([k, k * dt, 250] + [d[key] for key in DIAG_KEYS])
            eng.step(dt, 1)
    got = np.loadtxt(tmp_path / "diagnostics.csv", delimiter=",", skiprows=1, ndmin=2)
    assert got.shape == (len(want), 16)
    assert np.array_equal(got, np.array(want))
    assert len(list(tmp_path.glob("snapshot_*.csv"))) == 4
