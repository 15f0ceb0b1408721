"""Run-output writer against the reference's own output files (CPU).

The goldens under tests/golden/ref_outputs are files the reference's
SnapshotWriter wrote (cpp/io/snapshot_writer.cpp): every snapshot,
diagnostics and metadata file is parsed, handed to the writer's host entry
points, and the file it writes must equal the golden byte for byte (the
goldens carry CRLF line ends from their checkout; the reference writes LF).
17 significant digits round-trip every FP64 value, so any difference in
number formatting, column order or header text fails here."""
import gzip
import json
import re

import numpy as np
import pytest

from conftest import GOLDEN
from gcsim import GcError, SnapshotWriter
from gcsim.writer import DIAG_KEYS

SNAPSHOTS = sorted(p.name for p in GOLDEN.glob("*_snapshot_*.csv.gz"))
RUNS = sorted({p.name.split("_diagnostics")[0] for p in GOLDEN.glob("*_diagnostics.csv*")})


def _text(path):
    opener = gzip.open if path.suffix == ".gz" else open
    with opener(path, "rb") as fh:
        return fh.read().replace(b"\r\n", b"\n")


def _parse_snapshot(raw):
    lines = raw.decode().splitlines()
    time = float(lines[0].split("=", 1)[1])
    data = np.loadtxt(lines[2:], delimiter=",", ndmin=2)
    return time, data


def test_golden_inventory():
    assert len(SNAPSHOTS) == 14 and len(RUNS) == 6


@pytest.mark.parametrize("name", SNAPSHOTS)
@pytest.mark.parametrize("threads", [1, 3])
def test_snapshot_bytes_match_reference(tmp_path, name, threads):
    raw = _text(GOLDEN / name)
    time, d = _parse_snapshot(raw)
    step = int(re.search(r"snapshot_(\d+)", name).group(1))
    assert np.array_equal(d[:, 0], np.arange(d.shape[0]))
    with SnapshotWriter(tmp_path) as out:
        out.set_threads(threads)
        out.write_snapshot(step, time, d[:, 3:6], d[:, 6:9], d[:, 9:12], d[:, 2],
                           group_ids=d[:, 1].astype(np.uint32))
    got = (tmp_path / f"snapshot_{step:06d}.csv").read_bytes()
    assert got == raw


@pytest.mark.parametrize("run", RUNS)
def test_diagnostics_bytes_match_reference(tmp_path, run):
    path = next(GOLDEN.glob(f"{run}_diagnostics.csv*"))
    raw = _text(path)
    lines = raw.decode().splitlines()
    assert lines[0].split(",") == ["step", "time", "n", *DIAG_KEYS]
    with SnapshotWriter(tmp_path) as out:
        for line in lines[1:]:
            f = line.split(",")
            vals = dict(zip(DIAG_KEYS, map(float, f[3:])))
            out.write_diagnostics(int(f[0]), float(f[1]), vals, particle_count=int(f[2]))
    assert (tmp_path / "diagnostics.csv").read_bytes() == raw


@pytest.mark.parametrize("run", RUNS)
def test_metadata_bytes_match_reference(tmp_path, run):
    raw = _text(GOLDEN / f"{run}_metadata.json")
    meta = json.loads(raw)
    with SnapshotWriter(tmp_path) as out:
        out.write_metadata(**meta)
    assert (tmp_path / "metadata.json").read_bytes() == raw


def test_empty_run_and_wide_step(tmp_path):
    # setw(6) is a minimum width: step 1234567 keeps all digits
    with SnapshotWriter(tmp_path / "nested" / "run") as out:
        z = np.zeros((0, 3))
        out.write_snapshot(1234567, 0.5, z, z, z, np.zeros(0))
    d = tmp_path / "nested" / "run"
    assert (d / "snapshot_1234567.csv").read_bytes() == (
        b"# time=0.5\nid,group_id,mass,x,y,z,vx,vy,vz,ax,ay,az\n")
    assert (d / "diagnostics.csv").read_bytes().count(b"\n") == 1


def test_metadata_escapes_quotes_and_backslashes(tmp_path):
    with SnapshotWriter(tmp_path) as out:
        out.write_metadata(name='a"b\\c', solver="hip-direct", particle_count=2, steps=1,
                           dt=0.5, snapshot_every=1)
    meta = json.loads((tmp_path / "metadata.json").read_bytes())
    assert meta["name"] == 'a"b\\c' and meta["solver"] == "hip-direct"
    # SimulationConfig defaults (cpp/core/config.hpp:19-34)
    assert (meta["dim"], meta["tree_theta"], meta["tree_leaf_capacity"],
            meta["fmm_expansion_order"], meta["gravitational_constant"],
            meta["softening"]) == (3, 0.6, 16, 4, 1, 0.02)


def test_snapshot_values_round_trip(tmp_path):
    # special values and extremes of FP64 survive the text form
    rng = np.random.default_rng(5)
    pos = rng.standard_normal((7, 3)) * 10.0 ** rng.integers(-300, 300, (7, 3))
    pos[0] = [0.0, -0.0, 5e-324]
    pos[1] = [np.finfo(float).max, -np.finfo(float).tiny, 1.0 / 3.0]
    vel, acc, mass = pos[::-1].copy(), -pos, np.abs(pos[:, 0]) + 1
    groups = np.array([0, 1, 2, 3, 4294967295, 7, 0], dtype=np.uint32)
    with SnapshotWriter(tmp_path) as out:
        out.write_snapshot(3, 1.0 / 3.0, pos, vel, acc, mass, group_ids=groups)
    time, d = _parse_snapshot((tmp_path / "snapshot_000003.csv").read_bytes())
    assert time == 1.0 / 3.0
    for got, want in ((d[:, 2], mass), (d[:, 3:6], pos), (d[:, 6:9], vel), (d[:, 9:12], acc)):
        assert np.array_equal(got, want)
    assert np.array_equal(np.signbit(d[:, 3:6]), np.signbit(pos))
    assert np.array_equal(d[:, 1].astype(np.uint64), groups.astype(np.uint64))


def test_many_chunks_in_row_order(tmp_path):
    # rows span several formatting chunks (8192 rows each) and threads; the
    # text must equal printf's %.17g of every value, in row order
    n = 3 * 8192 + 17
    rng = np.random.default_rng(11)
    bits = rng.integers(0, 2**63, (n, 10), dtype=np.int64).view(np.float64)
    bits[:, 0] = np.abs(rng.standard_normal(n))  # finite masses
    groups = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    for threads in (1, 2, 5):
        d = tmp_path / str(threads)
        with SnapshotWriter(d) as out:
            out.set_threads(threads)
            out.write_snapshot(7, 0.125, bits[:, 1:4], bits[:, 4:7], bits[:, 7:10], bits[:, 0],
                               group_ids=groups)
        lines = (d / "snapshot_000007.csv").read_text().splitlines()
        assert lines[0] == "# time=0.125" and len(lines) == n + 2
        for i in (0, 8191, 8192, 16384, n - 1):
            want = ",".join([str(i), str(groups[i])] + ["%.17g" % v for v in bits[i]])
            assert lines[i + 2] == want, (threads, i)
        if threads == 1:
            first = lines
        else:
            assert lines == first


def test_bad_arguments(tmp_path):
    blocker = tmp_path / "file"
    blocker.write_text("x")
    with pytest.raises(GcError):
        SnapshotWriter(blocker / "sub")
    with SnapshotWriter(tmp_path / "ok") as out:
        with pytest.raises(ValueError):
            out.write_snapshot(0, 0.0, np.zeros((2, 3)), np.zeros((2, 3)), np.zeros((2, 3)),
                               np.zeros(2), group_ids=np.zeros(3))
        with pytest.raises(ValueError):
            out.write_diagnostics(0, 0.0, dict.fromkeys(DIAG_KEYS, 0.0))
        with pytest.raises(GcError):
            out.set_threads(-1)
    with pytest.raises(GcError):
        out.flush()  # closed


def _run_writer_check(tmp_path, golden, step):
    import subprocess
    from pathlib import Path

    exe = Path(__file__).resolve().parent / "cpp" / "writer_check"
    assert exe.exists(), "build() compiles tests/cpp/writer_check"
    src = tmp_path / "in.csv"
    src.write_bytes(_text(GOLDEN / golden))
    out = tmp_path / "out"
    res = subprocess.run([str(exe), str(src), str(step), str(out)], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    return out, res.stdout


def test_cpp_mirror_writes_reference_files(tmp_path):
    """HipSnapshotWriter fed a SimulationConfig-shaped config and a
    particle vector, as the reference's main loop feeds SnapshotWriter."""
    import gcsim

    if gcsim.device_count() > 0:
        pytest.skip("GPU visible: covered by the GPU test")
    out, _ = _run_writer_check(tmp_path, "direct_250_snapshot_000020.csv.gz", 20)
    assert (out / "snapshot_000020.csv").read_bytes() == \
        _text(GOLDEN / "direct_250_snapshot_000020.csv.gz")
    assert (out / "metadata.json").read_bytes() == _text(GOLDEN / "direct_250_metadata.json")
    rows = (out / "diagnostics.csv").read_text().splitlines()
    assert len(rows) == 2
    f = rows[1].split(",")
    assert f[:3] == ["20", "0.20000000000000001", "250"]
    want = [2.0, 0.25, -1.0 / 3.0, 0.25 - 1.0 / 3.0, 1e-300, -0.0, 5e-324, 0.1, 0.2, 0.3,
            -1.5, 2.5e10, 3.0]
    got = [float(x) for x in f[3:]]
    assert got == want and np.signbit(got[5])


@pytest.mark.parametrize("seed", [0, 1])
def test_against_compiled_reference_writer(tmp_path, seed):
    """Every file against the reference's SnapshotWriter compiled from its
    source (oracle/_ref), on values the goldens do not hold: random FP64
    bit patterns (NaN, infinities, subnormals, -0.0), 32-bit group ids,
    unusual names and parameters."""
    import oracle as O

    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(seed)
    n = 9000
    bits = rng.integers(-2**63, 2**63, (n, 10), dtype=np.int64).view(np.float64)
    bits[:5, 1:4] = [[0.0, -0.0, 5e-324], [np.inf, -np.inf, np.nan], [-np.nan, 1e308, -1e-308],
                     [0.1, 1e16, 1e17], [123456789.0, 1e-5, 1e-4]]
    groups = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    diag = rng.integers(-2**63, 2**63, 13, dtype=np.int64).view(np.float64)
    meta = dict(name='run "x" \\ y', solver="hip-fmm", particle_count=n, steps=123456,
                dt=float(rng.uniform(1e-7, 1)), snapshot_every=7, dim=3,
                gravitational_constant=6.6743e-11, softening=1.0 / 3.0, tree_theta=0.58,
                tree_leaf_capacity=64, fmm_expansion_order=2)
    step, time = int(rng.integers(0, 10**7)), float(rng.uniform(0, 100))
    O.ref_write_run(tmp_path / "ref", meta, step, time, bits[:, 1:4], bits[:, 4:7],
                    bits[:, 7:10], bits[:, 0], groups, diag)
    with SnapshotWriter(tmp_path / "ours") as out:
        out.write_metadata(**meta)
        out.write_snapshot(step, time, bits[:, 1:4], bits[:, 4:7], bits[:, 7:10], bits[:, 0],
                           group_ids=groups)
        out.write_diagnostics(step, time, dict(zip(DIAG_KEYS, diag)), particle_count=n)
    names = sorted(p.name for p in (tmp_path / "ref").iterdir())
    assert names == sorted(p.name for p in (tmp_path / "ours").iterdir())
    assert len(names) == 3
    for f in names:
        assert (tmp_path / "ours" / f).read_bytes() == (tmp_path / "ref" / f).read_bytes(), f


def test_output_cadence_follows_reference_loop():
    # cpp/core/main.cpp:171-184: step 0, every snapshot_every-th step, the last
    from gcsim.writer import output_steps

    assert output_steps(25, 10) == [0, 10, 20, 25]
    assert output_steps(20, 20) == [0, 20]
    assert output_steps(3, 1) == [0, 1, 2, 3]
    assert output_steps(0, 5) == [0]
    brute = [0] + [s for s in range(1, 300) if s % 7 == 0 or s == 299]
    assert output_steps(299, 7) == brute
    with pytest.raises(ValueError):
        output_steps(10, 0)
