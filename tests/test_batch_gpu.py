"""Batched stepping: each run of a batch is bit-identical to the same run
stepped alone through hip-direct (mixed) with the matching pair form and
launch shape; every schedule is deterministic and within the
mixed-precision bound of the oracle; graph-replayed steps equal eager
steps, and the autotuner's choice keeps results within the bound."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _runs(gpu, k, seed=2026):
    specs = gpu.generators.collisions(k, seed=seed, n_min=512, n_max=3000)
    parts = [gpu.generators.collision_particles(s, softening=0.05) for s in specs]
    return specs, parts


_ORACLE = {}


def _oracle_acc(p):
    """FP64 oracle forces of one run (cached: the ~48K-particle run takes
    seconds on all host cores)."""
    key = (len(p[2]), p[0][:2].tobytes())
    if key not in _ORACLE:
        _ORACLE[key] = O.direct_targets(p[0], p[2], 1.0, 0.05, threads=0)
    return _ORACLE[key]


def _cat(parts):
    return tuple(np.concatenate([p[i] for p in parts]) for i in range(3))


def _alone(gpu, pos, vel, m, chunk, steps, dt=1e-3, ipt=None):
    ntiles = (len(m) + 255) // 256
    with gpu.Engine("hip-direct", "mixed", softening=0.05) as eng:
        if ipt is None:
            # the scheduled batch kernel: per-target form, LDS source tiles
            eng.set_pairs("targets")
            eng.set_sources("lds")
            eng.set_launch_shape(1, (ntiles + chunk - 1) // chunk)
        else:
            eng.set_pairs("symmetric")
            eng.set_launch_shape(ipt, 1)
        eng.upload(pos, vel, m)
        eng.compute_accelerations()
        eng.step(dt, steps)
        return eng.download()


@pytest.mark.parametrize("chunk", [1, 3, 8])
def test_batched_runs_bit_identical_to_single_runs(gpu, chunk):
    from gcsim.batch import Batch

    specs, parts = _runs(gpu, 6)
    with Batch([len(p[2]) for p in parts], softening=0.05) as b:
        b.set_pairs("targets")
        b.set_schedule(chunk, 97, graph_steps=0)
        b.upload(*_cat(parts))
        b.compute_accelerations()
        b.step(1e-3, 7)
        got = b.download()
    for r, p in enumerate(parts):
        want = _alone(gpu, *p, chunk, 7)
        for g, w in zip(got, want):
            assert np.array_equal(b.split(g)[r], w), (r, chunk)


@pytest.mark.parametrize("ipt,blocks", [(1, 0), (2, 97), (4, 0)])
def test_symmetric_batch_bit_identical_to_single_runs(gpu, ipt, blocks):
    from gcsim.batch import Batch

    # runs of 2-12 tiles plus one of ~190 tiles, whose rows split into
    # multi-tile work items
    specs, parts = _runs(gpu, 5, seed=11)
    big = gpu.generators.collisions(1, seed=12, n_min=48000, n_max=49000)[0]
    parts.insert(2, gpu.generators.collision_particles(big, softening=0.05))
    with Batch([len(p[2]) for p in parts], softening=0.05) as b:
        b.set_pairs("symmetric", ipt)
        b.set_schedule(0, blocks, graph_steps=0)
        assert b.pairs() == ("symmetric", ipt)
        assert b.schedule() == (0, blocks, 0)
        b.upload(*_cat(parts))
        b.compute_accelerations()
        b.step(1e-3, 5)
        got = b.download()
    for r, p in enumerate(parts):
        want = _alone(gpu, *p, 0, 5, ipt=ipt)
        for g, w in zip(got, want):
            assert np.array_equal(b.split(g)[r], w), (r, ipt)


@pytest.mark.parametrize("ipt,tiles,blocks", [(1, 0, 0), (2, 1, 97), (4, 0, 0), (4, 3, 5000)])
def test_symmetric_batch_against_oracle_and_deterministic(gpu, ipt, tiles, blocks):
    from gcsim.batch import Batch

    # runs of 2-12 tiles plus one of ~190 tiles; items of the default, 1 or
    # 3 source tiles; one block per item or persistent grids of 97 and 5000
    specs, parts = _runs(gpu, 5, seed=11)
    big = gpu.generators.collisions(1, seed=12, n_min=48000, n_max=49000)[0]
    parts.insert(2, gpu.generators.collision_particles(big, softening=0.05))
    out = []
    for _ in range(2):
        with Batch([len(p[2]) for p in parts], softening=0.05) as b:
            b.set_pairs("symmetric", ipt)
            b.set_schedule(tiles, blocks, graph_steps=0)
            assert b.schedule() == (tiles, blocks, 0)
            b.upload(*_cat(parts))
            b.compute_accelerations()
            out.append(b.download()[2])
    assert np.array_equal(out[0], out[1])
    for r, p in enumerate(parts):
        ref = _oracle_acc(p)
        assert rel_err(b.split(out[0])[r], ref).max() <= 1e-5, (r, ipt, blocks)


@pytest.mark.parametrize("pairs", ["targets", "symmetric"])
def test_graph_replay_equals_eager(gpu, pairs):
    from gcsim.batch import Batch

    specs, parts = _runs(gpu, 5, seed=7)
    data = _cat(parts)
    out = []
    for graph_steps in (0, 10, 4):
        with Batch([len(p[2]) for p in parts], softening=0.05) as b:
            b.set_pairs(pairs)
            b.set_schedule(4, 512, graph_steps=graph_steps)
            b.upload(*data)
            b.compute_accelerations()
            b.step(1e-3, 23)  # graph chunks plus an eager remainder
            b.step(1e-3, 10)  # replays the same graph again
            out.append(b.download())
    for other in out[1:]:
        for g, w in zip(other, out[0]):
            assert np.array_equal(g, w)


@pytest.mark.parametrize("pairs", ["auto", "targets"])
def test_autotune_and_forces_against_oracle(gpu, pairs):
    from gcsim.batch import Batch

    specs, parts = _runs(gpu, 12, seed=3)
    with Batch([len(p[2]) for p in parts], softening=0.05) as b:
        b.set_pairs(pairs)
        b.upload(*_cat(parts))
        chunk, blocks = b.autotune()
        form, ipt = b.pairs()
        if form == "targets":
            assert chunk >= 1 and blocks >= 1 and ipt == 1
        else:
            assert pairs == "auto" and chunk >= 0 and blocks >= 0 and ipt in (1, 2, 4)
        assert b.schedule()[:2] == (chunk, blocks)
        assert b.autotune() == (chunk, blocks)  # cached
        assert b.pairs() == (form, ipt)
        b.compute_accelerations()
        acc = b.download()[2]
        ms, pairs = b.time_force_kernel(3)
        assert ms > 0 and pairs == sum(len(p[2]) * (len(p[2]) - 1) for p in parts)
    for r, p in enumerate(parts):
        ref = O.direct_targets(p[0], p[2], 1.0, 0.05)
        assert rel_err(b.split(acc)[r], ref).max() <= 1e-5


@pytest.mark.parametrize("pairs", ["targets", "symmetric"])
def test_single_particle_runs_and_bad_input(gpu, pairs):
    from gcsim.batch import Batch

    with Batch([1, 255, 256, 257], softening=0.05) as b:
        b.set_pairs(pairs)
        rng = np.random.default_rng(1)
        n = b.n
        pos, mass = rng.normal(size=(n, 3)), rng.uniform(0.5, 1, n)
        b.upload(pos, None, mass)
        b.compute_accelerations()
        acc = b.download()[2]
        assert not acc[0].any()  # a lone particle feels nothing
        assert np.isfinite(acc).all()
        for p, m, a in zip(b.split(pos)[1:], b.split(mass)[1:], b.split(acc)[1:]):
            assert rel_err(a, O.direct_targets(p, m, 1.0, 0.05)).max() <= 1e-5
        with pytest.raises(gpu.GcError):
            b.set_pairs("symmetric", 3)
    with pytest.raises(gpu.GcError):
        Batch([0, 5])


_TUNE_SCRIPT = r'''
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import gcsim.generators as G
from gcsim.batch import Batch
specs = G.collisions(6, seed=5, n_min=512, n_max=2000)
parts = [G.collision_particles(s, softening=0.05) for s in specs]
with Batch([len(p[2]) for p in parts], softening=0.05) as b:
    b.set_pairs("targets")
    b.upload(*(np.concatenate([p[i] for p in parts]) for i in range(3)))
    chunk, blocks = b.autotune()
    print(chunk, blocks, b.pairs()[0], flush=True)
'''


def test_autotune_cache_file_reused_by_a_new_process(gpu, tmp_path):
    # The tuner's result for a (device model, batch shape, pair form) key is
    # appended to the cache file; a later process takes it from there
    # instead of timing.  Shown by editing the stored schedule to one the
    # tuner never picks (chunk 3 is not a candidate) and reading it back.
    import os
    import subprocess
    import sys
    from pathlib import Path

    pkg = Path(__file__).resolve().parent.parent / "galaxy-collision-simulation_amd"
    cache = tmp_path / "tune.txt"
    env = dict(os.environ, GCSIM_TUNE_CACHE=str(cache))

    def run():
        out = subprocess.run([sys.executable, "-c", _TUNE_SCRIPT, str(pkg)], env=env,
                             capture_output=True, text=True, timeout=180)
        assert out.returncode == 0, out.stderr[-2000:]
        return out.stdout.split()

    first = run()
    assert first[2] == "targets" and int(first[0]) >= 1
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and lines[0].startswith("gcsim-tune 1 ")
    f = lines[0].split()
    assert len(f) == 14 and f[8] == "1" and (f[9], f[10]) == (first[0], first[1])
    f[9], f[10] = "3", "777"
    cache.write_text(" ".join(f) + "\n")
    assert run() == ["3", "777", "targets"]
    cache.write_text("gcsim-tune 1 garbage\n" + " ".join(f[:9]) + " 0 -5 1 0 0\n")
    second = run()  # no usable line: tuned again and appended
    assert second[2] == "targets" and int(second[0]) >= 1
    assert len(cache.read_text().splitlines()) == 3
