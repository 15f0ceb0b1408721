"""Batched stepping: each run of a batch is bit-identical to the same run
stepped alone through hip-direct (mixed) with the matching source split, the
graph-replayed steps equal eager steps, and the autotuner's choice keeps
results within the mixed-precision bound."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _runs(gpu, k, seed=2026):
    specs = gpu.generators.collisions(k, seed=seed, n_min=512, n_max=3000)
    parts = [gpu.generators.collision_particles(s, softening=0.05) for s in specs]
    return specs, parts


def _cat(parts):
    return tuple(np.concatenate([p[i] for p in parts]) for i in range(3))


def _alone(gpu, pos, vel, m, chunk, steps, dt=1e-3):
    ntiles = (len(m) + 255) // 256
    with gpu.Engine("hip-direct", "mixed", softening=0.05) as eng:
        # the batch kernel runs the per-target form with LDS source tiles
        eng.set_pairs("targets")
        eng.set_sources("lds")
        eng.set_launch_shape(1, (ntiles + chunk - 1) // chunk)
        eng.upload(pos, vel, m)
        eng.compute_accelerations()
        eng.step(dt, steps)
        return eng.download()


@pytest.mark.parametrize("chunk", [1, 3, 8])
def test_batched_runs_bit_identical_to_single_runs(gpu, chunk):
    from gcsim.batch import Batch

    specs, parts = _runs(gpu, 6)
    with Batch([len(p[2]) for p in parts], softening=0.05) as b:
        b.set_schedule(chunk, 97, graph_steps=0)
        b.upload(*_cat(parts))
        b.compute_accelerations()
        b.step(1e-3, 7)
        got = b.download()
    for r, p in enumerate(parts):
        want = _alone(gpu, *p, chunk, 7)
        for g, w in zip(got, want):
            assert np.array_equal(b.split(g)[r], w), (r, chunk)


def test_graph_replay_equals_eager(gpu):
    from gcsim.batch import Batch

    specs, parts = _runs(gpu, 5, seed=7)
    data = _cat(parts)
    out = []
    for graph_steps in (0, 10, 4):
        with Batch([len(p[2]) for p in parts], softening=0.05) as b:
            b.set_schedule(4, 512, graph_steps=graph_steps)
            b.upload(*data)
            b.compute_accelerations()
            b.step(1e-3, 23)  # graph chunks plus an eager remainder
            b.step(1e-3, 10)  # replays the same graph again
            out.append(b.download())
    for other in out[1:]:
        for g, w in zip(other, out[0]):
            assert np.array_equal(g, w)


def test_autotune_and_forces_against_oracle(gpu):
    from gcsim.batch import Batch

    specs, parts = _runs(gpu, 12, seed=3)
    with Batch([len(p[2]) for p in parts], softening=0.05) as b:
        b.upload(*_cat(parts))
        chunk, blocks = b.autotune()
        assert chunk >= 1 and blocks >= 1
        assert b.schedule()[:2] == (chunk, blocks)
        assert b.autotune() == (chunk, blocks)  # cached
        b.compute_accelerations()
        acc = b.download()[2]
        ms, pairs = b.time_force_kernel(3)
        assert ms > 0 and pairs == sum(len(p[2]) * (len(p[2]) - 1) for p in parts)
    for r, p in enumerate(parts):
        ref = O.direct_targets(p[0], p[2], 1.0, 0.05)
        assert rel_err(b.split(acc)[r], ref).max() <= 1e-5


def test_single_particle_runs_and_bad_input(gpu):
    from gcsim.batch import Batch

    with Batch([1, 255, 256, 257], softening=0.05) as b:
        rng = np.random.default_rng(1)
        n = b.n
        b.upload(rng.normal(size=(n, 3)), None, rng.uniform(0.5, 1, n))
        b.compute_accelerations()
        acc = b.download()[2]
        assert not acc[0].any()  # a lone particle feels nothing
        assert np.isfinite(acc).all()
    with pytest.raises(gpu.GcError):
        Batch([0, 5])
