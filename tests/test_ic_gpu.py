"""Initial conditions drawn on the device, through the C ABI, against the
oracle's restatement of the same counter-based models (oracle.ic_generate).

The Philox draws are integer arithmetic and identical on both sides; the
transforms call log, exp, sin, cos, pow and sqrt, whose device and host
versions may differ in the last ulp, so positions, velocities and masses
are compared to 1e-13 relative (with an absolute floor of 1e-13)."""
import numpy as np
import pytest

import oracle as O
from gcsim import generators as G
from gcsim.batch import Batch

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _close(got, want):
    assert got.shape == want.shape
    assert np.all(np.abs(got - want) <= TOL * np.maximum(1.0, np.abs(want)))


def _ref_params(c, g, eps):
    return [c["n_particles"], c["mass"], c["radius"], *c["position"], *c["velocity"],
            c["orientation"], c["thickness"], c["inclination"], g, eps]


def _disk_params(gal, g, eps):
    nd = gal.n_particles - 1
    md = gal.disk_particle_mass if gal.disk_particle_mass > 0 else gal.disk_mass / nd
    return [gal.n_particles, gal.central_mass, md, md * nd, gal.scale_length, gal.r_min,
            gal.r_max, gal.z_sigma, gal.v_sigma, gal.inclination, *gal.position, *gal.velocity,
            g, eps]


def test_reference_galaxies_on_device(gpu):
    gal = [dict(n_particles=3000, mass=1.0, radius=0.85, position=[-0.72, -0.10, 0.06],
                velocity=[0.34, 0.10, -0.015], orientation=0.25, group_id=0, thickness=0.045,
                inclination=0.62),
           dict(n_particles=0, mass=1.0, radius=1.0, position=[0, 0, 0], velocity=[0, 0, 0],
                orientation=0.0, group_id=1, thickness=0.0, inclination=0.0),
           dict(n_particles=1001, mass=2.0, radius=0.5, position=[0.72, 0.10, -0.06],
                velocity=[-0.34, -0.10, 0.015], orientation=3.42, group_id=2, thickness=0.0,
                inclination=-0.72)]
    seed, g, eps = 20260502, 1.3, 0.025
    with gpu.Engine("hip-direct", "fp64", softening=eps, G=g) as eng:
        eng.generate_reference_galaxies(gal, seed)
        pos, vel, acc = eng.download()
        assert eng.n == 4001
    off = 0
    for k, c in enumerate(gal):
        if c["n_particles"] == 0:
            continue
        n = c["n_particles"]
        wp, wv, wm = O.ic_generate(O.IC_REF_DISK, _ref_params(c, g, eps), seed, k, n)
        _close(pos[off:off + n], wp)
        _close(vel[off:off + n], wv)
        off += n
    assert off == 4001 and np.all(acc == 0.0)


def test_two_disk_and_plummer_on_device(gpu):
    seed = 77
    with gpu.Engine("hip-direct", "mixed", softening=0.05) as eng:
        eng.generate_two_disk(700, seed=seed, inclination_deg=30.0, disk_particle_mass=2e-4)
        pos, vel, _ = eng.download()
    gals = G.two_disk_galaxies(700, inclination_deg=30.0, disk_particle_mass=2e-4)
    for k in range(2):
        wp, wv, _ = O.ic_generate(O.IC_TWO_DISK, _disk_params(gals[k], 1.0, 0.05), seed, k, 700)
        _close(pos[700 * k:700 * (k + 1)], wp)
        _close(vel[700 * k:700 * (k + 1)], wv)
    with gpu.Engine("hip-direct", "mixed", softening=0.01) as eng:
        eng.generate_plummer(20000, total_mass=2.0, scale=0.5, seed=3)
        pos, vel, _ = eng.download()
    wp, wv, _ = O.ic_generate(O.IC_PLUMMER, [20000, 2.0, 0.5, 1.0], 3, 0, 20000)
    _close(pos, wp)
    _close(vel, wv)


def test_device_generation_is_deterministic_and_masses_exact(gpu):
    with gpu.Engine("hip-direct", "mixed", softening=0.02) as eng:
        eng.generate_plummer(5000, seed=9)
        a = eng.download()
        eng.generate_plummer(5000, seed=9)
        b = eng.download()
        eng.generate_plummer(5000, seed=10)
        c = eng.download()
        eng.upload(np.zeros((3, 3)), None, np.ones(3))  # then a smaller state
        eng.generate_plummer(1, seed=9)
        d = eng.download()
        assert eng.n == 1
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    assert np.array_equal(d[0][0], a[0][0])  # particle 0 of the same stream


def test_batch_collisions_on_device(gpu):
    runs = G.collisions(runs=6, seed=4, n_min=300, n_max=1500)
    counts = [r["n"] for r in runs]
    eps = 0.05
    with Batch(counts, softening=eps) as b:
        b.generate_collisions(runs)
        pos, vel, acc = b.download()
        b.compute_accelerations()
        b.step(1e-3, 2)
        pos2, _, _ = b.download()
    off = 0
    for r, run in enumerate(runs):
        gals = G.collision_galaxies(run)
        half = run["n"] // 2
        for k in range(2):
            wp, wv, _ = O.ic_generate(O.IC_TWO_DISK, _disk_params(gals[k], 1.0, eps),
                                      run["seed"], k, half)
            _close(pos[off + k * half:off + (k + 1) * half], wp)
            _close(vel[off + k * half:off + (k + 1) * half], wv)
        off += run["n"]
    assert off == len(pos) and np.all(acc == 0.0)
    assert np.all(np.isfinite(pos2)) and not np.array_equal(pos, pos2)


def test_device_generation_bad_arguments(gpu):
    with gpu.Engine("hip-direct", "mixed") as eng:
        with pytest.raises(gpu.GcError):
            eng.generate_plummer(10, scale=0.0)
        # counts whose sum wraps to 0 in 64 bits are rejected before any launch
        with pytest.raises(gpu.GcError):
            eng.generate_reference_galaxies([dict(n_particles=2 ** 63), dict(n_particles=2 ** 63)])
        with pytest.raises(gpu.GcError):
            eng.generate_two_disk(n_per_galaxy=2 ** 31)
    runs = G.collisions(runs=2, seed=4, n_min=300, n_max=600)
    with Batch([r["n"] for r in runs]) as b:
        bad = [dict(runs[0]), dict(runs[1])]
        bad[1]["n"] += 2
        with pytest.raises(gpu.GcError):
            b.generate_collisions(bad)
