"""Synthetic initial conditions of the benchmark workloads (host side, via
the gcsim C ABI so every consumer sees the same particles for a seed).

two_disk(n_per_galaxy, ...)  central mass + exponential disk per galaxy,
                             galaxy B inclined about x
plummer(n, ...)              isotropic Plummer sphere (Aarseth-Henon-Wielen)
collisions(runs, ...)        per-run parameters of a batch of collisions
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L

_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _galaxy(n, centre, velocity, inclination, central_mass, disk_mass, disk_particle_mass,
            scale_length, r_min, r_max, z_sigma, v_sigma):
    g = L.DiskGalaxy()
    L.load().gc_default_disk_galaxy(C.byref(g))
    g.n_particles = int(n)
    g.central_mass = central_mass
    g.disk_mass = disk_mass
    g.disk_particle_mass = disk_particle_mass
    g.scale_length = scale_length
    g.r_min, g.r_max = r_min, r_max
    g.z_sigma, g.v_sigma = z_sigma, v_sigma
    g.inclination = inclination
    for d in range(3):
        g.position[d] = centre[d]
        g.velocity[d] = velocity[d]
    return g


def two_disk_galaxies(n_per_galaxy=500, central_mass=1.0, disk_mass=0.1, disk_particle_mass=0.0,
                      inclination_deg=45.0, separation=5.0, impact_parameter=1.5, speed=0.7,
                      scale_length=0.3, r_min=0.05, r_max=1.0, z_sigma=0.02, v_sigma=0.01):
    """The two galaxies of a two-disk collision (ctypes array of 2).
    Galaxy A at (-sep/2, -b/2, 0) moving +speed/2 in x, galaxy B at
    (+sep/2, +b/2, 0) moving -speed/2, B inclined about x."""
    gals = (L.DiskGalaxy * 2)()
    gals[0] = _galaxy(n_per_galaxy, (-separation / 2, -impact_parameter / 2, 0.0),
                      (speed / 2, 0.0, 0.0), 0.0, central_mass, disk_mass, disk_particle_mass,
                      scale_length, r_min, r_max, z_sigma, v_sigma)
    gals[1] = _galaxy(n_per_galaxy, (separation / 2, impact_parameter / 2, 0.0),
                      (-speed / 2, 0.0, 0.0), math.radians(inclination_deg), central_mass,
                      disk_mass, disk_particle_mass, scale_length, r_min, r_max, z_sigma, v_sigma)
    return gals


def two_disk(n_per_galaxy=500, softening=0.05, G=1.0, seed=42, **galaxy_kw):
    """Two-disk collision on the host (two_disk_galaxies' keywords).  The
    defaults give centres (-2.5, -0.75, 0), (2.5, 0.75, 0) and velocities
    (+-0.35, 0, 0)."""
    gals = two_disk_galaxies(n_per_galaxy, **galaxy_kw)
    n = 2 * int(n_per_galaxy)
    pos, vel, mass = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    got = L.load().gc_generate_disks(gals, 2, G, softening, seed, _ptr(pos), _ptr(vel),
                                     _ptr(mass))
    if got != n:
        raise RuntimeError("two_disk: generator rejected its arguments")
    return pos, vel, mass


def config0(seed=42):
    """N=1000: central 1.0 + 499 disk particles of mass 2e-4 per galaxy, eps 0.05."""
    return two_disk(500, softening=0.05, seed=seed, disk_particle_mass=2e-4)


def config1(seed=20260502, n=65536):
    """N=65,536: central 1.0 + disk mass 0.1 per galaxy, eps 0.02."""
    return two_disk(n // 2, softening=0.02, seed=seed, disk_mass=0.1)


def plummer(n, total_mass=1.0, scale=1.0, G=1.0, seed=1):
    pos, vel, mass = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    got = L.load().gc_generate_plummer(n, total_mass, scale, G, seed, _ptr(pos), _ptr(vel),
                                       _ptr(mass))
    if got != n:
        raise RuntimeError("plummer: generator rejected its arguments")
    return pos, vel, mass


def collisions(runs=256, seed=2026, n_min=512, n_max=8192):
    out = (L.Collision * runs)()
    rc = L.load().gc_sample_collisions(runs, seed, n_min, n_max, out)
    if rc != 0:
        raise RuntimeError("collisions: bad arguments")
    return [dict(n=int(c.n_particles), inclination=c.inclination,
                 impact_parameter=c.impact_parameter, speed=c.speed, seed=int(c.seed))
            for c in out]


def collision_galaxies(run: dict):
    """The two galaxies of one sampled collision (disk mass 0.1 each)."""
    return two_disk_galaxies(run["n"] // 2, inclination_deg=math.degrees(run["inclination"]),
                             impact_parameter=run["impact_parameter"], speed=run["speed"])


def collision_particles(run: dict, softening=0.05, G=1.0):
    """Particles of one sampled collision (two equal galaxies, disk mass 0.1)."""
    return two_disk(run["n"] // 2, softening=softening, G=G, seed=run["seed"],
                    inclination_deg=math.degrees(run["inclination"]),
                    impact_parameter=run["impact_parameter"], speed=run["speed"])


GALAXY_KEYS = ("n_particles", "mass", "radius", "position", "velocity", "orientation",
               "group_id", "thickness", "inclination")


def galaxy_configs(galaxies):
    """ctypes array of GalaxyConfig from dicts of its fields (missing fields
    take GalaxyConfig's defaults, cpp/core/initial_conditions.hpp:14-24)."""
    arr = (L.GalaxyConfig * len(galaxies))()
    for k, gdict in enumerate(galaxies):
        unknown = set(gdict) - set(GALAXY_KEYS)
        if unknown:
            raise ValueError(f"galaxy {k}: unknown fields {sorted(unknown)}")
        c = L.GalaxyConfig()
        L.load().gc_default_galaxy_config(C.byref(c))
        for key, val in gdict.items():
            if key in ("position", "velocity"):
                for d in range(3):
                    getattr(c, key)[d] = float(val[d])
            else:
                setattr(c, key, val)
        arr[k] = c
    return arr


def reference_galaxies(galaxies, G=1.0, softening=0.02, seed=42):
    """generate_galaxies (cpp/core/initial_conditions.cpp:94-114): the
    reference's disk galaxies, bit for bit.  galaxies: dicts of GalaxyConfig
    fields.  Returns pos (n, 3), vel (n, 3), mass (n,), group_id (n,)."""
    arr = galaxy_configs(galaxies)
    n = sum(int(c.n_particles) for c in arr if c.n_particles > 0 and c.mass > 0 and c.radius > 0)
    pos, vel, mass = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    grp = np.zeros(n, dtype=np.uint32)
    got = L.load().gc_generate_galaxies(arr, len(galaxies), float(G), float(softening), int(seed),
                                        _ptr(pos), _ptr(vel), _ptr(mass),
                                        grp.ctypes.data_as(C.POINTER(C.c_uint32)))
    if got != n:
        raise RuntimeError("reference_galaxies: generator rejected its arguments")
    return pos, vel, mass, grp

