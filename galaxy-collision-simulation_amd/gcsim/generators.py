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


def two_disk(n_per_galaxy=500, softening=0.05, G=1.0, seed=42, central_mass=1.0,
             disk_mass=0.1, disk_particle_mass=0.0, inclination_deg=45.0,
             separation=5.0, impact_parameter=1.5, speed=0.7, scale_length=0.3, r_min=0.05,
             r_max=1.0, z_sigma=0.02, v_sigma=0.01):
    """Two-disk collision.  Galaxy A at (-sep/2, -b/2, 0) moving +speed/2 in x,
    galaxy B at (+sep/2, +b/2, 0) moving -speed/2, B inclined about x.
    The defaults give centres (-2.5, -0.75, 0), (2.5, 0.75, 0) and
    velocities (+-0.35, 0, 0)."""
    gals = (L.DiskGalaxy * 2)()
    gals[0] = _galaxy(n_per_galaxy, (-separation / 2, -impact_parameter / 2, 0.0),
                      (speed / 2, 0.0, 0.0), 0.0, central_mass, disk_mass, disk_particle_mass,
                      scale_length, r_min, r_max, z_sigma, v_sigma)
    gals[1] = _galaxy(n_per_galaxy, (separation / 2, impact_parameter / 2, 0.0),
                      (-speed / 2, 0.0, 0.0), math.radians(inclination_deg), central_mass,
                      disk_mass, disk_particle_mass, scale_length, r_min, r_max, z_sigma, v_sigma)
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


def collision_particles(run: dict, softening=0.05, G=1.0):
    """Particles of one sampled collision (two equal galaxies, disk mass 0.1)."""
    return two_disk(run["n"] // 2, softening=softening, G=G, seed=run["seed"],
                    inclination_deg=math.degrees(run["inclination"]),
                    impact_parameter=run["impact_parameter"], speed=run["speed"])
