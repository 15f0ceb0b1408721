"""ctypes binding for the CPU oracle (liboracle.so) and the compiled reference
(_ref/libgcref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product path never imports
this module.  Arrays are numpy float64, particle-major (n, 3).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_dbl_p = C.POINTER(C.c_double)
_size_t = C.c_size_t


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dbl_p)


class FmmStats(C.Structure):
    _fields_ = [("node_count", _size_t), ("leaf_count", _size_t),
                ("far_interactions", _size_t), ("near_interactions", _size_t),
                ("max_depth_reached", C.c_int)]


class Solver(C.Structure):
    _fields_ = [("solver", C.c_int), ("G", C.c_double), ("eps", C.c_double),
                ("theta", C.c_double), ("leaf_capacity", _size_t),
                ("max_depth", C.c_int), ("expansion_order", C.c_int)]


SOLVER_DIRECT_SYMMETRIC, SOLVER_DIRECT_TARGETS, SOLVER_FMM, SOLVER_TREE = 0, 1, 2, 3

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        _lib = C.CDLL(str(path))
    return _lib


def ref_lib():
    """The compiled reference (None when it was not built here)."""
    global _ref
    if _ref is None:
        path = HERE / "_ref" / "libgcref.so"
        if not path.exists():
            return None
        _ref = C.CDLL(str(path))
    return _ref


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- direct
def direct_symmetric(pos, mass, G=1.0, eps=0.02):
    pos, mass = _c64(pos), _c64(mass)
    acc = np.zeros_like(pos)
    lib().orc_direct_symmetric(_size_t(len(mass)), _p(pos), _p(mass), C.c_double(G),
                               C.c_double(eps), _p(acc))
    return acc


def direct_targets(pos, mass, G=1.0, eps=0.02, begin=0, end=None, threads=1):
    pos, mass = _c64(pos), _c64(mass)
    n = len(mass)
    end = n if end is None else end
    acc = np.zeros_like(pos)
    if threads == 1:
        lib().orc_direct_targets(_size_t(n), _p(pos), _p(mass), C.c_double(G), C.c_double(eps),
                                 _size_t(begin), _size_t(end), _p(acc))
    else:
        lib().orc_direct_targets_omp.restype = C.c_int
        lib().orc_direct_targets_omp(_size_t(n), _p(pos), _p(mass), C.c_double(G),
                                     C.c_double(eps), _size_t(begin), _size_t(end), _p(acc),
                                     C.c_int(threads))
    return acc


def direct_targets_omp_threads(threads=0):
    """Threads the OpenMP targets form would use (0 = all host cores)."""
    pos = np.zeros((2, 3))
    pos[1, 0] = 1.0
    m = np.ones(2)
    acc = np.zeros_like(pos)
    lib().orc_direct_targets_omp.restype = C.c_int
    return int(lib().orc_direct_targets_omp(_size_t(2), _p(pos), _p(m), C.c_double(1.0),
                                            C.c_double(0.0), _size_t(0), _size_t(2), _p(acc),
                                            C.c_int(threads)))


def softened_acceleration(target, source, m, G=1.0, eps=0.02):
    t, s = _c64(target), _c64(source)
    out = np.zeros(3)
    lib().orc_softened_acceleration(_p(t), _p(s), C.c_double(m), C.c_double(G),
                                    C.c_double(eps), _p(out))
    return out


# ---------------------------------------------------------------- integrator
def make_solver(kind=SOLVER_DIRECT_SYMMETRIC, G=1.0, eps=0.02, theta=0.6, leaf_capacity=16,
                max_depth=32, expansion_order=4):
    return Solver(kind, G, eps, theta, leaf_capacity, max_depth, expansion_order)


def compute_accelerations(solver: Solver, pos, mass):
    pos, mass = _c64(pos), _c64(mass)
    acc = np.zeros_like(pos)
    rc = lib().orc_compute_accelerations(C.byref(solver), _size_t(len(mass)), _p(pos), _p(mass),
                                         _p(acc))
    if rc:
        raise RuntimeError(f"oracle compute failed: {rc}")
    return acc


def leapfrog(solver: Solver, pos, vel, acc, mass, dt, steps):
    """`steps` KDK steps; returns new (pos, vel, acc).  acc must be a(pos)."""
    pos, vel, acc, mass = (_c64(pos).copy(), _c64(vel).copy(), _c64(acc).copy(), _c64(mass))
    rc = lib().orc_leapfrog(C.byref(solver), _size_t(len(mass)), _p(pos), _p(vel), _p(acc),
                            _p(mass), C.c_double(dt), C.c_int(steps))
    if rc:
        raise RuntimeError(f"oracle leapfrog failed: {rc}")
    return pos, vel, acc


def kick(vel, acc, dt):
    vel = _c64(vel).copy()
    lib().orc_kick(_size_t(len(vel)), _p(vel), _p(_c64(acc)), C.c_double(dt))
    return vel


def drift(pos, vel, dt):
    pos = _c64(pos).copy()
    lib().orc_drift(_size_t(len(pos)), _p(pos), _p(_c64(vel)), C.c_double(dt))
    return pos


# ---------------------------------------------------------------- diagnostics
DIAG_FIELDS = ("total_mass", "kinetic_energy", "potential_energy", "total_energy",
               "momentum_x", "momentum_y", "momentum_z",
               "center_of_mass_x", "center_of_mass_y", "center_of_mass_z",
               "angular_momentum_x", "angular_momentum_y", "angular_momentum_z")


def diagnostics(pos, vel, mass, G=1.0, eps=0.02):
    out = np.zeros(16)
    lib().orc_diagnostics(_size_t(len(mass)), _p(_c64(pos)), _p(_c64(vel)), _p(_c64(mass)),
                          C.c_double(G), C.c_double(eps), _p(out))
    return dict(zip(DIAG_FIELDS, out[:13]))


# ---------------------------------------------------------------- multipoles
def exponents():
    out = np.zeros((35, 3), dtype=np.int32)
    lib().orc_multipole_exponents(out.ctypes.data_as(C.POINTER(C.c_int)))
    return out


def p2m(moments, offset, mass):
    mom = _c64(moments).copy()
    lib().orc_p2m(_p(mom), _p(_c64(offset)), C.c_double(mass))
    return mom


def m2m(parent, child, offset, child_mass):
    par = _c64(parent).copy()
    lib().orc_m2m(_p(par), _p(_c64(child)), _p(_c64(offset)), C.c_double(child_mass))
    return par


def m2p(target, com, mass, moments, G=1.0, eps=0.02, order=4):
    out = np.zeros(3)
    lib().orc_m2p(_p(_c64(target)), _p(_c64(com)), C.c_double(mass), _p(_c64(moments)),
                  C.c_double(G), C.c_double(eps), C.c_int(order), _p(out))
    return out


def fmm(pos, mass, G=1.0, eps=0.02, theta=0.6, leaf_capacity=16, max_depth=32,
        expansion_order=4, begin=0, end=None):
    pos, mass = _c64(pos), _c64(mass)
    n = len(mass)
    end = n if end is None else end
    acc = np.zeros_like(pos)
    st = FmmStats()
    rc = lib().orc_fmm(_size_t(n), _p(pos), _p(mass), C.c_double(G), C.c_double(eps),
                       C.c_double(theta), _size_t(leaf_capacity), C.c_int(max_depth),
                       C.c_int(expansion_order), _size_t(begin), _size_t(end), _p(acc),
                       C.byref(st))
    if rc:
        raise RuntimeError(f"oracle fmm failed: {rc}")
    stats = {k: getattr(st, k) for k, _ in FmmStats._fields_}
    return acc, stats


class TreeStats(C.Structure):
    _fields_ = [("node_count", C.c_size_t), ("leaf_count", C.c_size_t),
                ("m2p_evaluations", C.c_size_t), ("p2p_pairs", C.c_size_t),
                ("max_depth_reached", C.c_int)]


def tree(pos, mass, G=1.0, eps=0.02, theta=0.6, leaf_capacity=16, max_depth=32,
         expansion_order=4, begin=0, end=None):
    """Barnes-Hut accelerations (cpp/fmm/quadtree.cpp) and walk statistics."""
    pos, mass = _c64(pos), _c64(mass)
    n = len(mass)
    end = n if end is None else end
    acc = np.zeros_like(pos)
    st = TreeStats()
    rc = lib().orc_tree(_size_t(n), _p(pos), _p(mass), C.c_double(G), C.c_double(eps),
                        C.c_double(theta), _size_t(leaf_capacity), C.c_int(max_depth),
                        C.c_int(expansion_order), _size_t(begin), _size_t(end), _p(acc),
                        C.byref(st))
    if rc:
        raise RuntimeError(f"oracle tree failed: {rc}")
    stats = {k: getattr(st, k) for k, _ in TreeStats._fields_}
    return acc, stats


# ---------------------------------------------------------------- reference
def ref_available() -> bool:
    """Whether the compiled reference exists (checked without loading it, so
    collecting tests that skip on it loads nothing)."""
    return (HERE / "_ref" / "libgcref.so").exists()


def ref_direct_symmetric(pos, mass, G=1.0, eps=0.02):
    pos, mass = _c64(pos), _c64(mass)
    acc = np.zeros_like(pos)
    ref_lib().ref_direct_symmetric(_size_t(len(mass)), _p(pos), _p(mass), C.c_double(G),
                                   C.c_double(eps), _p(acc))
    return acc


def ref_direct_targets(pos, mass, G=1.0, eps=0.02, begin=0, end=None):
    pos, mass = _c64(pos), _c64(mass)
    n = len(mass)
    end = n if end is None else end
    acc = np.zeros_like(pos)
    ref_lib().ref_direct_targets(_size_t(n), _p(pos), _p(mass), C.c_double(G), C.c_double(eps),
                                 _size_t(begin), _size_t(end), _p(acc))
    return acc


def ref_leapfrog(solver, pos, vel, acc, mass, G, eps, dt, steps, theta=0.6, leaf=16, order=4):
    pos, vel, acc, mass = (_c64(pos).copy(), _c64(vel).copy(), _c64(acc).copy(), _c64(mass))
    ref_lib().ref_leapfrog(C.c_int(solver), _size_t(len(mass)), _p(pos), _p(vel), _p(acc),
                           _p(mass), C.c_double(G), C.c_double(eps), C.c_double(theta),
                           _size_t(leaf), C.c_int(order), C.c_double(dt), C.c_int(steps))
    return pos, vel, acc


def ref_diagnostics(pos, vel, mass, G=1.0, eps=0.02):
    out = np.zeros(16)
    ref_lib().ref_diagnostics(_size_t(len(mass)), _p(_c64(pos)), _p(_c64(vel)), _p(_c64(mass)),
                              C.c_double(G), C.c_double(eps), _p(out))
    return dict(zip(DIAG_FIELDS, out[:13]))


def ref_fmm(pos, mass, G=1.0, eps=0.02, theta=0.6, leaf_capacity=16, max_depth=32,
            expansion_order=4, begin=0, end=None):
    pos, mass = _c64(pos), _c64(mass)
    n = len(mass)
    end = n if end is None else end
    acc = np.zeros_like(pos)
    st = (C.c_size_t * 4)()
    ref_lib().ref_fmm(_size_t(n), _p(pos), _p(mass), C.c_double(G), C.c_double(eps),
                      C.c_double(theta), _size_t(leaf_capacity), C.c_int(max_depth),
                      C.c_int(expansion_order), _size_t(begin), _size_t(end), _p(acc), st)
    return acc, dict(node_count=st[0], leaf_count=st[1], far_interactions=st[2],
                     near_interactions=st[3])


def ref_tree(pos, mass, G=1.0, eps=0.02, theta=0.6, leaf_capacity=16, expansion_order=4):
    pos, mass = _c64(pos), _c64(mass)
    acc = np.zeros_like(pos)
    ref_lib().ref_tree(_size_t(len(mass)), _p(pos), _p(mass), C.c_double(G), C.c_double(eps),
                       C.c_double(theta), _size_t(leaf_capacity), C.c_int(expansion_order),
                       _p(acc))
    return acc


def ref_p2m(moments, offset, mass):
    mom = _c64(moments).copy()
    ref_lib().ref_p2m(_p(mom), _p(_c64(offset)), C.c_double(mass))
    return mom


def ref_m2m(parent, child, offset, child_mass):
    par = _c64(parent).copy()
    ref_lib().ref_m2m(_p(par), _p(_c64(child)), _p(_c64(offset)), C.c_double(child_mass))
    return par


def ref_m2p(target, com, mass, moments, G=1.0, eps=0.02, order=4):
    out = np.zeros(3)
    ref_lib().ref_m2p(_p(_c64(target)), _p(_c64(com)), C.c_double(mass), _p(_c64(moments)),
                      C.c_double(G), C.c_double(eps), C.c_int(order), _p(out))
    return out


def ref_generate(galaxies, G, eps, seed):
    """galaxies: list of dicts with the reference GalaxyConfig fields."""
    keys = ("n_particles", "mass", "radius", "position", "velocity", "orientation", "group_id",
            "thickness", "inclination")
    rows = []
    for g in galaxies:
        row = [g["n_particles"], g["mass"], g["radius"], *g["position"], *g["velocity"],
               g["orientation"], g["group_id"], g["thickness"], g["inclination"]]
        assert len(row) == 13, keys
        rows.append(row)
    gal = np.ascontiguousarray(np.array(rows, dtype=np.float64))
    n = int(sum(g["n_particles"] for g in galaxies))
    pos, vel, mass = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    grp = np.zeros(n, dtype=np.uint32)
    ref_lib().ref_generate.restype = C.c_size_t
    got = ref_lib().ref_generate(_p(gal), C.c_int(len(galaxies)), C.c_double(G), C.c_double(eps),
                                 C.c_ulonglong(seed), _p(pos), _p(vel), _p(mass),
                                 grp.ctypes.data_as(C.POINTER(C.c_uint)))
    return pos[:got], vel[:got], mass[:got], grp[:got]


def ref_write_run(directory, meta, step, time, pos, vel, acc, mass, groups, diag, snapshots=1):
    """The reference's SnapshotWriter (cpp/io/snapshot_writer.cpp) on
    `directory`: metadata.json from `meta` (metadata.json keys), `snapshots`
    copies of one snapshot (steps step, step + 1, ...) and one
    diagnostics row (`diag`: 13 values in diagnostics.csv column order)."""
    n = len(mass)
    g = np.ascontiguousarray(groups if groups is not None else np.zeros(n), dtype=np.uint32)
    d = _c64(diag)
    assert d.shape == (13,)
    rc = ref_lib().ref_write_run(
        os.fsencode(str(directory)), meta["name"].encode(), meta["solver"].encode(),
        C.c_int(meta["steps"]), C.c_double(meta["dt"]), C.c_int(meta["snapshot_every"]),
        C.c_double(meta["gravitational_constant"]), C.c_double(meta["softening"]),
        C.c_double(meta["tree_theta"]), _size_t(meta["tree_leaf_capacity"]),
        C.c_int(meta["fmm_expansion_order"]), C.c_int(step), C.c_double(time), _size_t(n),
        _p(_c64(pos)), _p(_c64(vel)), _p(_c64(acc)), _p(_c64(mass)),
        g.ctypes.data_as(C.POINTER(C.c_uint)), _p(d), C.c_int(snapshots))
    if rc:
        raise RuntimeError("reference SnapshotWriter threw")


# ---------------------------------------------------------------- initial conditions
IC_REF_DISK, IC_TWO_DISK, IC_PLUMMER = 0, 1, 2


def mt64_nth(seed, n):
    lib().orc_mt64_nth.restype = C.c_uint64
    return int(lib().orc_mt64_nth(C.c_uint64(seed), _size_t(n)))


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def _gal_rows(galaxies):
    rows = [[g["n_particles"], g["mass"], g["radius"], *g["position"], *g["velocity"],
             g["orientation"], g["thickness"], g["inclination"]] for g in galaxies]
    return np.ascontiguousarray(np.array(rows, dtype=np.float64).reshape(-1, 12))


def generate_galaxies(galaxies, G, eps, seed):
    """generate_galaxies restated (MT19937-64 + libstdc++ draws); galaxies:
    dicts with the GalaxyConfig fields."""
    gal = _gal_rows(galaxies)
    n = int(sum(g["n_particles"] for g in galaxies))
    pos, vel, mass = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    lib().orc_generate_galaxies.restype = _size_t
    got = lib().orc_generate_galaxies(_p(gal), C.c_int(len(galaxies)), C.c_double(G),
                                      C.c_double(eps), C.c_uint64(seed), _p(pos), _p(vel), _p(mass))
    return pos[:got], vel[:got], mass[:got]


def ic_generate(model, params, seed, stream, n):
    """The device generators' models on the Philox stream (seed, stream)."""
    par = _c64(params)
    pos, vel, mass = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    rc = lib().orc_ic_generate(C.c_int(model), _p(par), C.c_uint64(seed), C.c_uint32(stream),
                               _size_t(n), _p(pos), _p(vel), _p(mass))
    if rc:
        raise ValueError(f"unknown model {model}")
    return pos, vel, mass


__all__ = [n for n in dir() if not n.startswith("_")] + ["HERE"]
os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")
