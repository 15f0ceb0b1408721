"""Run output in the reference's formats (cpp/io/snapshot_writer.cpp):
metadata.json, diagnostics.csv and snapshot_NNNNNN.csv, byte for byte as
the reference writes them, so python/utils/snapshots.py reads them as is.

    with SnapshotWriter("outputs/run") as out:
        out.write_metadata(name="run", solver="hip-direct", particle_count=n,
                           steps=40, dt=0.01, snapshot_every=10, softening=0.02)
        out.write_snapshot(0, 0.0, eng)          # resident state, asynchronous
        out.write_diagnostics(0, 0.0, eng)       # device diagnostics
        eng.step(0.01, 10)
        out.write_snapshot(10, 0.1, eng)

write_snapshot of an Engine stages the resident state with one device copy
into pinned memory and returns; the writer's own thread formats and writes
the file while the device steps on.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L
from .engine import Engine, _c64, _ptr

DIAG_KEYS = ("total_mass", "kinetic_energy", "potential_energy", "total_energy",
             "momentum_x", "momentum_y", "momentum_z", "center_of_mass_x",
             "center_of_mass_y", "center_of_mass_z", "angular_momentum_x",
             "angular_momentum_y", "angular_momentum_z")


def _groups(group_ids, n):
    if group_ids is None:
        return None, None
    g = np.ascontiguousarray(group_ids, dtype=np.uint32)
    if g.shape != (n,):
        raise ValueError(f"group_ids: expected shape ({n},), got {g.shape}")
    return g, g.ctypes.data_as(C.POINTER(C.c_uint32))


class SnapshotWriter:
    """SnapshotWriter (cpp/io/snapshot_writer.hpp): creates `directory` and
    diagnostics.csv with its header line."""

    def __init__(self, directory):
        self._lib = L.load()
        h = C.c_void_p()
        L.check(self._lib.gc_writer_open(C.byref(h), os.fsencode(directory)), None,
                "writer open")
        self._w = h
        self.directory = str(directory)

    def close(self):
        if self._w:
            w, self._w = self._w, None
            L.check(self._lib.gc_writer_close(w), None, "writer close")

    def __del__(self):
        try:
            if self._w:
                self._lib.gc_writer_close(self._w)
                self._w = None
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _live(self):
        if not self._w:
            raise L.GcError(L.GC_ERR_STATE, "writer: closed")
        return self._w

    def write_metadata(self, *, name, solver, particle_count, steps, dt, snapshot_every,
                       dim=3, gravitational_constant=1.0, softening=0.02, tree_theta=0.6,
                       tree_leaf_capacity=16, fmm_expansion_order=4):
        """SnapshotWriter::write_metadata (snapshot_writer.cpp:50-70); the
        defaults are SimulationConfig's (cpp/core/config.hpp:19-34)."""
        info = L.RunInfo(name.encode(), solver.encode(), int(particle_count), int(steps),
                         float(dt), int(snapshot_every), int(dim),
                         float(gravitational_constant), float(softening), float(tree_theta),
                         int(tree_leaf_capacity), int(fmm_expansion_order))
        L.check(self._lib.gc_writer_metadata(self._live(), C.byref(info)), None,
                "writer metadata")

    def write_snapshot(self, step, time, source, vel=None, acc=None, mass=None,
                       group_ids=None):
        """SnapshotWriter::write_snapshot (snapshot_writer.cpp:72-96).

        source is an Engine (its resident state, staged asynchronously) or
        host positions (n, 3) with vel, acc (n, 3) and mass (n,)."""
        if isinstance(source, Engine):
            _, gp = _groups(group_ids, source.n)
            st = self._lib.gc_writer_snapshot(self._live(), source._ctx, int(step), float(time),
                                              gp)
            L.check(st, None, "writer snapshot")
            return
        pos = _c64(source)
        n = pos.shape[0]
        pos = _c64(pos, (n, 3))
        vel, acc, mass = _c64(vel, (n, 3)), _c64(acc, (n, 3)), _c64(mass, (n,))
        _, gp = _groups(group_ids, n)
        st = self._lib.gc_writer_snapshot_host(self._live(), int(step), float(time), n,
                                               _ptr(pos), _ptr(vel), _ptr(acc), _ptr(mass), gp)
        L.check(st, None, "writer snapshot")

    def write_diagnostics(self, step, time, source, particle_count=None):
        """SnapshotWriter::write_diagnostics (snapshot_writer.cpp:98-121).

        source is an Engine (FP64 diagnostics computed on the device) or a
        mapping of the diagnostics.csv columns with particle_count."""
        if isinstance(source, Engine):
            st = self._lib.gc_writer_diagnostics(self._live(), source._ctx, int(step),
                                                 float(time))
            L.check(st, None, "writer diagnostics")
            return
        if particle_count is None:
            raise ValueError("write_diagnostics: particle_count is required with values")
        d = L.Diagnostics()
        d.total_mass = source["total_mass"]
        d.kinetic_energy = source["kinetic_energy"]
        d.potential_energy = source["potential_energy"]
        d.total_energy = source["total_energy"]
        for i, ax in enumerate("xyz"):
            d.momentum[i] = source[f"momentum_{ax}"]
            d.center_of_mass[i] = source[f"center_of_mass_{ax}"]
            d.angular_momentum[i] = source[f"angular_momentum_{ax}"]
        st = self._lib.gc_writer_diagnostics_values(self._live(), int(step), float(time),
                                                    int(particle_count), C.byref(d))
        L.check(st, None, "writer diagnostics")

    def set_threads(self, threads=0):
        """Host threads formatting one snapshot's rows (0 = default)."""
        L.check(self._lib.gc_writer_set_threads(self._live(), int(threads)), None,
                "writer threads")

    def flush(self):
        """Waits until every queued snapshot is written; raises the first
        file error."""
        L.check(self._lib.gc_writer_flush(self._live()), None, "writer flush")


def output_steps(steps, snapshot_every):
    """Steps the reference's run loop writes outputs at
    (cpp/core/main.cpp:171-184): 0, every snapshot_every-th step and the
    last step."""
    if snapshot_every <= 0:
        raise ValueError("snapshot_every must be positive")
    return [0] + [s for s in range(1, steps + 1) if s % snapshot_every == 0 or s == steps]


def run_with_output(eng, out, steps, dt, snapshot_every, group_ids=None):
    """The reference's run loop (cpp/core/main.cpp:163-184) on a resident
    engine: outputs (snapshot + diagnostics row, time = step * dt) at the
    steps output_steps names, the KDK steps between two outputs queued as
    one call.  The engine must hold a(x) of its current state."""
    done = 0
    for s in output_steps(steps, snapshot_every):
        if s > done:
            eng.step(dt, s - done)
            done = s
        t = float(s) * dt
        out.write_snapshot(s, t, eng, group_ids=group_ids)
        out.write_diagnostics(s, t, eng)
