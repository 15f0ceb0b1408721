"""Many independent collisions advanced together (gc_batch_* ABI).

    b = Batch(counts, softening=0.05)
    b.upload(pos, vel, mass)        # runs concatenated in order
    b.autotune()                    # pair form and its schedule
    b.compute_accelerations()
    b.step(1e-3, 500)               # hipGraph chunks of graph_steps steps
    pos, vel, acc = b.download()
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .engine import PAIRS, make_params

_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


class Batch:
    def __init__(self, counts, softening=0.05, G=1.0, device=0):
        self._lib = L.load()
        self.counts = [int(c) for c in counts]
        self.params = make_params("hip-direct", "mixed", G=G, softening=softening)
        arr = (C.c_size_t * len(self.counts))(*self.counts)
        self._b = C.c_void_p()
        st = self._lib.gc_batch_create(C.byref(self._b), device, C.byref(self.params),
                                       len(self.counts), arr)
        if st != L.GC_OK:
            msg = self._lib.gc_batch_last_error(None).decode(errors="replace")
            raise L.GcError(st, f"create batch: {msg}")
        self.n = sum(self.counts)

    def _check(self, st, what):
        if st != L.GC_OK:
            msg = self._lib.gc_batch_last_error(self._b).decode(errors="replace")
            raise L.GcError(st, f"{what}: {msg}")

    def close(self):
        if self._b:
            self._lib.gc_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, pos, vel, mass):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.n, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64).reshape(self.n)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(self.n, 3)
        self._check(self._lib.gc_batch_upload(self._b, _ptr(pos), None if v is None else _ptr(v),
                                              _ptr(mass)), "batch upload")

    def generate_collisions(self, runs):
        """Draw every run's two-disk collision on the device (runs: the
        dicts of generators.collisions, in batch order; each run's n must
        equal its particle count here)."""
        from .generators import collision_galaxies

        gals = (L.DiskGalaxy * (2 * len(runs)))()
        seeds = (C.c_uint64 * len(runs))()
        for r, run in enumerate(runs):
            pair = collision_galaxies(run)
            gals[2 * r], gals[2 * r + 1] = pair[0], pair[1]
            seeds[r] = int(run["seed"])
        self._check(self._lib.gc_batch_generate_disks(self._b, gals, 2, seeds),
                    "batch generate collisions")

    def download(self):
        pos, vel, acc = (np.empty((self.n, 3)) for _ in range(3))
        self._check(self._lib.gc_batch_download(self._b, _ptr(pos), _ptr(vel), _ptr(acc)),
                    "batch download")
        return pos, vel, acc

    def compute_accelerations(self):
        self._check(self._lib.gc_batch_compute_accelerations(self._b), "batch accelerations")

    def step(self, dt, steps=1):
        self._check(self._lib.gc_batch_step(self._b, float(dt), int(steps)), "batch step")

    def synchronize(self):
        self._check(self._lib.gc_batch_synchronize(self._b), "batch synchronize")

    def set_schedule(self, chunk_tiles, blocks, graph_steps=10):
        self._check(self._lib.gc_batch_set_schedule(self._b, chunk_tiles, blocks, graph_steps),
                    "batch schedule")

    def schedule(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._check(self._lib.gc_batch_get_schedule(self._b, C.byref(a), C.byref(b), C.byref(c)),
                    "batch schedule")
        return a.value, b.value, c.value

    def set_pairs(self, mode="auto", targets_per_lane=0):
        """Pair coverage within each run: "auto", "targets" (scheduled
        per-target kernel) or "symmetric" (each unordered pair once)."""
        self._check(self._lib.gc_batch_set_pairs(self._b, PAIRS[mode], targets_per_lane),
                    "batch pairs")

    def pairs(self):
        """(form in use, targets per lane)."""
        m, t = C.c_int(), C.c_int()
        self._check(self._lib.gc_batch_get_pairs(self._b, C.byref(m), C.byref(t)), "batch pairs")
        return {L.GC_PAIRS_TARGETS: "targets", L.GC_PAIRS_SYMMETRIC: "symmetric"}[m.value], t.value

    def autotune(self):
        a, b = C.c_int(), C.c_int()
        self._check(self._lib.gc_batch_autotune(self._b, C.byref(a), C.byref(b)), "autotune")
        return a.value, b.value

    def time_force_kernel(self, reps=10):
        ms, inter = C.c_double(), C.c_double()
        self._check(self._lib.gc_batch_time_force_kernel(self._b, reps, C.byref(ms),
                                                         C.byref(inter)), "batch timing")
        return ms.value, inter.value

    def split(self, arr):
        """Per-run views of a concatenated (n, ...) array."""
        out, o = [], 0
        for c in self.counts:
            out.append(arr[o:o + c])
            o += c
        return out
