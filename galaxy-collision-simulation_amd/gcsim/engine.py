"""Python view of the gcsim engine: a resident-state force solver on one
gfx950 device, selected like the reference's `solver = ...` config value.

    eng = Engine(solver="hip-direct", precision="mixed", softening=0.02)
    eng.upload(pos, vel, mass)          # numpy (n, 3), (n, 3), (n,)
    eng.compute_accelerations()         # a(x)
    eng.step(dt=5e-4, steps=10)         # KDK, state stays in HBM
    pos, vel, acc = eng.download()

Mirrors the reference's solver selection (cpp/core/main.cpp:62-139) and
FmmOptions (cpp/fmm/fmm_solver.hpp:13-18).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

_dp = C.POINTER(C.c_double)

SOLVERS = {"hip-direct": L.GC_SOLVER_DIRECT, "hip": L.GC_SOLVER_DIRECT,
           "rocm-direct": L.GC_SOLVER_DIRECT, "hip-fmm": L.GC_SOLVER_FMM,
           "rocm-fmm": L.GC_SOLVER_FMM, "hip-tree": L.GC_SOLVER_TREE,
           "rocm-tree": L.GC_SOLVER_TREE}
PAIRS = {"auto": L.GC_PAIRS_AUTO, "targets": L.GC_PAIRS_TARGETS, "symmetric": L.GC_PAIRS_SYMMETRIC}
SOURCES = {"auto": L.GC_SOURCES_AUTO, "lds": L.GC_SOURCES_LDS, "scalar": L.GC_SOURCES_SCALAR}
PRECISIONS = {"fp64": L.GC_PRECISION_FP64, "double": L.GC_PRECISION_FP64,
              "mixed": L.GC_PRECISION_MIXED, "fp32": L.GC_PRECISION_MIXED,
              "fp64-fast": L.GC_PRECISION_FP64_FAST}


def _c64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _ptr(a):
    return a.ctypes.data_as(_dp)


def device_count() -> int:
    return int(L.load().gc_device_count())


def solver_available() -> bool:
    return bool(L.load().gc_solver_available())


def make_params(solver="hip-direct", precision="fp64", G=1.0, softening=0.02, theta=0.6,
                leaf_capacity=16, max_depth=32, expansion_order=4) -> L.Params:
    key = solver.lower()
    if key not in SOLVERS:
        raise ValueError(f"Unknown solver '{solver}'. Use hip-direct, hip-fmm or hip-tree.")
    pkey = precision.lower()
    if pkey not in PRECISIONS:
        raise ValueError(f"Unknown precision '{precision}'. Use fp64, fp64-fast or mixed.")
    p = L.Params()
    L.load().gc_default_params(C.byref(p))
    p.gravitational_constant = G
    p.softening = softening
    p.solver = SOLVERS[key]
    p.precision = PRECISIONS[pkey]
    p.theta = theta
    p.leaf_capacity = int(leaf_capacity)
    p.max_depth = int(max_depth)
    p.expansion_order = int(expansion_order)
    return p


class Engine:
    def __init__(self, solver="hip-direct", precision="fp64", G=1.0, softening=0.02,
                 theta=0.6, leaf_capacity=16, max_depth=32, expansion_order=4, device=0):
        self._lib = L.load()
        self.params = make_params(solver, precision, G, softening, theta, leaf_capacity,
                                  max_depth, expansion_order)
        self._ctx = C.c_void_p()
        L.check(self._lib.gc_create(C.byref(self._ctx), device, C.byref(self.params)), None,
                "create engine")

    # -- lifecycle
    def close(self):
        if self._ctx:
            self._lib.gc_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, status, what):
        L.check(status, self._ctx, what)

    # -- configuration
    def set_params(self, **kw):
        cur = {f: getattr(self.params, f) for f, _ in L.Params._fields_}
        solver = kw.pop("solver", None)
        precision = kw.pop("precision", None)
        cur.update(kw)
        p = L.Params(**cur)
        if solver is not None:
            p.solver = SOLVERS[solver.lower()]
        if precision is not None:
            p.precision = PRECISIONS[precision.lower()]
        self._check(self._lib.gc_set_params(self._ctx, C.byref(p)), "set params")
        self.params = p

    def set_stream(self, stream_handle: int | None):
        """Run on the given HIP stream (0 = the null stream, e.g. torch's
        default stream); None returns to the engine's own stream."""
        if stream_handle is None:
            self._check(self._lib.gc_use_own_stream(self._ctx), "use own stream")
        else:
            self._check(self._lib.gc_set_stream(self._ctx, C.c_void_p(stream_handle)),
                        "set stream")

    def set_launch_shape(self, targets_per_thread=0, jsplit=0):
        self._check(self._lib.gc_set_launch_shape(self._ctx, targets_per_thread, jsplit),
                    "set launch shape")

    def launch_shape(self):
        a, b = C.c_int(), C.c_int()
        self._check(self._lib.gc_get_launch_shape(self._ctx, C.byref(a), C.byref(b)),
                    "get launch shape")
        return a.value, b.value

    def set_sources(self, mode="auto"):
        """Mixed-kernel source reads: "auto", "lds" (block-shared LDS tiles) or
        "scalar" (wave-uniform scalar loads)."""
        self._check(self._lib.gc_set_sources(self._ctx, SOURCES[mode]), "set sources")

    def sources(self) -> str:
        return {L.GC_SOURCES_LDS: "lds",
                L.GC_SOURCES_SCALAR: "scalar"}[self._lib.gc_get_sources(self._ctx)]

    def set_pairs(self, mode="auto"):
        """Mixed full-range pair coverage: "auto", "targets" (every target over
        all sources) or "symmetric" (each unordered pair once)."""
        self._check(self._lib.gc_set_pairs(self._ctx, PAIRS[mode]), "set pairs")

    def set_pair_items(self, tiles=0):
        """Source tiles per work item of the symmetric form (0 = automatic)."""
        self._check(self._lib.gc_set_pair_items(self._ctx, int(tiles)), "set pair items")

    def set_graph_steps(self, steps=10):
        """Steps per hipGraph-captured chunk of step() (0 or 1: eager)."""
        self._check(self._lib.gc_set_graph_steps(self._ctx, int(steps)), "set graph steps")

    def graph_steps(self) -> int:
        return int(self._lib.gc_get_graph_steps(self._ctx))

    def pairs(self) -> str:
        return {L.GC_PAIRS_TARGETS: "targets",
                L.GC_PAIRS_SYMMETRIC: "symmetric"}[self._lib.gc_get_pairs(self._ctx)]

    # -- state
    @property
    def n(self) -> int:
        return int(self._lib.gc_particle_count(self._ctx))

    def upload(self, pos, vel, mass, acc=None):
        mass = _c64(mass)
        n = mass.shape[0]
        pos = _c64(pos, (n, 3))
        vel = None if vel is None else _c64(vel, (n, 3))
        self._check(self._lib.gc_upload(self._ctx, n, _ptr(pos),
                                        None if vel is None else _ptr(vel), _ptr(mass)),
                    "upload")
        if acc is not None:
            acc = _c64(acc, (n, 3))
            self._check(self._lib.gc_set_accelerations(self._ctx, _ptr(acc)),
                        "set accelerations")

    def download(self):
        n = self.n
        pos, vel, acc = (np.empty((n, 3)) for _ in range(3))
        self._check(self._lib.gc_download(self._ctx, _ptr(pos), _ptr(vel), _ptr(acc)),
                    "download")
        return pos, vel, acc

    def device_view(self) -> L.DeviceView:
        v = L.DeviceView()
        self._check(self._lib.gc_device_state(self._ctx, C.byref(v)), "device state")
        return v

    # -- the hot path
    def compute_accelerations(self, begin=None, end=None):
        if begin is None and end is None:
            self._check(self._lib.gc_compute_accelerations(self._ctx), "compute accelerations")
        else:
            b = 0 if begin is None else begin
            e = self.n if end is None else end
            self._check(self._lib.gc_compute_accelerations_for_targets(self._ctx, b, e),
                        "compute accelerations for targets")

    # -- initial conditions drawn on the device (counter-based Philox stream;
    #    the state is replaced, accelerations zeroed)
    def generate_reference_galaxies(self, galaxies, seed=42):
        """generate_disk_galaxy's model for each GalaxyConfig dict."""
        from .generators import galaxy_configs

        arr = galaxy_configs(galaxies)
        self._check(self._lib.gc_generate_galaxies_device(self._ctx, arr, len(galaxies), int(seed)),
                    "generate galaxies")

    def generate_two_disk(self, n_per_galaxy=500, seed=42, **galaxy_kw):
        """The two-disk collision (generators.two_disk_galaxies' keywords)."""
        from .generators import two_disk_galaxies

        gals = two_disk_galaxies(n_per_galaxy, **galaxy_kw)
        self._check(self._lib.gc_generate_disks_device(self._ctx, gals, 2, int(seed)),
                    "generate two-disk")

    def generate_plummer(self, n, total_mass=1.0, scale=1.0, seed=1):
        self._check(self._lib.gc_generate_plummer_device(self._ctx, int(n), float(total_mass),
                                                         float(scale), int(seed)),
                    "generate plummer")

    def step(self, dt, steps=1):
        self._check(self._lib.gc_step(self._ctx, float(dt), int(steps)), "step")

    def synchronize(self):
        self._check(self._lib.gc_synchronize(self._ctx), "synchronize")

    def diagnostics(self) -> dict:
        d = L.Diagnostics()
        self._check(self._lib.gc_diagnostics_compute(self._ctx, C.byref(d)), "diagnostics")
        return {
            "total_mass": d.total_mass, "kinetic_energy": d.kinetic_energy,
            "potential_energy": d.potential_energy, "total_energy": d.total_energy,
            "momentum_x": d.momentum[0], "momentum_y": d.momentum[1],
            "momentum_z": d.momentum[2], "center_of_mass_x": d.center_of_mass[0],
            "center_of_mass_y": d.center_of_mass[1], "center_of_mass_z": d.center_of_mass[2],
            "angular_momentum_x": d.angular_momentum[0],
            "angular_momentum_y": d.angular_momentum[1],
            "angular_momentum_z": d.angular_momentum[2],
        }

    # -- device-resident interop (addresses of caller-owned device arrays,
    #    FP64, SoA [3][n]; used by the autograd layer)
    def set_state_device(self, n, pos_ptr, vel_ptr, mass_ptr, acc_ptr=None):
        self._check(self._lib.gc_set_state_device(self._ctx, n, C.c_void_p(pos_ptr),
                                                  C.c_void_p(vel_ptr or 0),
                                                  C.c_void_p(mass_ptr),
                                                  C.c_void_p(acc_ptr or 0)), "set state device")

    def get_state_device(self, pos_ptr=None, vel_ptr=None, acc_ptr=None):
        self._check(self._lib.gc_get_state_device(self._ctx, C.c_void_p(pos_ptr or 0),
                                                  C.c_void_p(vel_ptr or 0),
                                                  C.c_void_p(acc_ptr or 0)), "get state device")

    def step_reverse(self, dt, steps=1):
        self._check(self._lib.gc_step_reverse(self._ctx, float(dt), int(steps)), "reverse step")

    def force_vjp(self, lambda_ptr, xbar_ptr, mbar_ptr, accumulate=False):
        self._check(self._lib.gc_force_vjp(self._ctx, C.c_void_p(lambda_ptr), C.c_void_p(xbar_ptr),
                                           C.c_void_p(mbar_ptr), 1 if accumulate else 0),
                    "force vjp")

    def fmm_stats(self) -> dict:
        """FmmStats of the last hip-fmm evaluation (fmm_solver.hpp:20-25) plus
        M2P/P2P work counts and build/evaluation times in ms."""
        s = L.FmmStats()
        self._check(self._lib.gc_fmm_last_stats(self._ctx, C.byref(s)), "fmm stats")
        return {k: getattr(s, k) for k, _ in L.FmmStats._fields_}

    def time_force_kernel(self, reps=20):
        """(ms per launch, interactions per launch) of the dominant force kernel."""
        ms, inter = C.c_double(), C.c_double()
        self._check(self._lib.gc_time_force_kernel(self._ctx, int(reps), C.byref(ms),
                                                   C.byref(inter)), "time force kernel")
        return ms.value, inter.value
