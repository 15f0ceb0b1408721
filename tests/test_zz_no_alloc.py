"""gc_direct_accelerations_noalloc on the GPU, through the C ABI.

The entry point allocates nothing: it runs the per-target mixed kernels of
gc_compute_accelerations on caller-owned device arrays with a static device
scratch array.  Two calls in a row (the second smaller, on the scratch the
first left behind) must give accelerations bit-identical to a context on
the same pair form and split count, and within the mixed path's bound
(max relative force error <= 1e-5) of the FP64 oracle.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from gcsim import _lib as L
from helpers import rel_err

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-5  # max relative force error, mixed path vs FP64
MAX_N = 262144  # GC_NOALLOC_MAX_PARTICLES
PART_FLOATS = 3145728  # GC_NOALLOC_PARTIAL_FLOATS


def _splits(n):
    """The split count the header documents for the entry point."""
    ntiles = -(-n // 256)
    js = min(64, -(-16384 // ntiles), ntiles)
    js = max(1, min(js, PART_FLOATS // (6 * 256 * ntiles)))
    tps = -(-ntiles // js)
    return -(-ntiles // tps)


def _noalloc(gpu, pos, mass, eps, G=1.0):
    import torch

    lib = L.load()
    n = len(mass)
    dev = torch.device("cuda", 0)
    p = torch.from_numpy(np.ascontiguousarray(pos.T)).to(dev)  # [3][n]
    m = torch.from_numpy(np.ascontiguousarray(mass)).to(dev)
    a = torch.full((3, n), float("nan"), dtype=torch.float64, device=dev)
    params = gpu.make_params("hip-direct", "mixed", G=G, softening=eps)
    stream = torch.cuda.current_stream(dev).cuda_stream
    st = lib.gc_direct_accelerations_noalloc(C.byref(params), n, C.c_void_p(p.data_ptr()),
                                             C.c_void_p(m.data_ptr()), C.c_void_p(a.data_ptr()),
                                             C.c_void_p(stream))
    L.check(st, None, "direct_accelerations_noalloc")
    torch.cuda.synchronize(dev)
    return a.cpu().numpy().T.copy()


def _context(gpu, pos, mass, eps, precision, G=1.0, splits=None):
    with gpu.Engine("hip-direct", precision, G=G, softening=eps) as eng:
        if splits:
            eng.set_pairs("targets")
            eng.set_sources("scalar")
            eng.set_launch_shape(1, splits)
        eng.upload(pos, None, mass)
        eng.compute_accelerations()
        return eng.download()[2]


def test_two_calls_match_context_and_oracle(gpu):
    big_pos, _, big_m = gpu.generators.two_disk(10001, softening=0.02, seed=7, disk_mass=0.1)
    small_pos, _, small_m = gpu.generators.config0()
    for name, pos, m, eps in (("n=20002", big_pos, big_m, 0.02),
                              ("n=1000", small_pos, small_m, 0.05)):
        print(f"noalloc call {name}", flush=True)
        got = _noalloc(gpu, pos, m, eps)
        want = _context(gpu, pos, m, eps, "mixed", splits=_splits(len(m)))
        assert np.array_equal(got, want), name
        ref = O.direct_targets(pos, m, 1.0, eps) if len(m) <= 4096 else \
            _context(gpu, pos, m, eps, "fp64")
        assert rel_err(got, ref).max() <= MIXED_TOL, name


def test_largest_size_and_limits(gpu):
    import torch

    pos, _, m = gpu.generators.two_disk(MAX_N // 2, softening=0.02, seed=3, disk_mass=0.1)
    print("noalloc call n=262144", flush=True)
    got = _noalloc(gpu, pos, m, 0.02)
    want = _context(gpu, pos, m, 0.02, "mixed", splits=_splits(MAX_N))
    assert np.array_equal(got, want)
    # a second, smaller call after the largest one
    print("noalloc call n=513", flush=True)
    got = _noalloc(gpu, pos[:513], m[:513], 0.02)
    assert rel_err(got, O.direct_targets(pos[:513], m[:513], 1.0, 0.02)).max() <= MIXED_TOL
    lib = L.load()
    params = gpu.make_params("hip-direct", "mixed", softening=0.02)
    # beyond the maximum: rejected on the host before any launch (the arrays
    # are never touched, so small ones suffice)
    buf = torch.zeros(3, dtype=torch.float64, device="cuda")
    ptr = C.c_void_p(buf.data_ptr())
    st = lib.gc_direct_accelerations_noalloc(C.byref(params), MAX_N + 1, ptr, ptr, ptr, None)
    assert st == L.GC_ERR_INVALID_ARGUMENT
    assert b"GC_NOALLOC_MAX_PARTICLES" in lib.gc_last_error(None)
    # fp64 precision has no scratch to spare: rejected
    p64 = gpu.make_params("hip-direct", "fp64", softening=0.02)
    assert lib.gc_direct_accelerations_noalloc(C.byref(p64), 1, ptr, ptr, ptr, None) == \
        L.GC_ERR_INVALID_ARGUMENT
    # n = 0 is a no-op
    assert lib.gc_direct_accelerations_noalloc(C.byref(params), 0, None, None, None, None) == L.GC_OK
