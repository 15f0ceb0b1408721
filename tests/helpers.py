"""Test helpers: a pure-Python std::mt19937_64 and the libstdc++
uniform_real_distribution<double> transform, to rebuild the reference
smoke test's particle sets exactly (cpp/tests/smoke_tests.cpp:83-95)."""
from __future__ import annotations

import numpy as np

_MASK = (1 << 64) - 1


class MT19937_64:
    """The 64-bit Mersenne Twister as specified by the C++ standard."""

    n, m = 312, 156
    a = 0xB5026F5AA96619E9
    u, d = 29, 0x5555555555555555
    s, b = 17, 0x71D67FFFEDA60000
    t, c = 37, 0xFFF7EEE000000000
    l = 43  # noqa: E741
    f = 6364136223846793005

    def __init__(self, seed: int = 5489):
        self.mt = [0] * self.n
        self.mt[0] = seed & _MASK
        for i in range(1, self.n):
            prev = self.mt[i - 1]
            self.mt[i] = (self.f * (prev ^ (prev >> 62)) + i) & _MASK
        self.idx = self.n

    def _twist(self):
        upper, lower = 0xFFFFFFFF80000000, 0x7FFFFFFF
        mt = self.mt
        for i in range(self.n):
            x = (mt[i] & upper) | (mt[(i + 1) % self.n] & lower)
            xa = x >> 1
            if x & 1:
                xa ^= self.a
            mt[i] = mt[(i + self.m) % self.n] ^ xa
        self.idx = 0

    def __call__(self) -> int:
        if self.idx >= self.n:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> self.u) & self.d
        y ^= (y << self.s) & self.b
        y ^= (y << self.t) & self.c
        y ^= y >> self.l
        return y & _MASK


def uniform_real(rng: MT19937_64, a: float, b: float) -> float:
    """libstdc++ uniform_real_distribution<double>: generate_canonical with
    one 64-bit draw, then u * (b - a) + a."""
    u = float(rng()) / 18446744073709551616.0
    if u >= 1.0:
        u = np.nextafter(1.0, 0.0)
    return u * (b - a) + a


def smoke_cube(n: int = 80, seed: int = 7):
    """The reference smoke test's particles: uniform in [-1, 1]^3, mass 1/n."""
    rng = MT19937_64(seed)
    pos = np.array([[uniform_real(rng, -1.0, 1.0) for _ in range(3)] for _ in range(n)])
    return pos, np.full(n, 1.0 / n)


def rel_err(a, ref):
    """Per-particle |a - a_ref| / max(|a_ref|, 1e-12) (docs/testing.md:55-60)."""
    return np.linalg.norm(a - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-12)


def pair_magnitude_sums(pos, mass, G, eps, chunk=512):
    """S_i = sum_j |G m_j (x_j - x_i) / (|x_j - x_i|^2 + eps^2)^1.5| (j != i).

    The FP32 pair terms of a mixed-precision sum carry rounding of order
    2^-24 |term| each, so |a_mixed - a| <= k 2^-24 S_i with k of order ten
    bounds the error whatever cancellation a_i itself has; rel_err alone
    does not, where a_i is much smaller than S_i.
    """
    pos = np.asarray(pos, dtype=np.float64)
    mass = np.asarray(mass, dtype=np.float64)
    n = len(mass)
    out = np.empty(n)
    for a in range(0, n, chunk):
        d = pos[None, :, :] - pos[a:a + chunk, None, :]
        s2 = (d * d).sum(axis=2) + eps * eps
        with np.errstate(divide="ignore", invalid="ignore"):
            w = np.where(s2 > 0, G * mass[None, :] * np.sqrt(s2) / (s2 * s2), 0.0)
        idx = np.arange(a, min(a + chunk, n))
        w[idx - a, idx] = 0.0
        out[a:a + chunk] = (w * np.sqrt((d * d).sum(axis=2))).sum(axis=1)
    return out
