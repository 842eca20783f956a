"""Pins of the oracle's Omega streams (cpp/src/anchor_select_randomized_id.cpp:269-279).

The reference fills u[a*ell + s] and v[a*ell + s] for i = a*ell + s = 0 .. d*ell-1
with std::normal_distribution<double>(0, 1) drawn over std::mt19937_64(seed)
and std::mt19937_64(seed ^ 0x9e3779b97f4a7c15).  The oracle calls libstdc++
for both.  Here every piece is restated independently in Python from its
published definition and checked against the oracle bit for bit:

* MT19937-64 from Matsumoto and Nishimura's 64-bit Mersenne Twister, with the
  parameters of std::mt19937_64 ([rand.predef]); pinned by the value the C++
  standard requires of it (the 10000th draw of a default-seeded engine is
  9981545732273789042);
* the uniform double of std::generate_canonical<double, 53> over a 64-bit
  engine: one draw, converted to double and divided by 2^64, with 1.0 mapped
  to the largest double below 1;
* Marsaglia's polar method as libstdc++ documents it: x, y = 2U - 1, retry
  while r2 = x^2 + y^2 is > 1 or 0, mult = sqrt(-2 log(r2) / r2), return
  y * mult and hold x * mult for the next call.

A wrong seed mix, a generator shared by u and v, the two polar values
swapped, or a transposed fill each fail one of these.
"""
import math

import numpy as np
import pytest

import oracle

MASK64 = (1 << 64) - 1


class MT19937_64:
    """The 64-bit Mersenne Twister (w=64, n=312, m=156, r=31)."""

    N, M = 312, 156
    A = 0xB5026F5AA96619E9
    UPPER, LOWER = 0xFFFFFFFF80000000, 0x7FFFFFFF

    def __init__(self, seed=5489):
        mt = [seed & MASK64]
        for i in range(1, self.N):
            prev = mt[-1]
            mt.append((6364136223846793005 * (prev ^ (prev >> 62)) + i) & MASK64)
        self.mt, self.index = mt, self.N

    def _twist(self):
        mt, n, m = self.mt, self.N, self.M
        for i in range(n):
            x = (mt[i] & self.UPPER) | (mt[(i + 1) % n] & self.LOWER)
            xa = x >> 1
            if x & 1:
                xa ^= self.A
            mt[i] = mt[(i + m) % n] ^ xa
        self.index = 0

    def __call__(self):
        if self.index >= self.N:
            self._twist()
        y = self.mt[self.index]
        self.index += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000 & MASK64
        y ^= (y << 37) & 0xFFF7EEE000000000 & MASK64
        y ^= y >> 43
        return y


def canonical(engine):
    """std::generate_canonical<double, 53> over a full-range 64-bit engine."""
    u = float(engine()) / 18446744073709551616.0  # int -> double rounds to nearest, / 2^64 is exact
    return u if u < 1.0 else math.nextafter(1.0, 0.0)


class PolarNormal:
    """Marsaglia's polar method with one held value, as libstdc++ draws
    std::normal_distribution<double>(0, 1)."""

    def __init__(self):
        self.saved = None

    def __call__(self, engine):
        if self.saved is not None:
            out, self.saved = self.saved, None
            return out
        while True:
            x = 2.0 * canonical(engine) - 1.0
            y = 2.0 * canonical(engine) - 1.0
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0.0):
                break
        mult = math.sqrt(-2.0 * math.log(r2) / r2)
        self.saved = x * mult
        return y * mult


def restated_omega(d, ell, seed):
    gu, gv = MT19937_64(seed), MT19937_64(seed ^ 0x9E3779B97F4A7C15)
    nu, nv = PolarNormal(), PolarNormal()
    u = np.empty(d * ell)
    v = np.empty(d * ell)
    for i in range(d * ell):  # one draw from each stream per flat index a*ell + s
        u[i] = nu(gu)
        v[i] = nv(gv)
    return u.reshape(d, ell), v.reshape(d, ell)


def test_mt19937_64_standard_value():
    """[rand.predef]: the 10000th consecutive invocation of a default-constructed
    mt19937_64 produces 9981545732273789042."""
    engine = MT19937_64()
    for _ in range(9999):
        engine()
    assert engine() == 9981545732273789042


def test_mt19937_64_seed_initialisation():
    """A seeded engine differs from the default one from its first draw, and equal
    seeds give equal streams (the seed enters through the published
    initialisation recurrence, not as an offset)."""
    a, b, c = MT19937_64(7), MT19937_64(7), MT19937_64()
    first = [a() for _ in range(5)]
    assert first == [b() for _ in range(5)]
    assert first != [c() for _ in range(5)]


def test_canonical_maps_the_top_of_the_range_below_one():
    class Top:
        def __call__(self):
            return MASK64
    assert canonical(Top()) == math.nextafter(1.0, 0.0)
    class Zero:
        def __call__(self):
            return 0
    assert canonical(Zero()) == 0.0


@pytest.mark.parametrize("d, ell, seed", [(7, 13, 7), (3, 5, 0), (2, 9, 2**63 + 5), (16, 40, 31)])
def test_oracle_omega_equals_independent_restatement(d, ell, seed):
    u, v = restated_omega(d, ell, seed)
    ou, ov = oracle.omega(d, ell, seed)
    np.testing.assert_array_equal(ou, u)
    np.testing.assert_array_equal(ov, v)


def test_restatement_catches_plausible_mistakes():
    """Each plausible slip in the Omega generation changes the values, so the
    bitwise comparison above pins the seed mix, the separate streams, the
    polar method's value order and the row-major fill."""
    d, ell, seed = 4, 6, 7
    ou, ov = oracle.omega(d, ell, seed)
    # v on the unmixed seed (the mix constant dropped)
    nv, gv = PolarNormal(), MT19937_64(seed)
    assert not np.array_equal(ov.ravel(), [nv(gv) for _ in range(d * ell)])
    # u and v drawn alternately from one shared generator
    g, nu, nv = MT19937_64(seed), PolarNormal(), PolarNormal()
    shared = [(nu(g), nv(g)) for _ in range(d * ell)]
    assert not np.array_equal(ou.ravel(), [p[0] for p in shared])
    # the held value returned first
    class Swapped(PolarNormal):
        def __call__(self, engine):
            if self.saved is not None:
                out, self.saved = self.saved, None
                return out
            first = super().__call__(engine)
            self.saved, first = first, self.saved
            return first
    su, gu = Swapped(), MT19937_64(seed)
    assert not np.array_equal(ou.ravel(), [su(gu) for _ in range(d * ell)])
    # column-major fill u[s*d + a]
    ru, _ = restated_omega(d, ell, seed)
    assert not np.array_equal(ou, ru.ravel().reshape(ell, d).T)
