"""Pure-Python restatements of the C++ standard-library random-number objects
the reference's -r generator and -c colouring draw from (graph.cpp:76-110,
547-571), as libstdc++ 11 (the C++ library of this image, GCC 11.4)
implements them:

* std::mt19937_64 -- Matsumoto & Nishimura's 64-bit Mersenne Twister with the
  parameters fixed by the C++ standard ([rand.predef]).
* std::minstd_rand0 (libstdc++'s std::default_random_engine) -- Park &
  Miller's x <- 16807 x mod (2^31 - 1).
* std::uniform_int_distribution -- for an engine of range 2^64 - 1, Lemire's
  nearly divisionless method (ACM TOMACS 29(1), 2019) with a 128-bit
  product; for other engines wider than the interval, the "divide and
  reject" downscaling (bits/uniform_int_dist.h).
* std::generate_canonical<double, 53> -- ceil(53 / floor(log2 R)) engine
  draws (R = max - min + 1) summed in base R, divided by R^k
  (bits/random.tcc).
* std::normal_distribution<double> -- Marsaglia's polar method, the second
  value of each pair kept for the next call (bits/random.tcc).

Checked against the C++ standard's required values (10000th output of the
default-seeded engines) in tests/test_coloring.py.
"""
from __future__ import annotations

import math

_M64 = (1 << 64) - 1


class MT19937_64:
    n, m = 312, 156
    a = 0xB5026F5AA96619E9
    u, d = 29, 0x5555555555555555
    s, b = 17, 0x71D67FFFEDA60000
    t, c = 37, 0xFFF7EEE000000000
    l = 43
    f = 6364136223846793005
    lower = (1 << 31) - 1
    upper = _M64 & ~lower
    min, max = 0, _M64

    def __init__(self, seed: int = 5489):
        x = [0] * self.n
        x[0] = seed & _M64
        for i in range(1, self.n):
            x[i] = (self.f * (x[i - 1] ^ (x[i - 1] >> 62)) + i) & _M64
        self.x = x
        self.i = self.n

    def _twist(self):
        x, n, m = self.x, self.n, self.m
        for k in range(n):
            y = (x[k] & self.upper) | (x[(k + 1) % n] & self.lower)
            z = x[(k + m) % n] ^ (y >> 1)
            if y & 1:
                z ^= self.a
            x[k] = z
        self.i = 0

    def __call__(self) -> int:
        if self.i >= self.n:
            self._twist()
        y = self.x[self.i]
        self.i += 1
        y ^= (y >> self.u) & self.d
        y ^= (y << self.s) & self.b & _M64
        y ^= (y << self.t) & self.c & _M64
        y ^= y >> self.l
        return y & _M64


class MinStdRand0:
    min, max = 1, 2147483646

    def __init__(self, seed: int = 1):
        self.x = seed % 2147483647 or 1

    def __call__(self) -> int:
        self.x = (16807 * self.x) % 2147483647
        return self.x


def uniform_int(g, a: int, b: int) -> int:
    """std::uniform_int_distribution<int64_t>(a, b)(g), b - a below the
    engine's range."""
    urange = b - a
    urngrange = g.max - g.min
    assert urngrange > urange
    if urngrange == _M64:                        # Lemire, 128-bit product
        r = urange + 1
        prod = g() * r
        low = prod & _M64
        if low < r:
            thr = ((1 << 64) - r) % r            # -r % r in 64-bit arithmetic
            while low < thr:
                prod = g() * r
                low = prod & _M64
        return (prod >> 64) + a
    uerange = urange + 1
    scaling = urngrange // uerange
    past = uerange * scaling
    while True:
        ret = g() - g.min
        if ret < past:
            break
    return ret // scaling + a


def generate_canonical(g) -> float:
    r = g.max - g.min + 1
    log2r = int(math.log(r) / math.log(2.0))
    k = max(1, (53 + log2r - 1) // log2r)
    s, t = 0.0, 1.0
    for _ in range(k):
        s += float(g() - g.min) * t
        t = float(int(t) * r)          # long double product, exact, rounded once to double
    ret = s / t
    if ret >= 1.0:
        ret = math.nextafter(1.0, 0.0)
    return ret


class Normal:
    def __init__(self, mean: float = 0.0, stddev: float = 1.0):
        self.mean, self.stddev = mean, stddev
        self.saved = None

    def __call__(self, g) -> float:
        if self.saved is not None:
            ret, self.saved = self.saved, None
        else:
            while True:
                x = 2.0 * generate_canonical(g) - 1.0
                y = 2.0 * generate_canonical(g) - 1.0
                r2 = x * x + y * y
                if not (r2 > 1.0 or r2 == 0.0):
                    break
            mult = math.sqrt(-2 * math.log(r2) / r2)
            self.saved = x * mult
            ret = y * mult
        return ret * self.stddev + self.mean


def coloring_weights(nv: int, engines: int = 1):
    """The random weights of Graph::coloring (graph.cpp:558-571): engine t =
    mt19937_64(1 << t), vertex i draws uniform_int(0, 4 nv) from engine
    i % engines, each engine in vertex order."""
    eng = [MT19937_64(1 << t) for t in range(engines)]
    return [uniform_int(eng[i % engines], 0, 4 * nv) for i in range(nv)]


def ba_edges(nv: int, m0: int):
    """The (tail, head, weight) records of create_random_network_ba
    (graph.cpp:76-105), in draw order."""
    g1, g2 = MinStdRand0(), MinStdRand0()
    gauss = Normal(1.0, 1.0)
    el = []
    for i in range(m0):
        w = abs(gauss(g1))
        el.append((0, i + 1, w))
        el.append((i + 1, 0, w))
    for i in range(m0 + 1, nv):
        total = len(el)
        new = []
        for _ in range(m0):
            x = el[uniform_int(g2, 0, total - 1)][0]
            w = abs(gauss(g1))
            new.append((i, x, w))
            new.append((x, i, w))
        el.extend(new)
    return el
