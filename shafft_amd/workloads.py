"""Synthetic workloads of the batched and served forms.

`draw_shapes` is the batched uneven workload: 2-D and 3-D transforms
(probability 1/2 each), every axis drawn log-uniformly (weight 1/n) from the
7-smooth lengths in [lo, hi] (16..256 gives the 64 lengths 16, 18, 20, 21,
24, ..., 245, 250, 252, 256).

`draw_stream` is the mixed serving stream: Poisson arrivals, sizes
log-uniform between 2^lo and 2^hi elements (rounded to the nearest 7-smooth
size), 70% complex64 / 30% complex128, forward or backward with equal odds,
half FFTND (2-D or 3-D, balanced 7-smooth factorisation of the size) and
half FFT1D.  The seed fixes the whole stream.
"""
import bisect
import math

import numpy as np


def smooth7_between(lo: int, hi: int):
    """Sorted 7-smooth integers in [lo, hi]."""
    out = []
    p2 = 1
    while p2 <= hi:
        p3 = p2
        while p3 <= hi:
            p5 = p3
            while p5 <= hi:
                p7 = p5
                while p7 <= hi:
                    if p7 >= lo:
                        out.append(p7)
                    p7 *= 7
                p5 *= 5
            p3 *= 3
        p2 *= 2
    return sorted(out)


def factor7(n: int):
    f = []
    for p in (7, 5, 3, 2):
        while n % p == 0:
            f.append(p)
            n //= p
    assert n == 1
    return f


def balanced_dims(size: int, ndim: int):
    """Split a 7-smooth size into ndim axis lengths of similar magnitude
    (largest prime factors placed first, each on the currently smallest axis)."""
    dims = [1] * ndim
    for p in factor7(size):
        i = min(range(ndim), key=lambda k: dims[k])
        dims[i] *= p
    return sorted(dims, reverse=True)


class Request:
    __slots__ = ("kind", "dims", "n", "prec", "direction", "t")

    def __init__(self, kind, dims, n, prec, direction, t):
        self.kind, self.dims, self.n, self.prec, self.direction, self.t = kind, dims, n, prec, direction, t

    @property
    def elems(self) -> int:
        return self.n if self.kind == 1 else int(np.prod(self.dims))

    @property
    def gflop(self) -> float:
        n = self.elems
        return 5.0 * n * math.log2(n) * 1e-9 if n > 1 else 0.0

    def describe(self) -> str:
        shape = f"n={self.n}" if self.kind == 1 else "x".join(map(str, self.dims))
        return f"{'1d' if self.kind else 'nd'}:{shape}:{'z2z' if self.prec else 'c2c'}:{'bwd' if self.direction else 'fwd'}"


def draw_stream(count: int, seed: int = 46, lo: int = 10, hi: int = 26, rate: float = 2000.0,
                p_c128: float = 0.3):
    rng = np.random.default_rng(seed)
    sizes = smooth7_between(1 << lo, 1 << hi)
    logs = [math.log(s) for s in sizes]
    t = 0.0
    out = []
    for _ in range(count):
        t += rng.exponential(1.0 / rate)
        u = rng.uniform(lo * math.log(2), hi * math.log(2))
        k = bisect.bisect_left(logs, u)
        if k == len(logs) or (k > 0 and u - logs[k - 1] < logs[k] - u):
            k -= 1
        size = sizes[k]
        prec = 1 if rng.random() < p_c128 else 0
        direction = int(rng.random() < 0.5)
        if rng.random() < 0.5:
            ndim = 2 if rng.random() < 0.5 else 3
            out.append(Request(0, balanced_dims(size, ndim), 0, prec, direction, t))
        else:
            out.append(Request(1, None, size, prec, direction, t))
    return out


def smooth7(lo=16, hi=256):
    out = []
    for n in range(lo, hi + 1):
        m = n
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        if m == 1:
            out.append(n)
    return out


def draw_shapes(count, seed, lo=16, hi=256):
    rng = np.random.default_rng(seed)
    lens = np.array(smooth7(lo, hi))
    w = 1.0 / lens
    w = w / w.sum()
    shapes = []
    for _ in range(count):
        nd = 2 if rng.random() < 0.5 else 3
        shapes.append(tuple(int(v) for v in rng.choice(lens, size=nd, p=w)))
    return shapes
