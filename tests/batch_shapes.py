"""Shape generator for the batched uneven workload: 2-D and 3-D transforms
(probability 1/2 each), every axis drawn log-uniformly (weight 1/n) from the
7-smooth lengths in [lo, hi] (16..256 gives the 64 lengths 16, 18, 20, 21,
24, ..., 245, 250, 252, 256)."""
import numpy as np


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
