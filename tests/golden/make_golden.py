"""Writes the golden DFT fixtures in tests/golden/*.npz.

The reference ships no fixtures (its tests compute a naive DFT in-test), so
these vectors are produced here with mpmath at 40 significant digits,
independently of the oracle's C code, on the reference's own test inputs:
  * dft_partial cases, input sin(0.01 i) + i cos(0.02 i)
    (tests/unit/test_fft_partial.cpp:35-40, cases :231-237);
  * small full-DFT cases of tests/unit/test_fft_highdim.cpp:189-201 shapes
    (5-D {2,2,2,2,2}, 7-D all-2, prime 3x5x7);
  * one random complex input per precision on {4,6,5} with std::mt19937_64
    seed 42 (re then im, 2*(x>>11)*2^-53 - 1), generated here in pure Python.
Run: python tests/golden/make_golden.py
"""
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))


def mt19937_64(seed):
    """Pure-Python std::mt19937_64 (independent of oracle/dft_oracle.c)."""
    n, m = 312, 156
    mt = [0] * n
    mt[0] = seed & 0xFFFFFFFFFFFFFFFF
    for i in range(1, n):
        mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & 0xFFFFFFFFFFFFFFFF
    idx = n
    while True:
        if idx >= n:
            for i in range(n):
                x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % n] & 0x7FFFFFFF)
                xa = x >> 1
                if x & 1:
                    xa ^= 0xB5026F5AA96619E9
                mt[i] = mt[(i + m) % n] ^ xa
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        yield y & 0xFFFFFFFFFFFFFFFF


def uniform(shape, seed):
    g = mt19937_64(seed)
    n = int(np.prod(shape))
    v = np.empty(2 * n)
    for i in range(2 * n):
        v[i] = 2.0 * ((next(g) >> 11) * 2.0 ** -53) - 1.0
    return (v[0::2] + 1j * v[1::2]).reshape(shape)


def mp_dft_axes(x, axes, sign):
    """Direct DFT over `axes` in mpmath, returned as complex128."""
    a = np.empty(x.shape, dtype=object)
    for idx in np.ndindex(x.shape):
        a[idx] = mp.mpc(x[idx].real, x[idx].imag)
    for ax in axes:
        n = x.shape[ax]
        w = [mp.expjpi(sign * mp.mpf(2 * m) / n) for m in range(n)]
        b = np.empty(x.shape, dtype=object)
        for idx in np.ndindex(x.shape):
            k = idx[ax]
            s = mp.mpc(0)
            for j in range(n):
                src = list(idx)
                src[ax] = j
                s += a[tuple(src)] * w[(j * k) % n]
            b[idx] = s
        a = b
    out = np.empty(x.shape, np.complex128)
    for idx in np.ndindex(x.shape):
        out[idx] = complex(a[idx])
    return out


def main():
    cases = {}
    partial = [((2, 3, 4, 5), (0, 1, 2)), ((4, 3, 2, 3, 2), (1, 2, 3)), ((2, 3, 4, 5), (2, 3)),
               ((2, 3, 4, 2), (0, 1, 2, 3)), ((3, 2, 2, 2, 2), (1, 2, 3, 4))]
    for ci, (dims, axes) in enumerate(partial):
        i = np.arange(int(np.prod(dims)))
        x = (np.sin(0.01 * i) + 1j * np.cos(0.02 * i)).reshape(dims)
        for sign in (-1, 1):
            cases[f"partial{ci}_{'fwd' if sign < 0 else 'bwd'}"] = (x, axes, sign)
    for ci, dims in enumerate([(2, 2, 2, 2, 2), (2,) * 7, (3, 5, 7)]):
        i = np.arange(int(np.prod(dims)))
        x = (np.sin(0.01 * i) + 1j * np.cos(0.02 * i)).reshape(dims)
        cases[f"full{ci}_fwd"] = (x, tuple(range(len(dims))), -1)
    x = uniform((4, 6, 5), 42)
    cases["mt42_fwd"] = (x, (0, 1, 2), -1)
    cases["mt42_bwd"] = (x, (0, 1, 2), 1)
    out = {}
    for name, (x, axes, sign) in cases.items():
        out[name + "__in"] = x
        out[name + "__axes"] = np.array(axes, dtype=np.int64)
        out[name + "__sign"] = np.array(sign, dtype=np.int64)
        out[name + "__out"] = mp_dft_axes(x, axes, sign)
        print(name, x.shape, axes, sign)
    np.savez_compressed(os.path.join(HERE, "dft_golden.npz"), **out)


if __name__ == "__main__":
    main()
