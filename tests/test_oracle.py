"""Pins the FP64 oracle before GPU results are checked against it.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign, a transposed operand, a wrong index) fails it:
golden vectors computed with mpmath, closed forms, numpy's pocketfft as an
independent library routine, the published std::mt19937_64 check value, and
planted mistakes that the error metric must reject.
"""
import math
import os

import numpy as np
import pytest

import oracle
from oracle import oracle as orc
from tests.closed_forms import chirp, chirp_spectrum

HERE = os.path.dirname(os.path.abspath(__file__))


def test_mt19937_64_published_check_value():
    # C++ [rand.predef]: the 10000th output of a default-constructed
    # std::mt19937_64 (seed 5489) is 9981545732273789042.
    assert int(orc.mt64_raw(10000, 5489)[-1]) == 9981545732273789042


def test_uniform_input_recipe():
    raw = orc.mt64_raw(4, 42)
    x = oracle.uniform_input((2,), 42)
    u = (raw >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    assert np.array_equal(x.real, 2 * u[0::2] - 1) and np.array_equal(x.imag, 2 * u[1::2] - 1)
    x32 = oracle.uniform_input((2,), 42, np.complex64)
    assert np.array_equal(x32, x.astype(np.complex64))


def _golden():
    z = np.load(os.path.join(HERE, "golden", "dft_golden.npz"), allow_pickle=False)
    names = sorted({k.split("__")[0] for k in z.files})
    return [(n, z[n + "__in"], tuple(z[n + "__axes"]), int(z[n + "__sign"]), z[n + "__out"]) for n in names]


@pytest.mark.parametrize("case", _golden(), ids=lambda c: c[0])
def test_oracle_matches_golden(case):
    name, x, axes, sign, ref = case
    y = oracle.dftn(x, axes, sign)
    assert np.max(np.abs(y - ref)) < 1e-13 * max(1.0, np.max(np.abs(ref)))


def test_golden_mt_input_matches_c_generator():
    z = np.load(os.path.join(HERE, "golden", "dft_golden.npz"), allow_pickle=False)
    assert np.array_equal(z["mt42_fwd__in"], oracle.uniform_input((4, 6, 5), 42))


@pytest.mark.parametrize("dims", [(8,), (7, 5), (6, 4, 3), (13, 2, 9), (16, 16)])
def test_oracle_vs_pocketfft(dims):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
    assert np.allclose(oracle.dftn(x, None, -1), np.fft.fftn(x), atol=1e-12)
    assert np.allclose(oracle.dftn(x, None, 1), np.fft.ifftn(x) * x.size, atol=1e-12)


def test_closed_forms():
    dims = (8, 6, 4)
    n = int(np.prod(dims))
    d = np.zeros(dims, complex)
    d[0, 0, 0] = 1
    assert np.allclose(oracle.dftn(d), 1, atol=1e-15)
    k = (2, 3, 1)
    idx = np.indices(dims)
    x = np.exp(2j * np.pi * sum(k[i] * idx[i] / dims[i] for i in range(3)))
    ref = np.zeros(dims, complex)
    ref[k] = n
    assert np.allclose(oracle.dftn(x, None, -1), ref, atol=1e-12)
    c = np.full(dims, 2 - 1j)
    ref = np.zeros(dims, complex)
    ref[0, 0, 0] = n * (2 - 1j)
    assert np.allclose(oracle.dftn(c), ref, atol=1e-12)
    rng = np.random.default_rng(0)
    r = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
    X = oracle.dftn(r)
    assert abs(np.vdot(X, X).real / n - np.vdot(r, r).real) < 1e-12 * np.vdot(r, r).real
    # shift theorem along axis 1: x[n - s] <-> X[k] exp(-2 pi i k s / N)
    s = 2
    Xs = oracle.dftn(np.roll(r, s, axis=1))
    ph = np.exp(-2j * np.pi * np.arange(dims[1]) * s / dims[1])[None, :, None]
    assert np.allclose(Xs, X * ph, atol=1e-12)


def test_fft_pow2_matches_direct():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((16, 8, 32)) + 1j * rng.standard_normal((16, 8, 32))
    for sign in (-1, 1):
        assert np.allclose(oracle.fftn_pow2(x, None, sign), oracle.dftn(x, None, sign), atol=1e-12)
    with pytest.raises(RuntimeError):
        oracle.fftn_pow2(np.zeros((6,), complex))


def test_norm_factors():
    assert oracle.norm_factor_nd((64, 64, 32), 2, 1) == 1.0 / 131072
    assert oracle.norm_factor_nd((64, 64, 32), 1, 1) == float(1 / np.sqrt(np.longdouble(131072)))
    assert oracle.norm_factor_nd((4,), 0, 0) == 1.0
    assert oracle.norm_factor_1d(256, 1, 1) == 1.0 / 16
    assert oracle.norm_factor_1d(100, 2, 0) == np.float32(0.01)


def test_error_metric_rejects_planted_mistakes():
    """Mistakes planted in the oracle's output must fail check_rel_error."""
    rng = np.random.default_rng(9)
    x = rng.standard_normal((6, 5, 4)) + 1j * rng.standard_normal((6, 5, 4))
    good = oracle.dftn(x)
    assert oracle.check_rel_error(good, np.fft.fftn(x), 1e-12)
    wrong_sign = oracle.dftn(x, None, +1)
    transposed = oracle.dftn(x.reshape(5, 6, 4)).reshape(6, 5, 4)  # axes read in the wrong order
    dropped_axis = oracle.dftn(x, (0, 1))
    off_by_one = np.roll(good, 1, axis=2)
    for bad in (wrong_sign, transposed, dropped_axis, off_by_one):
        assert not oracle.check_rel_error(bad, np.fft.fftn(x), 1e-5)
    # reference scaling: tolerance grows with sqrt(ln N)
    assert math.isclose(math.sqrt(math.log(120)), 2.188, rel_tol=1e-3)


def test_gaussian_input_recipe():
    """Box-Muller on the pinned raw stream, re-derived in numpy, and the
    moments of 2^20 draws (mean 0, variance 1 per component)."""
    raw = orc.mt64_raw(8, 44)
    u1 = ((raw[0::2] >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (raw[1::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    x = oracle.gaussian_input((4,), 44)
    assert np.allclose(x.real, r * np.cos(2 * np.pi * u2), rtol=1e-15, atol=0)
    assert np.allclose(x.imag, r * np.sin(2 * np.pi * u2), rtol=1e-15, atol=0)
    big = oracle.gaussian_input((1 << 20,), 44)
    for comp in (big.real, big.imag):
        assert abs(comp.mean()) < 5e-3 and abs(comp.var() - 1) < 5e-3
    assert abs(np.mean(big.real * big.imag)) < 5e-3
    assert np.array_equal(oracle.gaussian_input((3,), 44, np.complex64), x[:3].astype(np.complex64))


@pytest.mark.parametrize("n", [8, 64, 1000, 4096])
def test_chirp_closed_form(n):
    """For even n the DFT of exp(i pi m^2/n) is sqrt(n) e^{i pi/4}
    exp(-i pi k^2/n) (quadratic Gauss sum; the chirp is n-periodic for even
    n).  The GPU tests use this closed form as the all-bins check of the
    2^27-point FFT1D."""
    x = chirp(n)
    ref = chirp_spectrum(n)
    assert np.max(np.abs(oracle.dftn(x, None, -1) - ref)) < 1e-12 * np.sqrt(n)
    m = np.arange(n)
    assert np.max(np.abs(x - np.exp(1j * np.pi * ((m * m) % (2 * n)) / n))) < 1e-14
    # the chunked forms used on large n match the whole-array forms
    assert np.array_equal(np.concatenate([chirp(n, s, min(n, s + 24) - s) for s in range(0, n, 24)]), x)
    assert np.array_equal(np.concatenate([chirp_spectrum(n, s, min(n, s + 24) - s) for s in range(0, n, 24)]), ref)


def test_separable_spectrum():
    """A sum of separable terms u (x) v (x) w transforms term by term to
    DFT(u) (x) DFT(v) (x) DFT(w): the all-bins check of the 1024^3 FFTND."""
    rng = np.random.default_rng(11)
    dims = (6, 10, 8)
    terms = [[rng.standard_normal(d) + 1j * rng.standard_normal(d) for d in dims] for _ in range(2)]
    x = sum(np.einsum("i,j,k->ijk", *t) for t in terms)
    spec = [[oracle.dftn(v, None, -1) for v in t] for t in terms]
    ref = sum(np.einsum("i,j,k->ijk", *s) for s in spec)
    assert np.max(np.abs(oracle.dftn(x, None, -1) - ref)) < 1e-11


@pytest.mark.parametrize("n", [4099, 9001])
def test_pocketfft_prime_lengths_vs_direct(n):
    """numpy's FFT is the full-array reference for the Bluestein lengths on
    the GPU; at prime lengths it runs its own chirp-z path, checked here
    against the direct DFT."""
    x = oracle.gaussian_input((n,), 44)
    ref = oracle.dftn(x, None, -1)
    assert oracle.rel_l2(np.fft.fft(x), ref) < 1e-14
    assert oracle.rel_l2(np.fft.ifft(x) * n, oracle.dftn(x, None, 1)) < 1e-14
