"""GPU parity over every 7-smooth line length from 2 to 4096 (and a few
longer ones), so that every radix plan the planner can pick (radices 2-10,
12, 14, 15, 16, 18, 20, 21, and 24, 25, 28 for complex double), with and
without register-resident first/last stages and on padded or unpadded LDS
rows, runs at least once.

Each length runs through the seam (the reference's fftndPlan/fftndExecute,
src/cpp/fft/hipfft_method/fftnd_method.hpp:14-31) on a 2-D tensor of 5 lines
in both layouts: (5, n) over axis 1 (unit-stride lines, k-fast lane mapping)
and (n, 5) over axis 0 (5 adjacent lines at stride 5, line-fast mapping),
forward and backward, against the FP64 direct DFT for n <= 1024 and an FP64
FFT above, with check_rel_error at TOL_F / TOL_D
(tests/include/test_utils.hpp:84-122).
"""
import numpy as np
import pytest

import oracle
from shafft_amd import api

TOL = {api.C2C: 1e-5, api.Z2Z: 1e-12}
DT = {api.C2C: np.complex64, api.Z2Z: np.complex128}


def smooth7(n: int) -> bool:
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


LENGTHS = [n for n in range(2, 4097) if smooth7(n)] + [4725, 6125, 7203, 8192, 10080, 12005]


def reference(x, axis, sign):
    if x.shape[axis] <= 1024:
        return oracle.dftn(x, (axis,), sign)
    return np.fft.fft(x, axis=axis) if sign < 0 else np.fft.ifft(x, axis=axis) * x.shape[axis]


def sweep(prec, lengths, L=None):
    worst = {}
    for n in lengths:
        for dims, axis in (((5, n), 1), ((n, 5), 0)):
            x = oracle.uniform_input(dims, 60 + n % 97, DT[prec])
            x64 = x.astype(np.complex128)
            p = api.NDBackendPlan(list(dims), [axis], prec, L)
            a, b = api.DeviceBuffer(x.size, prec, L), api.DeviceBuffer(x.size, prec, L)
            for direction, sign in ((api.FORWARD, -1), (api.BACKWARD, +1)):
                a.upload(x.ravel())
                d, _ = p.execute(a.ptr, b.ptr, direction)
                got = api.download_ptr(d, x.size, prec, L).reshape(dims)
                ref = reference(x64, axis, sign)
                e = oracle.rel_l2(got, ref)
                key = "k-fast" if axis == 1 else "line-fast"
                worst[key] = max(worst.get(key, 0.0), e)
                assert oracle.check_rel_error(got, ref, TOL[prec]), (n, dims, direction, e, p.describe())
    print(f"prec {prec}: {len(lengths)} lengths, worst relL2 " +
          ", ".join(f"{k} {v:.3e}" for k, v in worst.items()))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
def test_every_7smooth_length(gpu_lib, prec):
    sweep(prec, LENGTHS)


# The CPU emulation build replays every pass of the same plans with bounds
# checks on each LDS and buffer access (every eighth length and two of the
# longer ones, to keep the CPU suite short).
@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
def test_every_7smooth_length_emulated(emu_lib, prec):
    sweep(prec, LENGTHS[::8] + [4725, 8192], emu_lib)


def batch_sweep(prec, lengths, L=None):
    """The same lengths through the batched launch (each pass re-tiled for
    the batch kernel): one plan of (5, n) and (n, 5) transforms over both
    axes, forward and backward, each against the FP64 reference."""
    shapes = [s for n in lengths for s in ((5, n), (n, 5))]
    bp = api.BatchPlan(shapes, prec, L)
    xs = [oracle.uniform_input(sh, 70 + i % 89, DT[prec]) for i, sh in enumerate(shapes)]
    a, o, w = (api.DeviceBuffer(bp.total, prec, L) for _ in range(3))
    a.upload(np.concatenate([x.ravel() for x in xs]))
    bp.set_buffers(a.ptr, o.ptr, w.ptr)
    for direction, sign in ((api.FORWARD, -1), (api.BACKWARD, +1)):
        bp.execute(direction, use_graph=False)
        bp.synchronize()
        y = o.download()
        worst = 0.0
        for i, (sh, x) in enumerate(zip(shapes, xs)):
            got = y[bp.offsets[i]:bp.offsets[i] + x.size].reshape(sh)
            ref = reference(reference(x.astype(np.complex128), 0, sign), 1, sign)
            worst = max(worst, oracle.rel_l2(got, ref))
            assert oracle.check_rel_error(got, ref, TOL[prec]), (sh, direction)
        print(f"batched prec {prec} dir {direction}: {len(shapes)} transforms, worst relL2 {worst:.3e}")


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
def test_every_7smooth_length_batched(gpu_lib, prec):
    batch_sweep(prec, [n for n in LENGTHS if n <= 4096])


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
def test_every_7smooth_length_batched_emulated(emu_lib, prec):
    batch_sweep(prec, LENGTHS[:-6:8], emu_lib)


# Longer FFT1D lengths (four-step splits into two or three passes, the
# four-step twiddle fused into the stores): a seeded draw of 7-smooth lengths
# up to 2^20 plus the powers of two 2^13 .. 2^22.
def long_lengths(count, hi, seed):
    pool = [n for n in range(4097, hi + 1) if smooth7(n)]
    rng = np.random.default_rng(seed)
    return sorted(set(rng.choice(pool, count, replace=False).tolist()))


def fft1d_sweep(prec, lengths, L=None):
    worst = 0.0
    for n in lengths:
        x = oracle.uniform_input((n,), 80 + n % 89, DT[prec])
        x64 = x.astype(np.complex128)
        p = api.FFT1D(L)
        p.init(n, n, 0, prec)
        p.plan()
        a, b = api.DeviceBuffer(n, prec, L), api.DeviceBuffer(n, prec, L)
        for direction, sign in ((api.FORWARD, -1), (api.BACKWARD, +1)):
            a.upload(x)
            p.set_buffers(a.ptr, b.ptr)
            p.execute(direction)
            d, _ = p.get_buffers()
            got = api.download_ptr(d, n, prec, L)
            ref = np.fft.fft(x64) if sign < 0 else np.fft.ifft(x64) * n
            worst = max(worst, oracle.rel_l2(got, ref))
            assert oracle.check_rel_error(got, ref, TOL[prec]), (n, direction)
        p.release()
    print(f"FFT1D prec {prec}: {len(lengths)} lengths {lengths[0]}..{lengths[-1]}, worst relL2 {worst:.3e}")


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
def test_fft1d_long_7smooth_lengths(gpu_lib, prec):
    fft1d_sweep(prec, long_lengths(48, 1 << 20, 81) + [1 << k for k in range(13, 23)])


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
def test_fft1d_long_7smooth_lengths_emulated(emu_lib, prec):
    fft1d_sweep(prec, long_lengths(6, 1 << 16, 82), emu_lib)
