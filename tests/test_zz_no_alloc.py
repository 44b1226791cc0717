"""shafftAmd1DExecuteNoAlloc against shafftAmd1DExecute: the allocation-free
variant runs the same kernels with the plan's intermediates in the library's
fixed 64 MiB device array, so its output must equal the plain execute's bit
for bit, and the numpy FFT within TOL_F / TOL_D.  Two calls in a row, the
second with a smaller input (a 1,000,003-point complex128 Bluestein plan that
fills the whole array, then a 40,009-point complex64 one that reuses the
array's stale contents), then a four-step plan out of place and in place,
and the error for a plan whose intermediates exceed the array."""
import numpy as np
import pytest

import oracle
from shafft_amd import api

TOL = {api.C2C: 1e-5, api.Z2Z: 1e-12}
DT = {api.C2C: np.complex64, api.Z2Z: np.complex128}


def compare(n, prec, L=None, check_numpy=True):
    p = api.OneDBackendPlan(n, prec, L)
    x = oracle.gaussian_input((n,), 44, DT[prec])
    a, b, c = (api.DeviceBuffer(n, prec, L) for _ in range(3))
    a.upload(x)
    for direction in (api.FORWARD, api.BACKWARD):
        p.execute(a.ptr, b.ptr, direction)
        p.synchronize()
        ref = api.download_ptr(b.ptr, n, prec, L)
        p.execute_no_alloc(a.ptr, c.ptr, direction)
        p.synchronize()
        got = api.download_ptr(c.ptr, n, prec, L)
        assert np.array_equal(got, ref), (n, prec, direction)
        assert np.array_equal(api.download_ptr(a.ptr, n, prec, L), x)  # input kept
        if check_numpy:
            x128 = x.astype(np.complex128)
            want = np.fft.fft(x128) if direction == api.FORWARD else np.fft.ifft(x128) * n
            e = oracle.rel_l2(got, want)
            print(f"no-alloc n={n} prec={prec} dir={direction}: bit-identical, relL2 vs numpy {e:.3e}",
                  flush=True)
            assert oracle.check_rel_error(got, want, TOL[prec])


def compare_in_place(n, prec, L=None):
    p = api.OneDBackendPlan(n, prec, L)
    x = oracle.gaussian_input((n,), 45, DT[prec])
    a, c = api.DeviceBuffer(n, prec, L), api.DeviceBuffer(n, prec, L)
    a.upload(x)
    c.upload(x)
    p.execute(a.ptr, a.ptr, api.FORWARD)
    p.execute_no_alloc(c.ptr, c.ptr, api.FORWARD)
    p.synchronize()
    assert np.array_equal(api.download_ptr(c.ptr, n, prec, L), api.download_ptr(a.ptr, n, prec, L))


def oversize_refused(L=None):
    p = api.OneDBackendPlan(4194319, api.C2C, L)
    a, b = api.DeviceBuffer(8, api.C2C, L), api.DeviceBuffer(8, api.C2C, L)
    with pytest.raises(api.ShafftError) as e:
        p.execute_no_alloc(a.ptr, b.ptr, api.FORWARD)
    assert e.value.status == 9 and "needs" in str(e.value)


def test_no_alloc_emulated(emu_lib):
    compare(40009, api.C2C, emu_lib, check_numpy=False)
    compare(4099, api.Z2Z, emu_lib, check_numpy=False)
    compare_in_place(1 << 12, api.Z2Z, emu_lib)
    oversize_refused(emu_lib)


@pytest.mark.gpu
def test_no_alloc_matches_execute(gpu_lib):
    compare(1000003, api.Z2Z)  # intermediates 2 x 2^21 complex128: the whole 64 MiB
    compare(40009, api.C2C)  # second call, smaller input, stale array contents
    compare(1 << 20, api.C2C)  # four-step, input preserved through one temporary
    compare_in_place(1 << 18, api.Z2Z)


@pytest.mark.gpu
def test_no_alloc_rejects_oversize(gpu_lib):
    oversize_refused()
