"""Streams through the public API and the reference's high-dimensional
integration cases.

* Streams (tests/integration/cpp/test_backend.cpp:251-450): a forward /
  normalize / backward / normalize round trip on a created stream
  (shafftSetStream after plan, as :266-276), the same after switching to a
  second stream and back to the default stream: every run bit-identical to
  the default-stream run, and the round trip within tolerance; the same for
  FFT1D.
* tests/integration/cpp/test_highdim.cpp:171-247: {4,4,4,4,4} C2C and Z2Z,
  {4,4,4,2,2,2,2} C2C and Z2Z, {4,4,2,2,2,2,2}, 2^9, 2^11, {8,4,2,4,8} and
  {3,5,7,3,5}, with that file's index tensor ((gidx - mid)/N, -(gidx -
  mid)/N) (:21-65): forward, backward, normalize, check_rel_error at
  TOL_F / TOL_D; the forward spectrum is also compared with the oracle.  The
  CPU suite runs the same cases on the emulation build.
"""
import ctypes as ct

import numpy as np
import pytest

import oracle
from shafft_amd import api

TOL = {api.C2C: 1e-5, api.Z2Z: 1e-12}
DT = {api.C2C: np.complex64, api.Z2Z: np.complex128}

HIGHDIM = [((4, 4, 4, 4, 4), api.C2C), ((4, 4, 4, 4, 4), api.Z2Z), ((4, 4, 4, 2, 2, 2, 2), api.C2C),
           ((4, 4, 4, 2, 2, 2, 2), api.Z2Z), ((4, 4, 2, 2, 2, 2, 2), api.C2C), ((2,) * 9, api.C2C),
           ((2,) * 11, api.C2C), ((8, 4, 2, 4, 8), api.C2C), ((3, 5, 7, 3, 5), api.C2C)]


def index_tensor(dims, prec):
    n = int(np.prod(dims))
    val = (np.arange(n, dtype=np.float64) - (n - 1) / 2.0) * (1.0 / n)
    return (val - 1j * val).astype(DT[prec]).reshape(dims)


def highdim_roundtrip(dims, prec, L=None):
    x = index_tensor(dims, prec)
    plan = api.FFTND(L)
    plan.init([1] * len(dims), dims, prec)
    plan.plan()
    n = plan.alloc_size()
    a, b = api.DeviceBuffer(n, prec, L), api.DeviceBuffer(n, prec, L)
    a.upload(x.ravel())
    plan.set_buffers(a.ptr, b.ptr)
    plan.execute(api.FORWARD)
    d, _ = plan.get_buffers()
    spec = api.download_ptr(d, n, prec, L).reshape(dims)
    plan.execute(api.BACKWARD)
    plan.normalize()
    d, _ = plan.get_buffers()
    back = api.download_ptr(d, n, prec, L).reshape(dims)
    plan.release()
    return x, spec, back


def check_highdim(dims, prec, L=None):
    x, spec, back = highdim_roundtrip(dims, prec, L)
    ref = oracle.dftn(x.astype(np.complex128), None, -1)
    e1, e2 = oracle.rel_l2(spec, ref), oracle.rel_l2(back, x)
    print(f"highdim {dims} prec {prec}: forward relL2 {e1:.3e} roundtrip relL2 {e2:.3e}", flush=True)
    assert oracle.check_rel_error(back, x, TOL[prec])
    assert oracle.check_rel_error(spec, ref, TOL[prec])


@pytest.mark.parametrize("dims,prec", HIGHDIM)
def test_highdim_integration_cases_emulated(emu_lib, dims, prec):
    check_highdim(dims, prec, emu_lib)


@pytest.mark.gpu
@pytest.mark.parametrize("dims,prec", HIGHDIM)
def test_highdim_integration_cases(gpu_lib, dims, prec):
    check_highdim(dims, prec)


class _Hip:
    def __init__(self):
        self.L = ct.CDLL("libamdhip64.so")
        self.L.hipStreamCreate.argtypes = [ct.POINTER(ct.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [ct.c_void_p]

    def stream(self):
        s = ct.c_void_p()
        assert self.L.hipStreamCreate(ct.byref(s)) == 0
        return s.value

    def destroy(self, s):
        assert self.L.hipStreamDestroy(ct.c_void_p(s)) == 0


def _api_roundtrip(plan, a, b, x, prec):
    n = x.size
    a.upload(x.ravel())
    plan.set_buffers(a.ptr, b.ptr)
    plan.execute(api.FORWARD)
    d, _ = plan.get_buffers()
    spec = api.download_ptr(d, n, prec)
    plan.normalize()
    plan.execute(api.BACKWARD)
    plan.normalize()
    d, _ = plan.get_buffers()
    return spec, api.download_ptr(d, n, prec)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
@pytest.mark.parametrize("kind", ["nd", "1d"])
def test_streams_custom_switch_default(gpu_lib, kind, prec):
    hip = _Hip()
    if kind == "nd":
        dims = (64, 48, 40)
        plan = api.FFTND()
        plan.init([1, 1, 1], dims, prec)
    else:
        dims = (1 << 18,)
        plan = api.FFT1D()
        plan.init(dims[0], dims[0], 0, prec)
    plan.plan()
    x = oracle.uniform_input(dims, 42, DT[prec])
    n = plan.alloc_size()
    a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
    s1, s2 = hip.stream(), hip.stream()
    runs = {}
    for name, s in (("default", 0), ("stream1", s1), ("stream2", s2), ("stream1 again", s1),
                    ("default again", 0)):
        plan.set_stream(s)
        runs[name] = _api_roundtrip(plan, a, b, x, prec)
    spec0, back0 = runs["default"]
    for name, (spec, back) in runs.items():
        assert np.array_equal(spec, spec0) and np.array_equal(back, back0), name
    e = oracle.rel_l2(back0, x.ravel())
    print(f"{kind} prec {prec}: streams bit-identical, round trip relL2 {e:.3e}", flush=True)
    assert oracle.check_rel_error(back0, x.ravel(), TOL[prec])
    plan.release()
    hip.destroy(s1)
    hip.destroy(s2)

