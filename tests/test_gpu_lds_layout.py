"""Both LDS row layouts of the generic pass kernel on the GPU: every shape
runs with padded rows forced (SHAFFT_AMD_LDS_PAD=1) and with unpadded rows
forced (=0), whatever the planner's bank model would pick.  The layout only
moves where a stage keeps its data, not the arithmetic, so the two results
must be bit-identical; both are checked against the FP64 oracle at the
precision's tolerance.  The shapes cover k-fast and line-fast lane
mappings, radices that conflict on unpadded rows (power-of-two first
stages) and odd radices that do not, wide radices (one butterfly per
thread), single-line tiles and the four-step twiddle."""
import numpy as np
import pytest

import oracle
from shafft_amd import api

pytestmark = pytest.mark.gpu
TOL = {api.C2C: 1e-5, api.Z2Z: 1e-12}
DT = {api.C2C: np.complex64, api.Z2Z: np.complex128}

ND = [((9, 405), api.C2C), ((405, 9), api.C2C), ((7, 378), api.Z2Z), ((378, 7), api.Z2Z),
      ((3, 336), api.C2C), ((5, 6, 320), api.C2C), ((12, 250), api.Z2Z), ((2, 5488), api.C2C),
      ((4, 441), api.Z2Z), ((6, 625), api.C2C)]
ONE_D = [(6048, api.C2C), (5670, api.Z2Z), (405 * 405, api.C2C), (60000, api.Z2Z)]


def _nd(dims, prec, x, direction):
    p = api.NDBackendPlan(list(dims), list(range(len(dims))), prec)
    n = x.size
    a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
    a.upload(x.ravel())
    d, _ = p.execute(a.ptr, b.ptr, direction)
    return api.download_ptr(d, n, prec).reshape(dims), p.describe()


def _1d(n, prec, x, direction):
    p = api.OneDBackendPlan(n, prec)
    a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
    a.upload(x)
    p.execute(a.ptr, b.ptr, direction)
    p.synchronize()
    return b.download(), p.describe()


@pytest.mark.parametrize("dims,prec", ND)
def test_nd_padded_equals_unpadded(gpu_lib, monkeypatch, dims, prec):
    x = oracle.uniform_input(dims, 31, DT[prec])
    for direction, sign in ((api.FORWARD, -1), (api.BACKWARD, 1)):
        out = {}
        for pad in ("1", "0"):
            monkeypatch.setenv("SHAFFT_AMD_LDS_PAD", pad)
            out[pad], desc = _nd(dims, prec, x, direction)
            if pad == "0":
                assert "unpadded" in desc, desc
        assert np.array_equal(out["1"], out["0"]), (dims, direction)
        ref = oracle.dftn(x.astype(np.complex128), list(range(len(dims))), sign)
        e = oracle.rel_l2(out["0"], ref)
        print(f"{dims} prec {prec} dir {direction}: relL2 {e:.3e}", flush=True)
        assert oracle.check_rel_error(out["0"], ref, TOL[prec]), (dims, direction, e)


@pytest.mark.parametrize("n,prec", ONE_D)
def test_1d_padded_equals_unpadded(gpu_lib, monkeypatch, n, prec):
    x = oracle.uniform_input((n,), 32, DT[prec])
    out = {}
    for pad in ("1", "0"):
        monkeypatch.setenv("SHAFFT_AMD_LDS_PAD", pad)
        out[pad], desc = _1d(n, prec, x, api.FORWARD)
    assert np.array_equal(out["1"], out["0"]), n
    ref = np.fft.fft(x.astype(np.complex128))  # pocketfft, pinned against the direct DFT in test_oracle
    e = oracle.rel_l2(out["0"], ref)
    print(f"FFT1D {n} prec {prec}: relL2 {e:.3e} ({desc.count('pass')} passes)", flush=True)
    assert oracle.check_rel_error(out["0"], ref, TOL[prec]), (n, e)
