"""The fixed kernel's tile orders give the same bits.

Passes strided by at least 2^16 elements on both sides walk rows of 32
tiles down the rows first (the default); SHAFFT_AMD_XCD=0 keeps the plain
order everywhere, SHAFFT_AMD_XCD=2 walks rows on every pass and =1 gives
each XCD a contiguous eighth of the tiles.  Each tile is computed the same
way whichever workgroup runs it, so all four orders must agree bit for bit;
the default is also checked against numpy's FP64 FFT.
"""
import numpy as np
import pytest

import oracle
from shafft_amd import api

pytestmark = pytest.mark.gpu
TOL = {api.C2C: 1e-5, api.Z2Z: 1e-12}
DT = {api.C2C: np.complex64, api.Z2Z: np.complex128}
ORDERS = ["default", "0", "1", "2"]


def _run(monkeypatch, make, n, prec, x):
    outs = {}
    for o in ORDERS:
        if o == "default":
            monkeypatch.delenv("SHAFFT_AMD_XCD", raising=False)
        else:
            monkeypatch.setenv("SHAFFT_AMD_XCD", o)
        p = make()
        a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
        a.upload(x.ravel())
        d = p(a.ptr, b.ptr)
        outs[o] = (api.download_ptr(d, n, prec), p.describe)
    return outs


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
def test_fftnd_tile_orders(gpu_lib, monkeypatch, prec):
    dims = (512, 256, 256)  # the 512-point pass runs at stride 2^16
    n = int(np.prod(dims))
    x = oracle.uniform_input(dims, 61, DT[prec])

    def make():
        plan = api.NDBackendPlan(list(dims), [0, 1, 2], prec)

        def run(a, b):
            d, _ = plan.execute(a, b, api.FORWARD)
            return d
        run.describe = plan.describe()
        return run
    outs = _run(monkeypatch, make, n, prec, x)
    ref = outs["default"][0]
    for o in ORDERS[1:]:
        assert np.array_equal(outs[o][0], ref), o
    spec = np.fft.fftn(x.astype(np.complex128))
    assert oracle.check_rel_error(ref.reshape(dims), spec, TOL[prec])


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
def test_fft1d_tile_orders(gpu_lib, monkeypatch, prec):
    n = 1 << 24  # four-step 256 x 256 x 256: the first pass at stride 2^16
    x = oracle.uniform_input((n,), 62, DT[prec])

    def make():
        plan = api.OneDBackendPlan(n, prec)

        def run(a, b):
            plan.execute(a, b, api.FORWARD)
            plan.synchronize()
            return b
        run.describe = plan.describe()
        return run
    outs = _run(monkeypatch, make, n, prec, x)
    ref = outs["default"][0]
    for o in ORDERS[1:]:
        assert np.array_equal(outs[o][0], ref), o
    assert oracle.check_rel_error(ref, np.fft.fft(x.astype(np.complex128)), TOL[prec])
