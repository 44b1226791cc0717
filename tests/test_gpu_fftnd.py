"""GPU parity of the FFTND path against the FP64 oracle.

Cases follow the reference's own tests:
  * seam-level partial-axis DFTs (tests/unit/test_fft_partial.cpp:231-237),
    tolerance 1e-10 absolute as there;
  * closed forms of tests/integration/cpp/test_correctness.cpp (delta, plane
    wave, constant, Parseval) and test_delta.cpp;
  * round trips with check_rel_error (tests/include/test_utils.hpp:84-122),
    TOL_F = 1e-5, TOL_D = 1e-12;
  * odd extents 63x47x31 (test_roundtrip.cpp:140-168);
  * the headline shapes 512^3 complex128 and 1024^3 complex64 through
    size-independent properties.
"""
import numpy as np
import pytest

import oracle
from shafft_amd import api

pytestmark = pytest.mark.gpu

TOL = {api.C2C: 1e-5, api.Z2Z: 1e-12}
DT = {api.C2C: np.complex64, api.Z2Z: np.complex128}


def run_seam(dims, axes, prec, x, direction):
    p = api.NDBackendPlan(dims, axes, prec)
    n = int(np.prod(dims))
    a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
    a.upload(x.ravel())
    d, _ = p.execute(a.ptr, b.ptr, direction)
    return api.download_ptr(d, n, prec).reshape(dims)


def test_partial_axes_reference_cases(gpu_lib):
    # dft_partial input: sin(0.01 i) + i cos(0.02 i) (test_fft_partial.cpp:35-40)
    cases = [((2, 3, 4, 5), (0, 1, 2)), ((4, 3, 2, 3, 2), (1, 2, 3)), ((2, 3, 4, 5), (2, 3)),
             ((2, 3, 4, 2), (0, 1, 2, 3)), ((3, 2, 2, 2, 2), (1, 2, 3, 4))]
    for dims, axes in cases:
        i = np.arange(int(np.prod(dims)))
        x = (np.sin(0.01 * i) + 1j * np.cos(0.02 * i)).reshape(dims)
        for direction, sign in ((api.FORWARD, -1), (api.BACKWARD, +1)):
            y = run_seam(dims, list(axes), api.Z2Z, x, direction)
            ref = oracle.dftn(x, axes, sign)
            err = np.max(np.abs(y - ref))
            print(f"partial {dims} axes {axes} dir {direction}: max abs err {err:.2e}")
            assert err < 1e-10


@pytest.mark.parametrize("dims,axes,prec", [((3, 4099), (1,), api.Z2Z), ((2, 9001), (1,), api.C2C),
                                             ((3, 2, 4099), (1, 2), api.Z2Z)])
def test_long_prime_axis_three_pass_bluestein(gpu_lib, dims, axes, prec):
    """A prime contiguous axis past the LDS cap (three-pass Bluestein with
    other axes as outer lines), both directions, against the oracle."""
    x = oracle.uniform_input(dims, 17, DT[prec])
    for direction, sign in ((api.FORWARD, -1), (api.BACKWARD, +1)):
        y = run_seam(list(dims), list(axes), prec, x, direction)
        ref = oracle.dftn(x.astype(np.complex128), axes, sign)
        e = oracle.rel_l2(y, ref)
        print(f"{dims} axes {axes} prec {prec} dir {direction}: relL2 {e:.3e}", flush=True)
        assert oracle.check_rel_error(y, ref, TOL[prec])


HIGHDIM_CASES = [(3, 2, 4, 3, 2), (2, 3, 2, 3, 2, 3), (3, 2, 3, 2, 3, 2, 3), (2, 3, 2, 3, 2, 3, 2, 3),
                 (3, 2, 3, 2, 3, 2, 3, 2, 3), (2, 3) * 5, (2, 3) * 5 + (2,),
                 (2, 2, 2, 3, 2, 2, 2, 3, 2, 2, 2, 3)]


def test_highdim_reference_cases(gpu_lib):
    """5-D to 12-D complex128 forward transforms through the public C API,
    max abs error 1e-10 against the direct DFT (tests/unit/test_fft_highdim.cpp:
    189-213, input sin(0.01 i) + i cos(0.02 i) as :65-70)."""
    for dims in HIGHDIM_CASES:
        i = np.arange(int(np.prod(dims)))
        x = (np.sin(0.01 * i) + 1j * np.cos(0.02 * i)).reshape(dims)
        plan = api.FFTND()
        plan.init([1] * len(dims), dims, api.Z2Z)
        plan.plan()
        n = plan.alloc_size()
        a, b = api.DeviceBuffer(n, api.Z2Z), api.DeviceBuffer(n, api.Z2Z)
        a.upload(x.ravel())
        plan.set_buffers(a.ptr, b.ptr)
        plan.execute(api.FORWARD)
        d, _ = plan.get_buffers()
        y = api.download_ptr(d, n, api.Z2Z).reshape(dims)
        err = np.max(np.abs(y - oracle.dftn(x, None, -1)))
        print(f"highdim {len(dims)}-D {dims}: max abs err {err:.2e}", flush=True)
        assert err < 1e-10
        plan.release()


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
@pytest.mark.parametrize("dims", [(64, 64, 32), (128, 64), (63, 47, 31), (32, 32, 32), (1000,), (5, 7, 11, 13)])
def test_forward_backward_vs_oracle(gpu_lib, prec, dims):
    x = oracle.uniform_input(dims, 42, DT[prec])
    for direction, sign in ((api.FORWARD, -1), (api.BACKWARD, +1)):
        y = run_seam(dims, list(range(len(dims))), prec, x, direction)
        ref = oracle.dftn(x.astype(np.complex128), None, sign)
        e = oracle.rel_l2(y, ref)
        print(f"{dims} prec {prec} dir {direction}: relL2 {e:.3e}")
        assert oracle.check_rel_error(y, ref, TOL[prec])


def api_roundtrip(dims, prec, x):
    plan = api.FFTND()
    plan.init([1] * len(dims), dims, prec)
    plan.plan()
    n = plan.alloc_size()
    assert n == int(np.prod(dims))
    a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
    a.upload(x.ravel())
    plan.set_buffers(a.ptr, b.ptr)
    plan.execute(api.FORWARD)
    d, w = plan.get_buffers()
    spec = api.download_ptr(d, n, prec).reshape(dims)
    plan.normalize()
    plan.execute(api.BACKWARD)
    plan.normalize()
    d, w = plan.get_buffers()
    back = api.download_ptr(d, n, prec).reshape(dims)
    return spec, back


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
def test_config1_roundtrip_api(gpu_lib, prec):
    dims = (64, 64, 32)
    x = oracle.uniform_input(dims, 42, DT[prec])
    spec, back = api_roundtrip(dims, prec, x)
    ref = oracle.dftn(x.astype(np.complex128), None, -1)
    e1, e2 = oracle.rel_l2(spec, ref), oracle.rel_l2(back, x)
    print(f"C1 prec {prec}: forward relL2 {e1:.3e} roundtrip relL2 {e2:.3e}")
    assert oracle.check_rel_error(spec, ref, TOL[prec])
    assert oracle.check_rel_error(back, x, TOL[prec])


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
@pytest.mark.parametrize("dims", [(1, 1, 1), (1, 64, 1), (32, 1, 17), (1, 1, 4099), (2,), (1, 2, 1, 3)])
def test_degenerate_extents(gpu_lib, prec, dims):
    """Unit-length axes (identity transforms), a length-2 line and a prime
    Bluestein axis between unit axes, through the public API: forward against
    the FP64 oracle and the normalized round trip."""
    x = oracle.uniform_input(dims, 7, DT[prec])
    spec, back = api_roundtrip(dims, prec, x)
    ref = oracle.dftn(x.astype(np.complex128), None, -1)
    e1, e2 = oracle.rel_l2(spec, ref), oracle.rel_l2(back, x)
    print(f"{dims} prec {prec}: forward relL2 {e1:.3e} roundtrip relL2 {e2:.3e}")
    assert oracle.check_rel_error(spec, ref, TOL[prec])
    assert oracle.check_rel_error(back, x, TOL[prec])


def test_delta_and_plane_wave_closed_forms(gpu_lib):
    dims = (64, 64, 32)
    n = int(np.prod(dims))
    x = np.zeros(dims, np.complex128)
    x[0, 0, 0] = 1
    y = run_seam(dims, [0, 1, 2], api.Z2Z, x, api.FORWARD)
    assert np.max(np.abs(y - 1)) < 1e-10
    k = (2, 3, 1)
    idx = np.indices(dims)
    ph = sum(k[i] * idx[i] / dims[i] for i in range(3))
    x = np.exp(2j * np.pi * ph)
    y = run_seam(dims, [0, 1, 2], api.Z2Z, x, api.FORWARD)
    ref = np.zeros(dims, np.complex128)
    ref[k] = n
    assert np.max(np.abs(y - ref)) < 1e-10 * n
    x = np.full(dims, 1 + 0.5j)
    y = run_seam(dims, [0, 1, 2], api.Z2Z, x, api.FORWARD)
    ref = np.zeros(dims, np.complex128)
    ref[0, 0, 0] = n * (1 + 0.5j)
    assert np.max(np.abs(y - ref)) < 1e-10 * n


def test_normalize_factor_api(gpu_lib):
    # delta -> forward -> normalize gives 1/sqrt(N) everywhere (test_delta.cpp:12-177)
    dims = (64, 64, 32)
    for prec in (api.C2C, api.Z2Z):
        x = np.zeros(dims, DT[prec])
        x[0, 0, 0] = 1
        plan = api.FFTND()
        plan.init([1, 1, 1], dims, prec)
        plan.plan()
        n = plan.alloc_size()
        a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
        a.upload(x.ravel())
        plan.set_buffers(a.ptr, b.ptr)
        plan.execute(api.FORWARD)
        plan.normalize()
        d, _ = plan.get_buffers()
        y = api.download_ptr(d, n, prec)
        f = oracle.norm_factor_nd(dims, 1, prec)
        assert np.all(y == DT[prec](f))


def _bin_direct(x, dims, kk):
    """One DFT bin X[kk] by separable direct summation in FP64 (O(N))."""
    v = [np.exp(-2j * np.pi * np.mod(np.arange(d, dtype=np.float64) * k, d) / d) for d, k in zip(dims, kk)]
    xs = x.reshape(dims)
    acc = np.zeros(dims[0], np.complex128)
    for i0 in range(dims[0]):
        acc[i0] = (xs[i0].astype(np.complex128) @ v[2]) @ v[1]
    return acc @ v[0]


def _energy(a, chunk=1 << 24):
    e = 0.0
    for s in range(0, a.size, chunk):
        c = a[s:s + chunk].astype(np.complex128)
        e += float(np.vdot(c, c).real)
    return e


def _rel_l2_chunked(a, b, chunk=1 << 24):
    num = den = 0.0
    for s in range(0, a.size, chunk):
        ca, cb = a[s:s + chunk].astype(np.complex128), b[s:s + chunk].astype(np.complex128)
        num += float(np.vdot(ca - cb, ca - cb).real)
        den += float(np.vdot(cb, cb).real)
    return np.sqrt(num / den)


@pytest.mark.parametrize("prec,dims", [(api.Z2Z, (512, 512, 512)), (api.C2C, (1024, 1024, 1024))])
def test_headline_shapes_properties(gpu_lib, prec, dims):
    """Full-size properties on the C2 input (std::mt19937_64 seed 43):
    Parseval, spot bins by direct summation, forward+backward+normalize
    round trip (every bin is checked in test_gpu_parity_headline.py)."""
    n = int(np.prod(dims))
    x = oracle.uniform_input((n,), 43, DT[prec])  # C2 input recipe
    plan = api.FFTND()
    plan.init([1, 1, 1], dims, prec)
    plan.plan()
    a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
    a.upload(x)
    plan.set_buffers(a.ptr, b.ptr)
    plan.execute(api.FORWARD)
    d, _ = plan.get_buffers()
    spec = api.download_ptr(d, n, prec)
    ex = _energy(x)
    e_par = abs(_energy(spec) / n - ex) / ex
    for kk in [(0, 0, 0), (1, 2, 3), (dims[0] - 1, 5, dims[2] // 2)]:
        ref = _bin_direct(x, dims, kk)
        got = spec.reshape(dims)[kk]
        e_bin = abs(got - ref) / np.sqrt(ex)
        print(f"{dims} bin {kk}: |err|/||x|| {e_bin:.3e}")
        assert e_bin < TOL[prec] * 10
    del spec
    plan.normalize()
    plan.execute(api.BACKWARD)
    plan.normalize()
    d, _ = plan.get_buffers()
    back = api.download_ptr(d, n, prec)
    e_rt = _rel_l2_chunked(back, x)
    print(f"{dims} prec {prec}: Parseval rel {e_par:.3e} roundtrip relL2 {e_rt:.3e}")
    assert e_par < TOL[prec]
    assert e_rt < TOL[prec] * np.sqrt(np.log(n))


def test_fftnd_past_int32_indices(gpu_lib):
    """2304 x 1024 x 1024 complex64 = 2,415,919,104 elements (18 GB per
    buffer): element indices past 2^31 - 1 through the rotated-layout plan.
    A unit impulse at p = (2303, 1023, 1021) must transform to the plane
    wave exp(-2 pi i sum_a k_a p_a / N_a) (spot bins with exact integer
    phases, Parseval), and forward -> normalize -> backward -> normalize must
    return it to p: an index that wrapped anywhere would move or lose it."""
    dims, prec = (2304, 1024, 1024), api.C2C
    n = int(np.prod(dims))
    p = (2303, 1023, 1021)
    flat = (p[0] * dims[1] + p[1]) * dims[2] + p[2]
    x = np.zeros(n, np.complex64)
    x[flat] = 1
    print(api.NDBackendPlan(list(dims), [0, 1, 2], prec).describe(), flush=True)
    plan = api.FFTND()
    plan.init([1, 1, 1], dims, prec)
    plan.plan()
    a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
    a.upload(x)
    del x
    plan.set_buffers(a.ptr, b.ptr)
    plan.execute(api.FORWARD)
    d, _ = plan.get_buffers()
    spec = api.download_ptr(d, n, prec).reshape(dims)
    errs = []
    for kk in [(0, 0, 0), (1, 2, 3), (2303, 1023, 1023), (1500, 7, 900)]:
        turns = sum(((k * q) % m) / m for k, q, m in zip(kk, p, dims))
        errs.append(abs(complex(spec[kk]) - np.exp(-2j * np.pi * turns)))
    e_par = abs(_energy(spec.ravel()) / n - 1)
    del spec
    plan.normalize()
    plan.execute(api.BACKWARD)
    plan.normalize()
    d, _ = plan.get_buffers()
    back = api.download_ptr(d, n, prec)
    peak = complex(back[flat])
    back[flat] = 0
    chunk = 1 << 26
    rest = max(float(np.abs(back[s:s + chunk]).max()) for s in range(0, n, chunk))
    print(f"{dims}: bins {max(errs):.3e} Parseval {e_par:.3e} back[p] {peak:.7f} "
          f"max elsewhere {rest:.3e}", flush=True)
    assert max(errs) < 1e-4
    assert e_par < 1e-5
    assert abs(peak - 1) < 1e-5
    assert rest < 1e-5
