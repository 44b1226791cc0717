"""Element-wise GPU parity of the kernels behind the headline numbers.

Every result below is compared at every element, not through properties:
  * the 512-point fixed pass (complex float and double) on a strided, a
    contiguous and an outermost pass, against the direct DFT;
  * the 1024-point complex-float fixed pass (32 elements per thread) in the
    axis-by-axis and the rotated-layout plans, including a grid whose
    workgroups walk several tiles (the prefetching persistent loop);
  * the 512 x 512 four-step FFT1D (2^18) in both precisions and directions;
  * the full 512^3 complex128 forward and backward transform of the C2
    input recipe against the oracle's FP64 FFT;
  * 1024^3 complex64 through a sum of separable Gaussian-integer terms
    (exact in float) whose spectrum is known at every bin;
  * the 2^27-point FFT1D through the chirp closed form at every bin;
  * the Bluestein lengths 1,000,003 and 4,194,319 against numpy's FFT over
    the whole array.
Tolerances: check_rel_error (tests/include/test_utils.hpp:84-122) with
TOL_F = 1e-5, TOL_D = 1e-12, as the reference's tests; the small cases
follow tests/unit/test_fft_partial.cpp:231-237 and
tests/integration/cpp/test_correctness.cpp:349-420 at the real line lengths.
"""
import numpy as np
import pytest

import oracle
from shafft_amd import api
from tests.closed_forms import ChunkedError, chirp, chirp_spectrum

pytestmark = pytest.mark.gpu
TOL = {api.C2C: 1e-5, api.Z2Z: 1e-12}
DT = {api.C2C: np.complex64, api.Z2Z: np.complex128}
DIRS = ((api.FORWARD, -1), (api.BACKWARD, +1))


def seam_nd(dims, prec, x, direction):
    p = api.NDBackendPlan(list(dims), list(range(len(dims))), prec)
    desc = p.describe()
    n = int(np.prod(dims))
    a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
    a.upload(x.ravel())
    d, _ = p.execute(a.ptr, b.ptr, direction)
    return api.download_ptr(d, n, prec).reshape(dims), desc


def reference(x, sign):
    x = x.astype(np.complex128)
    if x.size <= 1 << 16:
        return oracle.dftn(x, None, sign)
    return oracle.fftn_pow2(x, None, sign)


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
@pytest.mark.parametrize("dims", [(5, 512, 7), (512, 8, 8), (8, 8, 512)])
def test_fixed_512_vs_direct_dft(gpu_lib, prec, dims):
    x = oracle.uniform_input(dims, 43, DT[prec])
    for direction, sign in DIRS:
        y, desc = seam_nd(dims, prec, x, direction)
        assert any("n=512" in ln and "fixed" in ln for ln in desc.splitlines()), desc
        ref = oracle.dftn(x.astype(np.complex128), None, sign)
        print(f"{dims} prec {prec} dir {direction}: relL2 {oracle.rel_l2(y, ref):.3e}", flush=True)
        assert oracle.check_rel_error(y, ref, TOL[prec])


@pytest.mark.parametrize("rotate", ["1", "0"])
@pytest.mark.parametrize("dims", [(1024, 4, 4), (4, 4, 1024), (8, 1024, 1024)])
def test_fixed_1024_complex64(gpu_lib, monkeypatch, rotate, dims):
    """1024-point complex-float fixed pass (16x8x8, 32 elements per thread);
    (8, 1024, 1024) runs 512 tiles on a 256-workgroup persistent grid."""
    monkeypatch.setenv("SHAFFT_AMD_ROTATE", rotate)
    x = oracle.uniform_input(dims, 43, np.complex64)
    for direction, sign in DIRS:
        y, desc = seam_nd(dims, api.C2C, x, direction)
        lines = desc.strip().splitlines()
        assert any("n=1024" in ln and "fixed" in ln for ln in lines), desc
        if rotate == "1":
            assert all("iks=1 " in ln for ln in lines), desc
        ref = reference(x, sign)
        print(f"{dims} rotate={rotate} dir {direction}: relL2 {oracle.rel_l2(y, ref):.3e}", flush=True)
        assert oracle.check_rel_error(y, ref, TOL[api.C2C])


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
def test_fft1d_four_step_512x512(gpu_lib, prec):
    """2^18 = 512 x 512 four-step: a strided pass with the fused twiddle and
    a contiguous pass with the fused transpose, both on the 512-point fixed
    kernel."""
    n = 1 << 18
    x = oracle.gaussian_input((n,), 44, DT[prec])
    p = api.OneDBackendPlan(n, prec)
    desc = p.describe()
    assert sum("n=512" in ln and "fixed" in ln for ln in desc.splitlines()) == 2, desc
    a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
    a.upload(x)
    for direction, sign in DIRS:
        p.execute(a.ptr, b.ptr, direction)
        p.synchronize()
        y = api.download_ptr(b.ptr, n, prec)
        ref = oracle.fftn_pow2(x.astype(np.complex128), None, sign)
        print(f"FFT1D 2^18 prec {prec} dir {direction}: relL2 {oracle.rel_l2(y, ref):.3e}", flush=True)
        assert oracle.check_rel_error(y, ref, TOL[prec])


def test_512cubed_z2z_full_vs_oracle(gpu_lib):
    """The metric's workload: 512^3 complex128, input recipe of C2
    (std::mt19937_64 seed 43, uniform [-1,1)), forward through the public
    API and backward through the seam, every element against the oracle's
    FP64 FFT (pinned against the direct DFT) at TOL_D."""
    dims = (512, 512, 512)
    n = int(np.prod(dims))
    x = oracle.uniform_input(dims, 43)
    plan = api.FFTND()
    plan.init([1, 1, 1], dims, api.Z2Z)
    plan.plan()
    a, b = api.DeviceBuffer(n, api.Z2Z), api.DeviceBuffer(n, api.Z2Z)
    a.upload(x.ravel())
    plan.set_buffers(a.ptr, b.ptr)
    plan.execute(api.FORWARD)
    d, _ = plan.get_buffers()
    y = api.download_ptr(d, n, api.Z2Z).reshape(dims)
    ref = oracle.fftn_pow2(x, None, -1)
    e = oracle.rel_l2(y, ref)
    print(f"512^3 Z2Z forward: relL2 {e:.3e}", flush=True)
    assert oracle.check_rel_error(y, ref, TOL[api.Z2Z])
    del y, ref
    plan.release()
    seam = api.NDBackendPlan(list(dims), [0, 1, 2], api.Z2Z)
    a.upload(x.ravel())
    d, _ = seam.execute(a.ptr, b.ptr, api.BACKWARD)
    y = api.download_ptr(d, n, api.Z2Z).reshape(dims)
    ref = oracle.fftn_pow2(x, None, +1)
    print(f"512^3 Z2Z backward: relL2 {oracle.rel_l2(y, ref):.3e}", flush=True)
    assert oracle.check_rel_error(y, ref, TOL[api.Z2Z])


def _gaussian_integer_vec(rng, n):
    return (rng.integers(-4, 5, n) + 1j * rng.integers(-4, 5, n)).astype(np.complex128)


def test_1024cubed_c2c_all_bins(gpu_lib):
    """1024^3 complex64 (rotated-layout plan, 2^30 elements): input
    x = sum_t u_t (x) v_t (x) w_t with Gaussian-integer entries in [-4,4],
    so every element is exact in complex64; the spectrum
    sum_t U_t (x) V_t (x) W_t (U = oracle direct DFT of u) is formed plane by
    plane and compared with the GPU result at every bin."""
    dims = (1024, 1024, 1024)
    n = int(np.prod(dims))
    rng = np.random.default_rng(45)
    terms = [[_gaussian_integer_vec(rng, d) for d in dims] for _ in range(2)]
    vw = [np.multiply.outer(t[1], t[2]) for t in terms]
    x = np.empty(dims, np.complex64)
    for i in range(dims[0]):
        x[i] = terms[0][0][i] * vw[0] + terms[1][0][i] * vw[1]
    assert np.array_equal(x[3].astype(np.complex128), terms[0][0][3] * vw[0] + terms[1][0][3] * vw[1])
    plan = api.FFTND()
    plan.init([1, 1, 1], dims, api.C2C)
    plan.plan()
    a, b = api.DeviceBuffer(n, api.C2C), api.DeviceBuffer(n, api.C2C)
    a.upload(x.ravel())
    del x
    plan.set_buffers(a.ptr, b.ptr)
    plan.execute(api.FORWARD)
    d, _ = plan.get_buffers()
    spec = api.download_ptr(d, n, api.C2C).reshape(dims)
    sp = [[oracle.dftn(v, None, -1) for v in t] for t in terms]
    VW = [np.multiply.outer(s[1], s[2]) for s in sp]
    err = ChunkedError()
    for k0 in range(dims[0]):
        err.add(spec[k0], sp[0][0][k0] * VW[0] + sp[1][0][k0] * VW[1])
    print(f"1024^3 C2C forward, all bins: relL2 {err.rel_l2:.3e} relLinf {err.rel_linf:.3e}", flush=True)
    assert err.ok(TOL[api.C2C])


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
def test_fft1d_2p27_chirp_all_bins(gpu_lib, prec):
    """2^27-point FFT1D (three 512-point passes): forward of the chirp
    exp(i pi m^2/N) against sqrt(N) e^{i pi/4} exp(-i pi k^2/N) at every bin,
    and backward of its conjugate against the conjugate spectrum."""
    n, chunk = 1 << 27, 1 << 24
    x = np.empty(n, DT[prec])
    for s in range(0, n, chunk):
        x[s:s + chunk] = chirp(n, s, chunk)
    p = api.OneDBackendPlan(n, prec)
    print(p.describe(), flush=True)
    a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
    for direction in (api.FORWARD, api.BACKWARD):
        a.upload(x if direction == api.FORWARD else np.conj(x))
        p.execute(a.ptr, b.ptr, direction)
        p.synchronize()
        y = api.download_ptr(b.ptr, n, prec)
        err = ChunkedError()
        for s in range(0, n, chunk):
            ref = chirp_spectrum(n, s, chunk)
            err.add(y[s:s + chunk], ref if direction == api.FORWARD else np.conj(ref))
        print(f"FFT1D 2^27 prec {prec} dir {direction}, all bins: relL2 {err.rel_l2:.3e} "
              f"relLinf {err.rel_linf:.3e}", flush=True)
        assert err.ok(TOL[prec])


@pytest.mark.parametrize("n,prec", [(1000003, api.C2C), (1000003, api.Z2Z), (4194319, api.C2C),
                                    (4194319, api.Z2Z)])
def test_fft1d_bluestein_full_vs_numpy(gpu_lib, n, prec):
    """Prime lengths (Bluestein: three passes over 2^21 for 1,000,003; two
    four-step chains over 2^24 for 4,194,319), input recipe of C3 (seed 44,
    Box-Muller), both directions, every bin against numpy's FP64 FFT."""
    x = oracle.gaussian_input((n,), 44, DT[prec])
    p = api.OneDBackendPlan(n, prec)
    ext = max(max(p.pass_extents(i)) for i in range(p.num_passes()))
    a, b = api.DeviceBuffer(max(ext, n), prec), api.DeviceBuffer(max(ext, n), prec)
    a.upload(x)
    x128 = x.astype(np.complex128)
    for direction in (api.FORWARD, api.BACKWARD):
        p.execute(a.ptr, b.ptr, direction)
        p.synchronize()
        y = api.download_ptr(b.ptr, n, prec)
        ref = np.fft.fft(x128) if direction == api.FORWARD else np.fft.ifft(x128) * n
        print(f"FFT1D n={n} prec {prec} dir {direction}: relL2 {oracle.rel_l2(y, ref):.3e}", flush=True)
        assert oracle.check_rel_error(y, ref, TOL[prec])
