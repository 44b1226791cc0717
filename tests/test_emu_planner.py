"""Pass programs of the planner, run through the CPU emulation build.

tests/emu/emu_pass.cpp replays each pass descriptor on the host with the
plan's own twiddle tables and radices, bounds-checks every access against
its allocation and checks the launch-shape assumptions of the kernel.  These
tests pin the planner's index arithmetic (digits, strides, fused
transposes, four-step twiddles, Bluestein hooks, buffer slots) against the
oracle without a GPU.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from shafft_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = {0: 1e-5, 1: 1e-12}
DT = {0: np.complex64, 1: np.complex128}


def nd_roundtrip(L, dims, axes, prec, seed=1):
    x = oracle.uniform_input(dims, seed, DT[prec])
    p = api.NDBackendPlan(dims, axes, prec, L=L)
    n = x.size
    a, b = api.DeviceBuffer(n, prec, L), api.DeviceBuffer(n, prec, L)
    out = {}
    for d, sign in ((api.FORWARD, -1), (api.BACKWARD, 1)):
        a.upload(x.ravel())
        dd, ww = p.execute(a.ptr, b.ptr, d)
        y = api.download_ptr(dd, n, prec, L).reshape(dims)
        out[d] = (y, oracle.dftn(x.astype(np.complex128), axes, sign))
    return p, out


ND_CASES = [
    ((64, 64, 32), 0, None), ((64, 64, 32), 1, None), ((2, 3, 4, 5), 1, (0, 1, 2)),
    ((4, 3, 2, 3, 2), 1, (1, 2, 3)), ((2, 3, 4, 5), 1, (2, 3)), ((3, 2, 2, 2, 2), 1, (1, 2, 3, 4)),
    ((63, 47, 31), 1, None), ((5, 11), 0, None), ((1, 1, 1), 1, None), ((128, 64), 0, None),
    ((360, 3), 1, None), ((12, 8192), 0, (1,)), ((2, 1009), 0, None), ((7, 5, 3, 2, 2, 2), 1, None),
]


@pytest.mark.parametrize("dims,prec,axes", ND_CASES)
def test_nd_plans(emu_lib, dims, prec, axes):
    axes = tuple(range(len(dims))) if axes is None else axes
    p, out = nd_roundtrip(emu_lib, dims, list(axes), prec)
    for d, (y, ref) in out.items():
        assert oracle.check_rel_error(y, ref, TOL[prec]), (dims, d, oracle.rel_l2(y, ref))


def _run_in_subprocess(code, env_extra):
    env = dict(os.environ)
    env.update(env_extra)
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


SUB = r"""
import sys, numpy as np
sys.path.insert(0, '.')
import oracle
from shafft_amd import api
L = api.bind('_build/emu/libshafft_emu.so')
for n, prec in CASES:
    p = api.OneDBackendPlan(n, prec, L=L)
    x = oracle.uniform_input((n,), 44, np.complex64 if prec == 0 else np.complex128)
    a = api.DeviceBuffer(n, prec, L); b = api.DeviceBuffer(n, prec, L); a.upload(x)
    for d in (0, 1):
        p.execute(a.ptr, b.ptr, d)
        y = b.download()
        assert np.array_equal(a.download(), x), 'input not preserved'
        r = np.fft.fft(x.astype(complex)) if d == 0 else np.fft.ifft(x.astype(complex)) * n
        e = np.linalg.norm(y - r) / np.linalg.norm(r)
        print(n, prec, d, p.num_passes(), e)
        assert e < (1e-5 if prec == 0 else 1e-12), (n, prec, d, e)
"""


def test_1d_plans_hardware_cap(emu_lib):
    out = _run_in_subprocess(SUB.replace("CASES", "[(1,1),(7,0),(4096,1),(8192,0),(16384,1),(100000,1),(40009,1)]"), {})
    assert "40009 1 1 3" in out  # long Bluestein: 3 passes (transposed spectrum)


def test_long_bluestein_three_pass(emu_lib):
    """Prime lengths past the LDS cap on a contiguous axis run Bluestein's
    convolution in 3 passes (forward N1-point columns with the chirp and
    W_M twiddle; per-row N2-point forward DFT, transposed filter, inverse DFT
    and conjugate twiddle; inverse N1-point columns with the output chirp),
    with other axes as outer lines; both directions, both precisions, against
    the direct DFT."""
    L = emu_lib
    for dims, axes, prec, want in [((2, 4099), [1], 1, 3), ((9001,), [0], 0, 3),
                                   ((3, 2, 4099), [1, 2], 1, 4)]:
        x = oracle.uniform_input(dims, 12, DT[prec])
        p = api.NDBackendPlan(list(dims), axes, prec, L=L)
        assert p.num_passes() == want, p.describe()
        a, b = api.DeviceBuffer(x.size, prec, L), api.DeviceBuffer(x.size, prec, L)
        for d, sign in ((api.FORWARD, -1), (api.BACKWARD, 1)):
            a.upload(x.ravel())
            dd, _ = p.execute(a.ptr, b.ptr, d)
            y = api.download_ptr(dd, x.size, prec, L).reshape(dims)
            ref = oracle.dftn(x.astype(np.complex128), axes, sign)
            assert oracle.check_rel_error(y, ref, TOL[prec]), (dims, d, oracle.rel_l2(y, ref))


def test_bluestein_fixed_kernels(emu_lib):
    """Three-pass Bluestein on the compile-time-length kernels (column
    passes of 32..512 points with the chirp hooks, row passes of 1024..8192
    points with the filter): the emulation checks each launch against what
    those kernels assume, and the results against numpy's FFT (pinned
    against the direct DFT at primes in test_oracle), with outer lines."""
    L = emu_lib
    for dims, prec, kinds in [((40009,), 0, ("blue1", "blue2", "blue3")), ((2, 40009), 1, ("blue1", "blue2", "blue3")),
                              ((3, 12007), 0, ("blue2",)), ((70001,), 1, ("blue1", "blue2", "blue3"))]:
        x = oracle.uniform_input(dims, 13, DT[prec])
        p = api.NDBackendPlan(list(dims), [len(dims) - 1], prec, L=L)
        desc = p.describe()
        assert p.num_passes() == 3, desc
        for k in kinds:
            assert k in desc, desc
        a, b = api.DeviceBuffer(x.size, prec, L), api.DeviceBuffer(x.size, prec, L)
        x128 = x.astype(np.complex128)
        for d in (api.FORWARD, api.BACKWARD):
            a.upload(x.ravel())
            dd, _ = p.execute(a.ptr, b.ptr, d)
            y = api.download_ptr(dd, x.size, prec, L).reshape(dims)
            ref = np.fft.fft(x128, axis=-1) if d == api.FORWARD else np.fft.ifft(x128, axis=-1) * dims[-1]
            assert oracle.check_rel_error(y, ref, TOL[prec]), (dims, d, oracle.rel_l2(y, ref))


def test_1d_plans_small_cap_multilevel(emu_lib):
    # A tiny LDS cap forces three-level four-step and three-level Bluestein.
    out = _run_in_subprocess(SUB.replace("CASES", "[(4096,1),(32768,1),(1000,0),(997,1),(30030,1),(6561,1)]"),
                             {"SHAFFT_AMD_LDS_CAP": "64"})
    assert "30030 1 0 6" in out


def test_nd_small_cap_multilevel(emu_lib):
    code = r"""
import sys, numpy as np
sys.path.insert(0, '.')
import oracle
from shafft_amd import api
L = api.bind('_build/emu/libshafft_emu.so')
for dims in [(64, 48, 30), (7, 1000), (3, 257, 2)]:
    p = api.NDBackendPlan(dims, list(range(len(dims))), 1, L=L)
    x = oracle.uniform_input(dims, 45)
    a = api.DeviceBuffer(x.size, 1, L); b = api.DeviceBuffer(x.size, 1, L); a.upload(x.ravel())
    d, w = p.execute(a.ptr, b.ptr, 0)
    y = api.download_ptr(d, x.size, 1, L).reshape(dims)
    e = oracle.rel_l2(y, np.fft.fftn(x))
    print(dims, p.num_passes(), e)
    assert e < 1e-12
"""
    _run_in_subprocess(code, {"SHAFFT_AMD_LDS_CAP": "32"})


def test_nd_strided_axis_split(emu_lib):
    """A strided axis longer than the strided-line cap (1024) is split
    (four-step, or a four-pass Bluestein chain for a prime), with the
    unit-stride row digit kept innermost so every pass is line-fast (C
    adjacent rows per element) on both sides; strided axes up to the cap
    stay one pass.  (Axis blocks that do not include the last axis: those
    keep the axis-by-axis plan.)"""
    L = emu_lib
    for dims, axes, prec, npass in [((1200, 6), [0], 0, 2), ((6, 1400, 5), [0, 1], 1, 3),
                                    ((1031, 4), [0], 1, 4), ((2100, 22), [0], 0, 2),
                                    ((6, 700, 5), [0, 1], 1, 2), ((600, 22), [0], 0, 1)]:
        x = oracle.uniform_input(dims, 7, DT[prec])
        p = api.NDBackendPlan(dims, axes, prec, L=L)
        assert p.num_passes() == npass, (dims, p.describe())
        for line in p.describe().strip().splitlines():
            if "iks=1 " not in line or "oks=1 " not in line:
                assert "ld=line st=line" in line, (dims, line)
        a, b = api.DeviceBuffer(x.size, prec, L), api.DeviceBuffer(x.size, prec, L)
        for d, sign in ((api.FORWARD, -1), (api.BACKWARD, 1)):
            a.upload(x.ravel())
            dd, _ = p.execute(a.ptr, b.ptr, d)
            y = api.download_ptr(dd, x.size, prec, L).reshape(dims)
            ref = oracle.dftn(x.astype(np.complex128), axes, sign)
            assert oracle.check_rel_error(y, ref, TOL[prec]), (dims, d, oracle.rel_l2(y, ref))


ROT_SUB = r"""
import sys, numpy as np
sys.path.insert(0, '.')
import oracle
from shafft_amd import api
L = api.bind('_build/emu/libshafft_emu.so')
cases = [((16, 12, 10), [0, 1, 2], 1), ((3, 8, 1, 6), [1, 2, 3], 0), ((2, 5, 7, 9, 4), [1, 2, 3, 4], 1),
         ((33, 20), [0, 1], 0), ((5, 1, 1), [0, 1, 2], 1), ((4, 600, 10), [1, 2], 0)]
for dims, axes, prec in cases:
    dt = np.complex64 if prec == 0 else np.complex128
    x = oracle.uniform_input(dims, 9, dt)
    p = api.NDBackendPlan(list(dims), axes, prec, L=L)
    for line in p.describe().strip().splitlines():
        if "twiddle" not in line and "bluestein" not in line:
            assert "iks=1 " in line, (dims, line)
    a, b = api.DeviceBuffer(x.size, prec, L), api.DeviceBuffer(x.size, prec, L)
    for d, sign in ((0, -1), (1, 1)):
        a.upload(x.ravel())
        dd, _ = p.execute(a.ptr, b.ptr, d)
        y = api.download_ptr(dd, x.size, prec, L).reshape(dims)
        ref = oracle.dftn(x.astype(np.complex128), axes, sign)
        assert oracle.check_rel_error(y, ref, 1e-5 if prec == 0 else 1e-12), (dims, axes, d)
    print(dims, p.num_passes())
"""


def test_nd_rotated_plans(emu_lib):
    """Axis blocks ending at the last axis can rotate the layout one axis per
    pass (forced here with SHAFFT_AMD_ROTATE=1; automatic for complex64
    outermost strides of 4 MiB or more): every pass reads its lines
    contiguously (iks = 1) and the last pass restores row-major order.
    Length-1 axes are skipped; leading untouched axes stay outermost."""
    out = _run_in_subprocess(ROT_SUB, {"SHAFFT_AMD_ROTATE": "1"})
    assert "(16, 12, 10) 3" in out and "(5, 1, 1) 1" in out
    L = emu_lib
    # automatic choice: 1024^3 complex64 rotates, 512^3 complex128 does not
    p = api.NDBackendPlan([1024] * 3, [0, 1, 2], 0, L=L)
    assert all("iks=1 " in ln for ln in p.describe().strip().splitlines())
    q = api.NDBackendPlan([512] * 3, [0, 1, 2], 1, L=L)
    assert "iks=262144" in q.describe()


def test_nd_small_strided_cap(emu_lib):
    code = r"""
import sys, numpy as np
sys.path.insert(0, '.')
import oracle
from shafft_amd import api
L = api.bind('_build/emu/libshafft_emu.so')
for dims, prec in [((40, 36, 20), 1), ((3, 97, 8), 1), ((2, 64, 2, 30), 0)]:
    p = api.NDBackendPlan(dims, list(range(len(dims))), prec, L=L)
    x = oracle.uniform_input(dims, 46, np.complex64 if prec == 0 else np.complex128)
    a = api.DeviceBuffer(x.size, prec, L); b = api.DeviceBuffer(x.size, prec, L); a.upload(x.ravel())
    d, w = p.execute(a.ptr, b.ptr, 1)
    y = api.download_ptr(d, x.size, prec, L).reshape(dims)
    e = oracle.rel_l2(y, np.fft.ifftn(x.astype(complex)) * x.size)
    print(dims, p.num_passes(), e)
    assert e < (1e-5 if prec == 0 else 1e-12)
"""
    out = _run_in_subprocess(code, {"SHAFFT_AMD_LF_CAP": "8"})
    assert "(40, 36, 20) 5" in out and "(3, 97, 8) 3" in out


def test_inplace_buffers(emu_lib):
    L = emu_lib
    dims = (16, 12, 10)
    x = oracle.uniform_input(dims, 2)
    p = api.NDBackendPlan(dims, [0, 1, 2], 1, L=L)
    a = api.DeviceBuffer(x.size, 1, L)
    a.upload(x.ravel())
    d, w = p.execute(a.ptr, a.ptr, 0)
    assert d == a.ptr
    assert oracle.rel_l2(a.download().reshape(dims), oracle.dftn(x)) < 1e-13
    p1 = api.OneDBackendPlan(4096 * 3, 1, L=L)
    y = oracle.uniform_input((4096 * 3,), 3)
    a1 = api.DeviceBuffer(y.size, 1, L)
    a1.upload(y)
    p1.execute(a1.ptr, a1.ptr, 1)
    assert oracle.rel_l2(a1.download(), np.fft.ifft(y) * y.size) < 1e-13


def test_c_api_state_machine_and_errors(emu_lib):
    """Status codes of the reference's state machine (src/cpp/shafft.cpp:55-75,
    172-186, shafft_internal.cpp:198-210,229-333)."""
    L = emu_lib
    import ctypes as ct
    h = ct.c_void_p()
    assert L.shafftNDCreate(ct.byref(h)) == 0
    assert L.shafftPlan(h) == 4            # not configured -> ERR_PLAN_NOT_INIT
    assert L.shafftExecute(h, 0) == 4
    dims = (ct.c_int * 3)(8, 8, 8)
    bad = (ct.c_int * 3)(2, 1, 1)
    assert L.shafftNDInit(h, 3, bad, dims, 0, 0, 0) == 7   # grid > 1 rank -> ERR_INVALID_DECOMP
    last = (ct.c_int * 3)(1, 1, 2)
    assert L.shafftNDInit(h, 3, last, dims, 0, 0, 0) == 7  # last axis distributed
    ones = (ct.c_int * 3)(1, 1, 1)
    assert L.shafftNDInit(h, 3, ones, dims, 5, 0, 0) == 8  # bad type
    zero = (ct.c_int * 3)(8, 0, 8)
    assert L.shafftNDInit(h, 3, ones, zero, 0, 0, 0) == 5  # bad dim
    assert L.shafftNDInit(h, 3, ones, dims, 0, 7, 0) == 2  # communicator other than one rank
    assert L.shafftNDInit(h, 3, ones, dims, 0, 0, 0) == 0
    assert L.shafftNDInit(h, 3, ones, dims, 0, 0, 0) == 15  # twice -> ERR_INVALID_STATE
    assert L.shafftSetStream(h, None) == 4                  # before plan()
    assert L.shafftPlan(h) == 0
    assert L.shafftPlan(h) == 15
    assert L.shafftExecute(h, 0) == 3                       # no buffers
    n = ct.c_size_t()
    assert L.shafftGetAllocSize(h, ct.byref(n)) == 0 and n.value == 512
    sub = (ct.c_size_t * 3)()
    off = (ct.c_size_t * 3)()
    assert L.shafftGetLayout(h, sub, off, 0) == 0 and list(sub) == [8, 8, 8] and list(off) == [0, 0, 0]
    ca = (ct.c_int * 3)()
    assert L.shafftGetAxes(h, ca, None, 0) == 0 and list(ca) == [0, 1, 2]
    assert L.shafftLastErrorStatus() == 0
    assert L.shafftDestroy(ct.byref(h)) == 0 and h.value is None
    assert L.shafftDestroy(ct.byref(h)) == 1
    assert L.shafftLastErrorStatus() == 1
    assert L.shafftErrorSourceName(2) == b"HIP"
    h1 = ct.c_void_p()
    assert L.shafft1DCreate(ct.byref(h1)) == 0
    assert L.shafft1DInit(h1, 100, 50, 0, 1, 0) == 12       # layout of another rank count
    assert L.shafft1DInit(h1, 100, 100, 0, 1, 0) == 0
    ln, ls, la = ct.c_size_t(), ct.c_size_t(), ct.c_size_t()
    assert L.shafftConfiguration1D(100, ct.byref(ln), ct.byref(ls), ct.byref(la), 1, 0) == 0
    assert (ln.value, ls.value, la.value) == (100, 0, 100)
    assert L.shafftDestroy(ct.byref(h1)) == 0


def test_c_api_configuration_nd_and_fortran_entry_points(emu_lib):
    """shafftConfigurationND at one rank (src/cpp/configuration.cpp:530-600:
    a grid of ones is taken as given, :478-498; any other request falls
    through to the automatic search, which packs to nda 0 and has no candidate
    for a 1-D tensor, :318-410) and the Fortran-handle companions
    (src/f03/shafftf03.cpp:11-60)."""
    L = emu_lib
    import ctypes as ct
    Z, ONE = ct.c_size_t(0), ct.c_size_t(1)   # size_t / intptr_t arguments on the stack
    C0, C3 = ct.c_ssize_t(0), ct.c_ssize_t(3)
    size = (ct.c_int * 3)(64, 48, 32)
    sub, off = (ct.c_size_t * 3)(), (ct.c_size_t * 3)()
    nda, cs = ct.c_int(), ct.c_int()
    for grid, want_nda in (((1, 1, 1), 0), ((4, 2, 1), 0), ((0, 0, 0), 0)):
        cd = (ct.c_int * 3)(*grid)
        nda.value = 2
        assert L.shafftConfigurationND(3, size, 1, cd, ct.byref(nda), sub, off, ct.byref(cs),
                                       0, ONE, C0) == 0   # memLimit 1 byte: not applied at 1 rank
        assert list(cd) == [1, 1, 1] and nda.value == want_nda and cs.value == 1
        assert list(sub) == [64, 48, 32] and list(off) == [0, 0, 0]
    one = (ct.c_int * 1)(100)
    s1, o1 = (ct.c_size_t * 1)(), (ct.c_size_t * 1)()
    cd1 = (ct.c_int * 1)(0)
    assert L.shafftConfigurationND(1, one, 0, cd1, ct.byref(nda), s1, o1, ct.byref(cs), 1, Z, C0) == 7
    cd1[0] = 1
    assert L.shafftConfigurationND(1, one, 0, cd1, ct.byref(nda), s1, o1, ct.byref(cs), 1, Z, C0) == 0
    assert s1[0] == 100 and cs.value == 1
    cd = (ct.c_int * 3)(1, 1, 1)
    assert L.shafftConfigurationND(3, size, 9, cd, ct.byref(nda), sub, off, ct.byref(cs), 0, Z, C0) == 8
    assert L.shafftConfigurationND(3, size, 0, cd, ct.byref(nda), sub, off, ct.byref(cs), 0, Z, C3) == 2
    bad = (ct.c_int * 3)(64, 0, 32)
    assert L.shafftConfigurationND(3, bad, 0, cd, ct.byref(nda), sub, off, ct.byref(cs), 0, Z, C0) == 5
    assert L.shafftConfigurationND(3, size, 0, cd, None, sub, off, ct.byref(cs), 0, Z, C0) == 1

    # Fortran companions: handle by reference, 0 = the one-rank handle.
    fh, fbad = ct.c_int(0), ct.c_int(1140850688)
    assert L.shafftConfigurationNDf03(3, size, 0, cd, ct.byref(nda), sub, off, ct.byref(cs),
                                      1, Z, ct.byref(fh)) == 0
    assert L.shafftConfigurationNDf03(3, size, 0, cd, ct.byref(nda), sub, off, ct.byref(cs),
                                      1, Z, ct.byref(fbad)) == 2
    assert L.shafftConfigurationNDf03(3, size, 0, cd, ct.byref(nda), sub, off, ct.byref(cs),
                                      1, Z, None) == 1
    h = ct.c_void_p()
    assert L.shafftNDCreate(ct.byref(h)) == 0
    pol = ct.c_int(1)
    assert L.shafftNDInitf03(h, 3, cd, size, 0, ct.byref(fh), None) == 1
    assert L.shafftNDInitf03(h, 3, cd, size, 0, ct.byref(fbad), ct.byref(pol)) == 2
    assert L.shafftNDInitf03(h, 3, cd, size, 0, ct.byref(fh), ct.byref(pol)) == 0
    n = ct.c_size_t()
    assert L.shafftGetAllocSize(h, ct.byref(n)) == 0 and n.value == 64 * 48 * 32
    assert L.shafftDestroy(ct.byref(h)) == 0
    ln, ls, la = ct.c_size_t(), ct.c_size_t(), ct.c_size_t()
    assert L.shafftConfiguration1Df03(77, ct.byref(ln), ct.byref(ls), ct.byref(la), 1,
                                      ct.byref(fh)) == 0
    assert (ln.value, ls.value, la.value) == (77, 0, 77)
    h1 = ct.c_void_p()
    assert L.shafft1DCreate(ct.byref(h1)) == 0
    assert L.shafft1DInitf03(h1, 77, 77, 0, 1, ct.byref(fbad)) == 2
    assert L.shafft1DInitf03(h1, 77, 77, 0, 1, ct.byref(fh)) == 0
    assert L.shafftDestroy(ct.byref(h1)) == 0


def test_c_api_roundtrip_and_normalize(emu_lib):
    L = emu_lib
    for prec in (0, 1):
        dims = (8, 6, 4)
        x = oracle.uniform_input(dims, 42, DT[prec])
        plan = api.FFTND(L=L)
        plan.init([1, 1, 1], dims, prec)
        plan.plan()
        n = plan.alloc_size()
        a, b = api.DeviceBuffer(n, prec, L), api.DeviceBuffer(n, prec, L)
        a.upload(x.ravel())
        plan.set_buffers(a.ptr, b.ptr)
        plan.execute(api.FORWARD)
        d, w = plan.get_buffers()
        assert (d, w) == (b.ptr, a.ptr)   # odd pass count: result behind data after the swap
        plan.execute(api.BACKWARD)
        plan.normalize()                  # exponent 2 -> 1/N
        d, _ = plan.get_buffers()
        y = api.download_ptr(d, n, prec, L).reshape(dims)
        assert oracle.check_rel_error(y, x, TOL[prec])
        f = oracle.norm_factor_nd(dims, 2, prec)
        assert f == np.float32(1 / 192) if prec == 0 else f == 1 / 192
    # FFT1D: data -> work then swap; normalize factor in working precision
    x = oracle.uniform_input((100,), 44)
    p = api.FFT1D(L=L)
    p.init(100, 100, 0, 1)
    p.plan()
    a, b = api.DeviceBuffer(100, 1, L), api.DeviceBuffer(100, 1, L)
    a.upload(x)
    p.set_buffers(a.ptr, b.ptr)
    p.execute(api.FORWARD)
    d, w = p.get_buffers()
    assert (d, w) == (b.ptr, a.ptr)
    assert np.array_equal(api.download_ptr(w, 100, 1, L), x)  # input preserved
    p.normalize()
    y = api.download_ptr(d, 100, 1, L)
    assert oracle.rel_l2(y, np.fft.fft(x) / 10) < 1e-14


def test_c_api_highdim_reference_cases(emu_lib):
    """The 5-D..12-D cases of tests/unit/test_fft_highdim.cpp:189-201 through
    the C API state machine, max abs error 1e-10 against the direct DFT."""
    L = emu_lib
    cases = [(3, 2, 4, 3, 2), (2, 3, 2, 3, 2, 3), (3, 2, 3, 2, 3, 2, 3), (2, 3, 2, 3, 2, 3, 2, 3),
             (3, 2, 3, 2, 3, 2, 3, 2, 3), (2, 3) * 5, (2, 3) * 5 + (2,), (2, 2, 2, 3) * 3]
    for dims in cases:
        i = np.arange(int(np.prod(dims)))
        x = (np.sin(0.01 * i) + 1j * np.cos(0.02 * i)).reshape(dims)
        plan = api.FFTND(L=L)
        plan.init([1] * len(dims), dims, 1)
        plan.plan()
        n = plan.alloc_size()
        a, b = api.DeviceBuffer(n, 1, L), api.DeviceBuffer(n, 1, L)
        a.upload(x.ravel())
        plan.set_buffers(a.ptr, b.ptr)
        plan.execute(api.FORWARD)
        d, _ = plan.get_buffers()
        y = api.download_ptr(d, n, 1, L).reshape(dims)
        assert np.max(np.abs(y - oracle.dftn(x, None, -1))) < 1e-10, dims
        plan.release()


def test_single_pass_extents(emu_lib):
    """Single-pass launches (profiling aid) check the caller's buffer sizes:
    a Bluestein chain's intermediate passes span the padded length M > n,
    and undersized buffers are refused instead of overrun.  With buffers of
    the reported extents every pass runs bounds-checked."""
    L = emu_lib
    n = 40009
    p = api.OneDBackendPlan(n, 1, L=L)
    ext = [p.pass_extents(i) for i in range(p.num_passes())]
    assert p.num_passes() == 3 and ext[0] == (n, 131072) and ext[-1] == (131072, n)
    big = max(max(e) for e in ext)
    a, b = api.DeviceBuffer(big, 1, L), api.DeviceBuffer(big, 1, L)
    a.upload(oracle.uniform_input((big,), 3))
    for i in range(p.num_passes()):
        p.execute_pass(i, a.ptr, big, b.ptr, big, 0)
    small_a, small_b = api.DeviceBuffer(n, 1, L), api.DeviceBuffer(n, 1, L)
    with pytest.raises(api.ShafftError, match="needs buffers"):
        p.execute_pass(0, small_a.ptr, n, small_b.ptr, n, 0)
    q = api.NDBackendPlan([8, 12], [0, 1], 0, L=L)
    assert [q.pass_extents(i) for i in range(q.num_passes())] == [(96, 96)] * q.num_passes()
    with pytest.raises(api.ShafftError):
        q.execute_pass(0, small_a.ptr, 95, small_b.ptr, 96, 0)


def test_lds_row_padding_choices(emu_lib):
    """The bank model's padded / unpadded LDS rows per pass: odd radices and
    line-fast lanes run unpadded; a k-fast pass whose first stage has a
    power-of-two radix of 16 keeps the padding; the knob forces either."""
    L = emu_lib

    def passes(make):
        return [ln for ln in make().describe().strip().splitlines()]

    nd = lambda dims, prec: (lambda: api.NDBackendPlan(list(dims), list(range(len(dims))), prec, L=L))  # noqa: E731
    one = lambda n, prec: (lambda: api.OneDBackendPlan(n, prec, L=L))  # noqa: E731
    p = passes(nd((9, 405), 0))
    assert "radices=9x9x5" in p[0] and "unpadded" in p[0], p  # odd radices, k-fast
    p = passes(nd((2, 5488), 0))
    assert "radices=16x7x7x7" in p[0] and "unpadded" not in p[0], p  # radix-16 first stage, k-fast
    p = passes(one(6048, 0))
    assert "radices=16x21x18" in p[0] and "unpadded" not in p[0], p
    p = passes(one(405 * 405, 0))
    assert all("unpadded" in ln for ln in p), p  # line-fast and k-fast passes of 9x9x5
    # fixed-kernel (power-of-two) passes always keep padded rows
    p = passes(nd((64, 64, 32), 0))
    assert all("unpadded" not in ln for ln in p if "fixed" in ln), p


def test_lds_row_padding_knob(emu_lib, monkeypatch):
    L = emu_lib
    for knob, want in (("1", False), ("0", True)):
        monkeypatch.setenv("SHAFFT_AMD_LDS_PAD", knob)
        d = api.NDBackendPlan([2, 5488], [0, 1], 0, L=L).describe().strip().splitlines()
        assert ("unpadded" in d[0]) == want, (knob, d)
