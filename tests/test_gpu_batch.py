"""GPU parity of the batched uneven-shape persistent launch.

Every transform of a mixed 2-D/3-D batch (7-smooth axes, plus odd primes
through in-LDS Bluestein) against the FP64 oracle; repeated graph replays
with eager work between them must each reproduce the eager result; the
full 4096-transform workload (axes 16..256) is checked per transform by
Parseval and by a forward/backward round trip, 96 of its transforms element
by element against an FP64 FFT, and its graph replays (long levels split
over per-XCD work queues) bit for bit against the first run.
"""
import numpy as np
import pytest

import oracle
from shafft_amd import api
from shafft_amd.workloads import draw_shapes

pytestmark = pytest.mark.gpu
TOL = {api.C2C: 1e-5, api.Z2Z: 1e-12}
DT = {api.C2C: np.complex64, api.Z2Z: np.complex128}


def setup(shapes, prec, seed=45):
    bp = api.BatchPlan(shapes, prec)
    xs = [oracle.uniform_input(sh, seed + i, DT[prec]) for i, sh in enumerate(shapes)]
    a, o, w = (api.DeviceBuffer(bp.total, prec) for _ in range(3))
    a.upload(np.concatenate([x.ravel() for x in xs]))
    bp.set_buffers(a.ptr, o.ptr, w.ptr)
    return bp, xs, a, o, w


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
def test_batch_vs_oracle(gpu_lib, prec):
    shapes = draw_shapes(48, 7, lo=16, hi=64) + [(47, 31), (5, 7, 11), (1, 128), (256, 256)]
    bp, xs, a, o, w = setup(shapes, prec)
    print(bp.info())
    for direction, sign in ((api.FORWARD, -1), (api.BACKWARD, +1)):
        bp.execute(direction)
        bp.synchronize()
        y = o.download()
        worst = 0.0
        for i, (sh, x) in enumerate(zip(shapes, xs)):
            got = y[bp.offsets[i]:bp.offsets[i] + x.size].reshape(sh)
            ref = oracle.dftn(x.astype(np.complex128), None, sign)
            worst = max(worst, oracle.rel_l2(got, ref))
            assert oracle.check_rel_error(got, ref, TOL[prec]), (sh, direction)
        print(f"batch prec {prec} dir {direction}: worst relL2 {worst:.3e}")


def test_batch_graph_replays(gpu_lib):
    shapes = draw_shapes(64, 11, lo=16, hi=128)
    bp, xs, a, o, w = setup(shapes, api.C2C)
    bp.execute(api.FORWARD, use_graph=False)
    bp.synchronize()
    ref = o.download()
    for r in range(4):
        o.upload(np.zeros(bp.total, np.complex64))       # eager work between replays
        bp.execute(api.FORWARD, use_graph=True)
        bp.synchronize()
        assert np.array_equal(o.download(), ref), f"replay {r} differs"


def test_batch_inplace(gpu_lib):
    shapes = [(64, 48), (16, 20, 18), (256,), (7, 9, 5)]
    bp = api.BatchPlan(shapes, api.Z2Z)
    xs = [oracle.uniform_input(sh, 50 + i) for i, sh in enumerate(shapes)]
    a, w = api.DeviceBuffer(bp.total, api.Z2Z), api.DeviceBuffer(bp.total, api.Z2Z)
    a.upload(np.concatenate([x.ravel() for x in xs]))
    bp.set_buffers(a.ptr, a.ptr, w.ptr)
    bp.execute(api.FORWARD)
    bp.synchronize()
    y = a.download()
    for i, (sh, x) in enumerate(zip(shapes, xs)):
        got = y[bp.offsets[i]:bp.offsets[i] + x.size].reshape(sh)
        assert oracle.rel_l2(got, oracle.dftn(x)) < 1e-13


def test_batch_full_workload_properties(gpu_lib):
    """The 4096-transform complex64 workload: per-transform Parseval and
    round trip (forward, backward, divide by N), repeated graph replays
    bit-identical, and a sample of transforms element by element."""
    shapes = draw_shapes(4096, 45)
    bp = api.BatchPlan(shapes, api.C2C)
    tot = bp.total
    rng = np.random.default_rng(45)
    x = np.empty(tot, np.complex64)
    x.real = rng.uniform(-1, 1, tot)
    x.imag = rng.uniform(-1, 1, tot)
    a, o, w = (api.DeviceBuffer(tot, api.C2C) for _ in range(3))
    a.upload(x)
    bp.set_buffers(a.ptr, o.ptr, w.ptr)
    print(bp.autotune())
    bp.execute(api.FORWARD)
    bp.synchronize()
    spec = o.download()
    for r in range(2):  # graph replays (per-XCD queue heads cleared in the graph) repeat it bit for bit
        o.upload(np.zeros(tot, np.complex64))
        bp.execute(api.FORWARD)
        bp.synchronize()
        assert np.array_equal(o.download(), spec), f"replay {r} differs"
    bp.set_buffers(o.ptr, o.ptr, w.ptr)   # backward in place on the spectrum
    bp.execute(api.BACKWARD)
    bp.synchronize()
    back = o.download()
    worst_p = worst_r = 0.0
    for i, sh in enumerate(shapes):
        s, e = bp.offsets[i], bp.offsets[i] + int(np.prod(sh))
        n = e - s
        xi = x[s:e].astype(np.complex128)
        ex = float(np.vdot(xi, xi).real)
        ep = abs(float(np.vdot(spec[s:e].astype(np.complex128), spec[s:e]).real) / n - ex) / ex
        er = oracle.rel_l2(back[s:e] / n, xi)
        worst_p, worst_r = max(worst_p, ep), max(worst_r, er)
    print(f"4096 transforms, {tot} elements: worst Parseval {worst_p:.3e}, worst round trip {worst_r:.3e}")
    assert worst_p < 1e-5 and worst_r < 1e-5 * np.sqrt(np.log(2 ** 24))
    # Every element of a seeded sample of the transforms (the 8 largest and
    # 88 others, drawn from every part of the work list the per-XCD queues
    # split) against an FP64 FFT of its input.
    sizes = np.array([int(np.prod(sh)) for sh in shapes])
    pick = set(np.argsort(sizes)[-8:].tolist())
    pick |= set(np.random.default_rng(46).choice(len(shapes), 88, replace=False).tolist())
    worst = 0.0
    for i in sorted(pick):
        sh = shapes[i]
        s, e = bp.offsets[i], bp.offsets[i] + int(sizes[i])
        ref = np.fft.fftn(x[s:e].astype(np.complex128).reshape(sh))
        got = spec[s:e].reshape(sh)
        worst = max(worst, oracle.rel_l2(got, ref))
        assert oracle.check_rel_error(got, ref, 1e-5), (i, sh)
    print(f"{len(pick)} transforms element by element vs FP64: worst relL2 {worst:.3e}")


# Split-precision matrix-core stages: bound on the relative L2 error per
# transform against the FP64 oracle.  bfloat16 hi/lo keeps ~16 significand
# bits per radix-8/16 stage (rounding 2^-17 per operand; measured 1.5e-6 ..
# 5e-6 per transform); float16 hi/lo with per-butterfly scaling keeps ~22
# bits (measured ~1e-7, the FP32 level), under the complex64 bound of 1e-5.
SPLIT_TOL = {api.SPLIT_BF16: 2e-5, api.SPLIT_FP16: 1e-5}


@pytest.mark.parametrize("mode", [api.SPLIT_BF16, api.SPLIT_FP16])
def test_batch_split_precision(gpu_lib, mode):
    shapes = draw_shapes(40, 9, lo=16, hi=64) + [(256, 256), (16, 128, 8), (240, 48), (7, 9, 5), (128,)]
    bp, xs, a, o, w = setup(shapes, api.C2C)
    bp.execute(api.FORWARD)
    bp.synchronize()
    base = o.download()
    bp.set_precision_mode(mode)
    for direction, sign in ((api.FORWARD, -1), (api.BACKWARD, +1)):
        bp.execute(direction)
        bp.synchronize()
        y = o.download()
        errs = []
        for i, (sh, x) in enumerate(zip(shapes, xs)):
            got = y[bp.offsets[i]:bp.offsets[i] + x.size].reshape(sh)
            ref = oracle.dftn(x.astype(np.complex128), None, sign)
            errs.append(oracle.rel_l2(got, ref))
        errs = np.array(errs)
        print(f"mode {mode} dir {direction}: relL2 vs FP64 median {np.median(errs):.2e} max {errs.max():.2e}")
        assert errs.max() < SPLIT_TOL[mode], (mode, direction, shapes[int(errs.argmax())], errs.max())
        if direction == api.FORWARD:
            # the matrix-core path ran: shapes with a radix-8/16 stage differ from FP32
            assert not np.array_equal(y, base)
    # back to FP32: the plain result again, bit for bit
    bp.set_precision_mode(api.FP32)
    bp.execute(api.FORWARD)
    bp.synchronize()
    assert np.array_equal(o.download(), base)


def test_batch_split_precision_rejects_c128(gpu_lib):
    bp = api.BatchPlan([(16, 16)], api.Z2Z)
    with pytest.raises(api.ShafftError):
        bp.set_precision_mode(api.SPLIT_BF16)
