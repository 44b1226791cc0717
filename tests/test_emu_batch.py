"""Batched uneven transforms through the CPU emulation build: per-transform
results against the oracle, graph replay, in-place operation, and the work
list's dependency order (checked by the emulator on every launch)."""
import numpy as np

import oracle
from shafft_amd import api
from shafft_amd.workloads import draw_shapes, smooth7


def _run(L, shapes, prec, inplace=False, use_graph=True, direction=0):
    bp = api.BatchPlan(shapes, prec, L=L)
    dt = np.complex64 if prec == 0 else np.complex128
    xs = [oracle.uniform_input(sh, 45 + i, dt) for i, sh in enumerate(shapes)]
    packed = np.concatenate([x.ravel() for x in xs])
    assert packed.size == bp.total
    a = api.DeviceBuffer(bp.total, prec, L)
    w = api.DeviceBuffer(bp.total, prec, L)
    o = a if inplace else api.DeviceBuffer(bp.total, prec, L)
    a.upload(packed)
    bp.set_buffers(a.ptr, o.ptr, w.ptr)
    bp.execute(direction, use_graph)
    bp.synchronize()
    y = o.download()
    return bp, xs, y


def test_smooth7_lengths():
    s = smooth7()
    assert len(s) == 64 and s[:5] == [16, 18, 20, 21, 24] and s[-3:] == [250, 252, 256]


def test_batch_mixed_shapes_vs_oracle(emu_lib):
    shapes = draw_shapes(12, 1, lo=2, hi=40) + [(1, 1), (7,), (5, 1, 3), (47, 3)]
    for prec in (0, 1):
        bp, xs, y = _run(emu_lib, shapes, prec)
        for i, (sh, x) in enumerate(zip(shapes, xs)):
            got = y[bp.offsets[i]:bp.offsets[i] + x.size].reshape(sh)
            ref = oracle.dftn(x.astype(np.complex128))
            assert oracle.check_rel_error(got, ref, 1e-5 if prec == 0 else 1e-12), (sh, prec)


def test_batch_backward_inplace_and_direct_launch(emu_lib):
    shapes = [(8, 12), (6, 5, 4), (16,), (3, 9, 2)]
    for inplace, use_graph in ((True, True), (False, False)):
        bp, xs, y = _run(emu_lib, shapes, 1, inplace=inplace, use_graph=use_graph, direction=1)
        for i, (sh, x) in enumerate(zip(shapes, xs)):
            got = y[bp.offsets[i]:bp.offsets[i] + x.size].reshape(sh)
            assert oracle.rel_l2(got, oracle.dftn(x, None, +1)) < 1e-13


def test_batch_prime_axes_both_directions(emu_lib):
    """Prime axes inside a batch run the in-LDS Bluestein pass of the batch
    kernel with its direction-dependent chirp and filter tables.  (Axes
    longer than one LDS pass are refused by the batch planner.)"""
    shapes = [(1009,), (2, 1031), (8, 12)]
    for direction, sign in ((0, -1), (1, +1)):
        bp, xs, y = _run(emu_lib, shapes, 1, direction=direction)
        for i, (sh, x) in enumerate(zip(shapes, xs)):
            got = y[bp.offsets[i]:bp.offsets[i] + x.size].reshape(sh)
            assert oracle.check_rel_error(got, oracle.dftn(x, None, sign), 1e-12), (sh, direction)


def test_batch_graph_replays_match(emu_lib):
    L = emu_lib
    shapes = draw_shapes(6, 3, lo=4, hi=32)
    bp = api.BatchPlan(shapes, 0, L=L)
    x = oracle.uniform_input((bp.total,), 9, np.complex64)
    a, o, w = (api.DeviceBuffer(bp.total, 0, L) for _ in range(3))
    a.upload(x)
    bp.set_buffers(a.ptr, o.ptr, w.ptr)
    bp.execute(0, False)
    ref = o.download()
    for _ in range(3):
        o.upload(np.zeros(bp.total, np.complex64))
        bp.execute(0, True)
        assert np.array_equal(o.download(), ref)
    info = bp.info()
    assert info["passes"] == sum(sum(1 for d in sh if d > 1) for sh in shapes)
