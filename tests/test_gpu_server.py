"""GPU tests of the transform server: every served job gives the same bits
alone (one slot, one job at a time) and in a mix whose jobs overlap on the
device, and the oracle's result within the precision's tolerance (direct DFT
for the smaller jobs; round trip for the larger ones).  Jobs that run their
own launches give the same bits as the plain backend call; micro-batched
jobs (up to 2^20 elements, run by the batch kernel) equal it within the
precision's tolerance, and bit for bit with micro-batching switched off."""
import numpy as np
import pytest

import oracle
from shafft_amd import api
from shafft_amd.workloads import draw_stream

pytestmark = pytest.mark.gpu
TOL = {0: 1e-5, 1: 1e-12}
DT = {0: np.complex64, 1: np.complex128}

# (kind, dims or n, prec, direction, normalize)
MIX = [
    (0, (512, 512), 1, 0, False), (1, 1 << 20, 0, 1, False), (0, (64, 64, 32), 0, 0, True),
    (1, 1000, 1, 0, False), (0, (96, 80, 70), 1, 1, False), (1, 40009, 0, 0, False),
    (0, (2048, 1024), 0, 1, False), (1, 360, 0, 1, True), (0, (64, 64, 32), 0, 0, True),
    (1, 1 << 21, 1, 0, False), (0, (30, 28, 25), 0, 0, False), (1, 12000, 1, 1, True),
    (0, (4096, 256), 1, 0, False), (0, (7, 11, 13), 1, 1, False), (1, 3 * 5 * 7 * 512, 0, 0, False),
    (0, (128, 128, 128), 0, 1, False), (1, 4096, 1, 0, False), (0, (250, 240), 0, 0, True),
]
ORACLE_MAX = 1 << 16
MB_MAX = 1 << 20  # the server's default micro-batch limit (elements)


def _batchable(elems, normalize):
    return elems <= MB_MAX and not normalize


def _shape(job):
    return (job[1],) if job[0] == 1 else tuple(job[1])


def _submit(srv, job, src, dst):
    kind, d, prec, direction, norm = job
    if kind == 1:
        return srv.submit_1d(d, prec, direction, src, dst, normalize=norm)
    return srv.submit_nd(d, prec, direction, src, dst, normalize=norm)


@pytest.fixture(scope="module")
def served(gpu_lib):
    xs, ins, oa, om = [], [], [], []
    for i, job in enumerate(MIX):
        x = oracle.uniform_input(_shape(job), 46 + i, DT[job[2]])
        b = api.DeviceBuffer(x.size, job[2])
        b.upload(x.ravel())
        xs.append(x)
        ins.append(b)
        oa.append(api.DeviceBuffer(x.size, job[2]))
        om.append(api.DeviceBuffer(x.size, job[2]))
    alone = api.Server(slots=1, arena_bytes=1 << 30, planners=1)
    for job, i, o in zip(MIX, ins, oa):
        r = alone.wait(_submit(alone, job, i.ptr, o.ptr), 60)
        assert r is not None and r.status == 0, (job, r and r.message)
    alone.close()
    mixed = api.Server(slots=8, arena_bytes=1 << 30, planners=2)
    # Build the plans first so the mix itself runs from resident plans.
    warm = [api.DeviceBuffer(x.size, j[2]) for x, j in zip(xs, MIX)]
    for job, i, o in zip(MIX, ins, warm):
        _submit(mixed, job, i.ptr, o.ptr)
    mixed.drain()
    mixed.reset_stats()
    ids = [_submit(mixed, job, i.ptr, o.ptr) for job, i, o in zip(MIX, ins, om)]
    res = [mixed.wait(j, 60) for j in ids]
    stats = mixed.stats()
    stats["batches"], stats["batched_jobs"] = mixed.batch_stats()
    mixed.close()
    ya = [api.download_ptr(o.ptr, x.size, j[2]).reshape(x.shape) for o, x, j in zip(oa, xs, MIX)]
    ym = [api.download_ptr(o.ptr, x.size, j[2]).reshape(x.shape) for o, x, j in zip(om, xs, MIX)]
    kept = [np.array_equal(api.download_ptr(i.ptr, x.size, j[2]).reshape(x.shape), x)
            for i, x, j in zip(ins, xs, MIX)]
    return xs, ya, ym, res, stats, kept


def test_alone_equals_mixed(served):
    xs, ya, ym, res, stats, kept = served
    for job, r in zip(MIX, res):
        assert r is not None and r.status == 0, (job, r and r.message)
    for job, a, m in zip(MIX, ya, ym):
        assert np.array_equal(a, m), job
    assert all(kept)
    assert stats["completed"] == len(MIX) and stats["failed"] == 0
    # the small jobs ran micro-batched (how many share a group depends on
    # how many are pending when the scheduler admits)
    print(f"micro-batches {stats['batches']}, jobs in them {stats['batched_jobs']}")
    assert stats["batched_jobs"] >= 6 and stats["batches"] >= 1, stats


def test_mix_overlaps_on_device(served):
    _, _, _, res, stats, _ = served
    ev = sorted([(r.admit, 1) for r in res] + [(r.done, -1) for r in res], key=lambda e: (e[0], e[1]))
    cur = peak = 0
    for _, d in ev:
        cur += d
        peak = max(peak, cur)
    print(f"peak concurrent jobs {peak}; p50 {stats['p50_ms']:.3f} ms p99 {stats['p99_ms']:.3f} ms")
    assert peak >= 2


def test_served_vs_oracle(served):
    xs, ya, _, _, _, _ = served
    for x, y, job in zip(xs, ya, MIX):
        if x.size > ORACLE_MAX:
            continue
        sign = -1 if job[3] == 0 else 1
        ref = oracle.dftn(x.astype(np.complex128), list(range(x.ndim)), sign)
        if job[4]:
            ref = ref / np.sqrt(x.size)
        assert oracle.check_rel_error(y, ref, TOL[job[2]]), (job, oracle.rel_l2(y, ref))


def _plain_mix(job, x):
    prec = job[2]
    if job[0] == 0:
        p = api.NDBackendPlan(list(job[1]), list(range(len(job[1]))), prec)
        a, b = api.DeviceBuffer(x.size, prec), api.DeviceBuffer(x.size, prec)
        a.upload(x.ravel())
        d, _ = p.execute(a.ptr, b.ptr, job[3])
        return api.download_ptr(d, x.size, prec).reshape(x.shape)
    p = api.OneDBackendPlan(job[1], prec)
    a, b = api.DeviceBuffer(x.size, prec), api.DeviceBuffer(x.size, prec)
    a.upload(x.ravel())
    p.execute(a.ptr, b.ptr, job[3])
    return api.download_ptr(b.ptr, x.size, prec).reshape(x.shape)


def test_served_vs_plain_calls(served):
    xs, ya, _, _, _, _ = served
    for x, y, job in zip(xs, ya, MIX):
        if job[4]:
            continue
        z = _plain_mix(job, x)
        if _batchable(x.size, job[4]):
            assert oracle.check_rel_error(y, z, TOL[job[2]]), (job, oracle.rel_l2(y, z))
        else:
            assert np.array_equal(z, y), job


def test_unbatched_mix_equals_plain_calls(gpu_lib):
    """Micro-batching off: every job of the mix runs its own launches, in
    overlap on 8 slots, and gives the plain backend call's bits."""
    srv = api.Server(slots=8, arena_bytes=1 << 30, planners=2)
    srv.set_batching(0)
    xs, ins, outs, ids = [], [], [], []
    for i, job in enumerate(MIX):
        x = oracle.uniform_input(_shape(job), 46 + i, DT[job[2]])
        b = api.DeviceBuffer(x.size, job[2])
        b.upload(x.ravel())
        xs.append(x)
        ins.append(b)
        outs.append(api.DeviceBuffer(x.size, job[2]))
    for job, i, o in zip(MIX, ins, outs):
        ids.append(_submit(srv, job, i.ptr, o.ptr))
    for jid, job in zip(ids, MIX):
        r = srv.wait(jid, 60)
        assert r is not None and r.status == 0, (job, r and r.message)
    assert srv.batch_stats() == (0, 0)
    srv.close()
    for x, job, o in zip(xs, MIX, outs):
        if job[4]:
            continue
        y = api.download_ptr(o.ptr, x.size, job[2]).reshape(x.shape)
        assert np.array_equal(_plain_mix(job, x), y), job


def test_served_round_trip(gpu_lib):
    """Large jobs: forward then backward (normalised) through the server
    returns the input within the precision's tolerance."""
    srv = api.Server(slots=4, arena_bytes=4 << 30, planners=2)
    for k, r in enumerate(draw_stream(12, seed=46, lo=20, hi=24)):
        shape = (r.n,) if r.kind == 1 else tuple(r.dims)
        x = oracle.uniform_input(shape, 460 + k, DT[r.prec])
        a, b, c = (api.DeviceBuffer(x.size, r.prec) for _ in range(3))
        a.upload(x.ravel())
        sub = srv.submit_1d if r.kind == 1 else srv.submit_nd
        arg = r.n if r.kind == 1 else r.dims
        j1 = sub(arg, r.prec, 0, a.ptr, b.ptr, normalize=True)
        assert srv.wait(j1, 60).status == 0
        j2 = sub(arg, r.prec, 1, b.ptr, c.ptr, normalize=True)
        assert srv.wait(j2, 60).status == 0
        y = api.download_ptr(c.ptr, x.size, r.prec).reshape(shape)
        assert oracle.check_rel_error(y, x, TOL[r.prec]), (r.describe(), oracle.rel_l2(y, x))
    srv.close()


def _plain(r, x):
    """The plain backend call on a copy of x (execute may overwrite its input)."""
    shape = (r.n,) if r.kind == 1 else tuple(r.dims)
    if r.kind == 1:
        p = api.OneDBackendPlan(r.n, r.prec)
        a, b = api.DeviceBuffer(x.size, r.prec), api.DeviceBuffer(x.size, r.prec)
        a.upload(x.ravel())
        p.execute(a.ptr, b.ptr, r.direction)
        p.synchronize()
        return api.download_ptr(b.ptr, x.size, r.prec).reshape(shape)
    p = api.NDBackendPlan(list(r.dims), list(range(len(r.dims))), r.prec)
    a, b = api.DeviceBuffer(x.size, r.prec), api.DeviceBuffer(x.size, r.prec)
    a.upload(x.ravel())
    d, _ = p.execute(a.ptr, b.ptr, r.direction)
    return api.download_ptr(d, x.size, r.prec).reshape(shape)


def test_served_c5_stream_paced(gpu_lib):
    """A draw of the C5 serving stream (Poisson arrivals at 2000 jobs/s,
    sizes 2^10..2^26, 70% complex64) served at its arrival times, with the
    2^25..2^26-element jobs of both precisions in it: every result bit for bit
    the plain backend call on the same input (within TOL for the
    micro-batched jobs up to 2^20 elements), and every job's inverse (served
    too, normalised) back to its input within TOL."""
    import time

    reqs = draw_stream(600, seed=46)
    big = {p: [r for r in reqs if r.prec == p and r.elems >= 1 << 25][:3] for p in (0, 1)}
    assert all(len(v) >= 2 for v in big.values()), "stream draw lacks 2^25+ jobs of a precision"
    jobs = sorted(big[0] + big[1] + reqs[:24], key=lambda r: r.t)
    assert max(r.elems for r in jobs) > 60_000_000
    xs, ins, outs = [], [], []
    for k, r in enumerate(jobs):
        shape = (r.n,) if r.kind == 1 else tuple(r.dims)
        x = oracle.uniform_input(shape, 4600 + k, DT[r.prec])
        b = api.DeviceBuffer(x.size, r.prec)
        b.upload(x.ravel())
        xs.append(x)
        ins.append(b)
        outs.append(api.DeviceBuffer(x.size, r.prec))
    srv = api.Server(slots=8, arena_bytes=24 << 30, planners=2)
    # plans first, so the paced phase runs from resident plans
    for r, i, o in zip(jobs, ins, outs):
        sub = srv.submit_1d if r.kind == 1 else srv.submit_nd
        sub(r.n if r.kind == 1 else r.dims, r.prec, r.direction, i.ptr, o.ptr)
    srv.drain()
    t0 = api.Server.now_ns() + 2_000_000
    base = jobs[0].t
    ids = []
    for r, i, o in zip(jobs, ins, outs):
        arr = t0 + int((r.t - base) * 1e9)
        while time.monotonic_ns() < arr:
            pass
        sub = srv.submit_1d if r.kind == 1 else srv.submit_nd
        ids.append(sub(r.n if r.kind == 1 else r.dims, r.prec, r.direction, i.ptr, o.ptr, arrival_ns=arr))
    res = [srv.wait(j, 120) for j in ids]
    for r, s in zip(jobs, res):
        assert s is not None and s.status == 0, (r.describe(), s and s.message)
    lat = sorted(s.latency_ms for s in res)
    print(f"paced C5 draw: {len(jobs)} jobs, largest {max(r.elems for r in jobs)} elements, "
          f"p50 {lat[len(lat) // 2]:.3f} ms, max {lat[-1]:.3f} ms", flush=True)
    backs = [api.DeviceBuffer(x.size, r.prec) for x, r in zip(xs, jobs)]
    inv = []
    for r, o, b in zip(jobs, outs, backs):
        sub = srv.submit_1d if r.kind == 1 else srv.submit_nd
        inv.append(sub(r.n if r.kind == 1 else r.dims, r.prec, 1 - r.direction, o.ptr, b.ptr, normalize=True))
    for j in inv:
        assert srv.wait(j, 120).status == 0
    srv.close()
    for r, x, o, b in zip(jobs, xs, outs, backs):
        y = api.download_ptr(o.ptr, x.size, r.prec).reshape(x.shape)
        z0 = _plain(r, x)
        if _batchable(x.size, False):
            assert oracle.check_rel_error(y, z0, TOL[r.prec]), (r.describe(), oracle.rel_l2(y, z0))
        else:
            assert np.array_equal(y, z0), r.describe()
        z = api.download_ptr(b.ptr, x.size, r.prec).reshape(x.shape) / np.sqrt(x.size)
        assert oracle.check_rel_error(z, x, TOL[r.prec]), (r.describe(), oracle.rel_l2(z, x))


def test_micro_batched_stream_vs_fp64(gpu_lib):
    """300 jobs of the serving stream's generator with sizes 2^10..2^16
    (both precisions, FFTND 2-D/3-D and FFT1D, both directions), queued at
    once so they run micro-batched several to a group: every result against
    an FP64 FFT of its input (numpy's pocketfft, pinned against the direct
    DFT in test_oracle) within the precision's tolerance, every input
    unchanged."""
    reqs = draw_stream(300, seed=47, lo=10, hi=16)
    srv = api.Server(slots=4, arena_bytes=1 << 30, planners=4)
    xs, ins, outs, ids = [], [], [], []
    for k, r in enumerate(reqs):
        shape = (r.n,) if r.kind == 1 else tuple(r.dims)
        x = oracle.uniform_input(shape, 7000 + k, DT[r.prec])
        a, b = api.DeviceBuffer(x.size, r.prec), api.DeviceBuffer(x.size, r.prec)
        a.upload(x.ravel())
        xs.append(x)
        ins.append(a)
        outs.append(b)
    # plans first, then every job at once
    for r, i, o in zip(reqs, ins, outs):
        sub = srv.submit_1d if r.kind == 1 else srv.submit_nd
        sub(r.n if r.kind == 1 else r.dims, r.prec, r.direction, i.ptr, o.ptr)
    srv.drain()
    srv.reset_stats()
    for r, i, o in zip(reqs, ins, outs):
        sub = srv.submit_1d if r.kind == 1 else srv.submit_nd
        ids.append(sub(r.n if r.kind == 1 else r.dims, r.prec, r.direction, i.ptr, o.ptr))
    for r, j in zip(reqs, ids):
        s = srv.wait(j, 60)
        assert s is not None and s.status == 0, (r.describe(), s and s.message)
    groups, grouped = srv.batch_stats()
    srv.close()
    print(f"{len(reqs)} jobs, {grouped} micro-batched in {groups} groups", flush=True)
    assert grouped == len(reqs) and groups < len(reqs)
    worst = {0: 0.0, 1: 0.0}
    for r, x, i, o in zip(reqs, xs, ins, outs):
        y = api.download_ptr(o.ptr, x.size, r.prec).reshape(x.shape)
        xd = x.astype(np.complex128)
        ref = np.fft.fftn(xd) if r.direction == 0 else np.fft.ifftn(xd) * x.size
        e = oracle.rel_l2(y, ref)
        worst[r.prec] = max(worst[r.prec], e)
        assert oracle.check_rel_error(y, ref, TOL[r.prec]), (r.describe(), e)
        assert np.array_equal(api.download_ptr(i.ptr, x.size, r.prec).reshape(x.shape), x), r.describe()
    print(f"worst relL2: complex64 {worst[0]:.2e}, complex128 {worst[1]:.2e}", flush=True)
