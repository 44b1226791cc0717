"""Transform server host logic, run through the CPU emulation build.

The emulated launches complete synchronously, so these tests cover the
request loop's bookkeeping (plan cache, arena admission, job states, error
paths, statistics) and that a served job computes exactly what the plain
backend call computes; overlap on a real device is covered by the GPU test.
The concurrent-client check also runs on the GPU (marked gpu).
"""
import numpy as np
import pytest

import oracle
from shafft_amd import api
from shafft_amd.workloads import balanced_dims, draw_stream, factor7, smooth7_between

TOL = {0: 1e-5, 1: 1e-12}
DT = {0: np.complex64, 1: np.complex128}

# (kind, dims or n, prec, direction, normalize)
MIX = [
    (0, (16, 12), 0, 0, False), (0, (8, 6, 10), 1, 1, True), (1, 1024, 0, 1, False),
    (1, 360, 1, 0, True), (0, (64,), 1, 0, False), (1, 97, 0, 0, False),
    (0, (5, 7, 9), 0, 1, False), (1, 4096, 1, 1, False), (0, (16, 12), 0, 0, False),
    (0, (30, 28), 1, 0, True), (1, 12000, 0, 0, False), (0, (3, 250), 0, 1, True),
]


def _shape(job):
    kind, d = job[0], job[1]
    return (d,) if kind == 1 else tuple(d)


def _inputs(L):
    xs, bufs = [], []
    for i, job in enumerate(MIX):
        x = oracle.uniform_input(_shape(job), 100 + i, DT[job[2]])
        b = api.DeviceBuffer(x.size, job[2], L)
        b.upload(x.ravel())
        xs.append(x)
        bufs.append(b)
    return xs, bufs


def _submit(srv, job, src, dst):
    kind, d, prec, direction, norm = job
    if kind == 1:
        return srv.submit_1d(d, prec, direction, src, dst, normalize=norm)
    return srv.submit_nd(d, prec, direction, src, dst, normalize=norm)


def _reference(x, job):
    kind, d, prec, direction, norm = job
    y = oracle.dftn(x.astype(np.complex128), list(range(x.ndim)), -1 if direction == 0 else 1)
    return y / np.sqrt(x.size) if norm else y


def test_served_jobs_match_oracle_and_plain_calls(emu_lib):
    L = emu_lib
    xs, ins = _inputs(L)
    srv = api.Server(slots=4, arena_bytes=8 << 20, planners=2, L=L)
    outs = [api.DeviceBuffer(x.size, job[2], L) for x, job in zip(xs, MIX)]
    ids = [_submit(srv, job, i.ptr, o.ptr) for job, i, o in zip(MIX, ins, outs)]
    for k, jid in enumerate(ids):
        r = srv.wait(jid, timeout_s=120)
        assert r is not None and r.status == 0, (MIX[k], r and r.message)
        assert r.arrival <= r.admit <= r.done
    for k, (x, job) in enumerate(zip(xs, MIX)):
        y = api.download_ptr(outs[k].ptr, x.size, job[2], L).reshape(x.shape)
        ref = _reference(x, job)
        assert oracle.check_rel_error(y, ref, TOL[job[2]]), (job, oracle.rel_l2(y, ref))
        # input preserved
        xi = api.download_ptr(ins[k].ptr, x.size, job[2], L).reshape(x.shape)
        assert np.array_equal(xi, x)
        # the same bits as the plain backend call on the same shape
        if job[0] == 0 and not job[4]:
            p = api.NDBackendPlan(list(job[1]), list(range(len(job[1]))), job[2], L=L)
            a, b = api.DeviceBuffer(x.size, job[2], L), api.DeviceBuffer(x.size, job[2], L)
            a.upload(x.ravel())
            dd, _ = p.execute(a.ptr, b.ptr, job[3])
            z = api.download_ptr(dd, x.size, job[2], L).reshape(x.shape)
            assert np.array_equal(z, y), job
        if job[0] == 1 and not job[4]:
            p = api.OneDBackendPlan(job[1], job[2], L=L)
            b = api.DeviceBuffer(x.size, job[2], L)
            p.execute(ins[k].ptr, b.ptr, job[3])
            z = api.download_ptr(b.ptr, x.size, job[2], L)
            assert np.array_equal(z, y.ravel()), job
    st = srv.stats()
    assert st["completed"] == len(MIX) and st["failed"] == 0
    assert st["plans"] == len({(j[0], _shape(j), j[2]) for j in MIX})  # repeated shape reuses its plan
    # every job up to 2^20 elements without normalize ran micro-batched
    assert srv.batch_stats()[1] == sum(1 for j in MIX if not j[4])
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["arena_peak"] <= st["arena_bytes"]
    want = sum(5 * x.size * np.log2(x.size) for x in xs) * 1e-9
    assert abs(st["gflop"] - want) < 1e-9 * max(1.0, want)
    srv.close()


def test_alone_equals_mixed(emu_lib):
    L = emu_lib
    xs, ins = _inputs(L)
    alone = api.Server(slots=1, arena_bytes=8 << 20, planners=1, L=L)
    mixed = api.Server(slots=6, arena_bytes=8 << 20, planners=3, L=L)
    oa = [api.DeviceBuffer(x.size, j[2], L) for x, j in zip(xs, MIX)]
    om = [api.DeviceBuffer(x.size, j[2], L) for x, j in zip(xs, MIX)]
    for job, i, o in zip(MIX, ins, oa):
        assert alone.wait(_submit(alone, job, i.ptr, o.ptr), 120).status == 0
    ids = [_submit(mixed, job, i.ptr, o.ptr) for job, i, o in zip(MIX, ins, om)]
    for jid in ids:
        assert mixed.wait(jid, 120).status == 0
    for x, job, a, m in zip(xs, MIX, oa, om):
        ya = api.download_ptr(a.ptr, x.size, job[2], L)
        ym = api.download_ptr(m.ptr, x.size, job[2], L)
        assert np.array_equal(ya, ym), job


def test_micro_batches_group_jobs(emu_lib):
    """Jobs pending together are grouped per precision, at most max_jobs a
    group; with micro-batching off every job runs its own launches; the
    results are the same either way (the emulation runs both on the host's
    generic pass)."""
    L = emu_lib
    xs, ins = _inputs(L)
    outs = {}
    for mode in ("grouped", "off"):
        srv = api.Server(slots=2, arena_bytes=8 << 20, planners=2, L=L)
        if mode == "off":
            srv.set_batching(0)
        else:
            srv.set_batching(1 << 20, 3)
        # plans first, then the whole mix at once so it is pending together
        warm = [api.DeviceBuffer(x.size, j[2], L) for x, j in zip(xs, MIX)]
        for job, i, o in zip(MIX, ins, warm):
            _submit(srv, job, i.ptr, o.ptr)
        srv.drain()
        srv.reset_stats()
        outs[mode] = [api.DeviceBuffer(x.size, j[2], L) for x, j in zip(xs, MIX)]
        ids = [_submit(srv, job, i.ptr, o.ptr) for job, i, o in zip(MIX, ins, outs[mode])]
        for jid in ids:
            assert srv.wait(jid, 120).status == 0
        b, nj = srv.batch_stats()
        if mode == "off":
            assert (b, nj) == (0, 0)
        else:
            assert nj == sum(1 for j in MIX if not j[4])
            assert b >= -(-nj // 3)  # at most 3 jobs a group
        srv.close()
    for x, job, a, m in zip(xs, MIX, outs["grouped"], outs["off"]):
        assert np.array_equal(api.download_ptr(a.ptr, x.size, job[2], L),
                              api.download_ptr(m.ptr, x.size, job[2], L)), job


@pytest.mark.parametrize("max_jobs,max_inflight", [(1, 1), (2, 1), (64, 2), (5, 8)])
def test_micro_batch_limits(emu_lib, max_jobs, max_inflight):
    """Every combination of group size and groups in flight serves the whole
    mix (normalised jobs, which never group, included) with the results of
    the server with micro-batching off; no group exceeds max_jobs."""
    L = emu_lib
    xs, ins = _inputs(L)
    res = {}
    for mode in ("off", "on"):
        srv = api.Server(slots=3, arena_bytes=8 << 20, planners=2, L=L)
        if mode == "off":
            srv.set_batching(0)
        else:
            srv.set_batching(1 << 20, max_jobs, max_inflight)
        outs = [api.DeviceBuffer(x.size, j[2], L) for x, j in zip(xs, MIX)]
        ids = [_submit(srv, job, i.ptr, o.ptr) for job, i, o in zip(MIX, ins, outs)]
        for jid in ids:
            r = srv.wait(jid, 120)
            assert r is not None and r.status == 0, r and r.message
        b, nj = srv.batch_stats()
        if mode == "on":
            assert nj == sum(1 for j in MIX if not j[4])
            assert b * max_jobs >= nj
        srv.close()
        res[mode] = [api.download_ptr(o.ptr, x.size, j[2], L) for o, x, j in zip(outs, xs, MIX)]
    for a, m, job in zip(res["off"], res["on"], MIX):
        assert np.array_equal(a, m), job


def test_micro_batch_limits_rejected(emu_lib):
    srv = api.Server(slots=1, arena_bytes=1 << 20, planners=1, L=emu_lib)
    for args in ((-1, 64, 2), (1 << 20, 0, 2), (1 << 20, 64, 0)):
        with pytest.raises(api.ShafftError):
            srv.set_batching(*args)
    srv.close()


def test_arena_admission(emu_lib):
    L = emu_lib
    # (64,64) complex128 needs one 64 KiB scratch block; (4096,) 1-D in one pass needs none.
    big = np.ones((64, 64), np.complex128)
    small = np.ones(4096, np.complex64)
    bi, bo = api.DeviceBuffer(big.size, 1, L), api.DeviceBuffer(big.size, 1, L)
    si, so = api.DeviceBuffer(small.size, 0, L), api.DeviceBuffer(small.size, 0, L)
    bi.upload(big.ravel())
    si.upload(small)
    tiny = api.Server(slots=2, arena_bytes=4096, planners=1, L=L)
    j1 = tiny.submit_nd((64, 64), 1, 0, bi.ptr, bo.ptr)
    j2 = tiny.submit_1d(4096, 0, 0, si.ptr, so.ptr)
    r1, r2 = tiny.wait(j1, 120), tiny.wait(j2, 120)
    assert r1.status == 9 and "exceeds the server arena" in r1.message  # SHAFFT_ERR_ALLOC
    assert r2.status == 0
    y = api.download_ptr(so.ptr, small.size, 0, L)
    assert abs(y[0] - 4096) < 1e-3 and np.max(np.abs(y[1:])) < 1e-3
    st = tiny.stats()
    assert st["completed"] == 1 and st["failed"] == 1
    # An arena that holds exactly one scratch block serialises the big jobs.
    one = api.Server(slots=4, arena_bytes=64 * 64 * 16, planners=1, L=L)
    outs = [api.DeviceBuffer(big.size, 1, L) for _ in range(5)]
    ids = [one.submit_nd((64, 64), 1, k % 2, bi.ptr, o.ptr) for k, o in enumerate(outs)]
    for jid in ids:
        assert one.wait(jid, 120).status == 0
    st = one.stats()
    assert st["arena_peak"] == 64 * 64 * 16 and st["plans"] == 1 and st["plan_builds"] == 1


def test_server_errors(emu_lib):
    L = emu_lib
    srv = api.Server(slots=1, arena_bytes=1 << 20, planners=1, L=L)
    b = api.DeviceBuffer(64, 0, L)
    c = api.DeviceBuffer(64, 0, L)
    with pytest.raises(api.ShafftError):
        srv.submit_1d(64, 0, 0, b.ptr, b.ptr)  # in == out
    with pytest.raises(api.ShafftError):
        srv.submit_1d(64, 2, 0, b.ptr, c.ptr)  # bad precision
    with pytest.raises(api.ShafftError):
        srv.submit_nd((8, 0), 0, 0, b.ptr, c.ptr)  # zero dim
    with pytest.raises(api.ShafftError):
        srv.submit_1d(64, 0, 0, 0, c.ptr)  # null input
    with pytest.raises(api.ShafftError):
        srv.wait(12345, 0.1)  # unknown id
    jid = srv.submit_1d(64, 0, 1, b.ptr, c.ptr)
    assert srv.wait(jid, 120).status == 0
    with pytest.raises(api.ShafftError):
        srv.finished(jid)  # released by wait
    with pytest.raises(api.ShafftError):
        api.Server(slots=0, L=L)
    srv.drain()
    srv.close()


def test_stream_generator():
    s = smooth7_between(1 << 10, 1 << 14)
    assert all(factor7(v) is not None for v in s) and s[0] == 1024 and s[-1] == 16384
    assert 1029 in s and 1031 not in s  # 3*7^3 is 7-smooth, 1031 is prime
    assert balanced_dims(1 << 26, 2) == [8192, 8192]
    assert balanced_dims(2 * 3 * 5 * 7 * 64, 3) == [28, 24, 20]
    reqs = draw_stream(4000, seed=46)
    assert len(reqs) == 4000
    assert all(1 << 9 < r.elems < 1 << 27 for r in reqs)
    frac128 = np.mean([r.prec for r in reqs])
    assert 0.26 < frac128 < 0.34
    assert 0.45 < np.mean([r.kind for r in reqs]) < 0.55
    gaps = np.diff([r.t for r in reqs])
    assert 1 / 2300 < gaps.mean() < 1 / 1700
    logs = np.log2([r.elems for r in reqs])
    assert abs(np.median(logs) - 18) < 0.6  # log-uniform on [10, 26]
    for r in reqs[:200]:
        if r.kind == 0:
            assert int(np.prod(r.dims)) == r.elems and len(r.dims) in (2, 3)


def concurrent_clients(L, clients=4, rounds=3):
    """Several host threads submit and wait on one server at once (the
    submit/wait entry points are called without Python's lock held); every
    client's results equal those of one client on a one-slot server, bit for
    bit, and the statistics count every job once."""
    import threading

    xs, ins = _inputs(L)
    alone = api.Server(slots=1, arena_bytes=8 << 20, planners=1, L=L)
    ref = []
    for job, i, x in zip(MIX, ins, xs):
        o = api.DeviceBuffer(x.size, job[2], L)
        assert alone.wait(_submit(alone, job, i.ptr, o.ptr), 120).status == 0
        ref.append(api.download_ptr(o.ptr, x.size, job[2], L))
    alone.close()
    srv = api.Server(slots=4, arena_bytes=64 << 20, planners=2, L=L)
    outs = [[[api.DeviceBuffer(x.size, j[2], L) for x, j in zip(xs, MIX)] for _ in range(rounds)]
            for _ in range(clients)]
    errors = []

    def client(c):
        try:
            for r in range(rounds):
                order = list(range(len(MIX)))[c % len(MIX):] + list(range(len(MIX)))[:c % len(MIX)]
                ids = [(k, _submit(srv, MIX[k], ins[k].ptr, outs[c][r][k].ptr)) for k in order]
                for k, jid in ids:
                    res = srv.wait(jid, 120)
                    if res is None or res.status != 0:
                        errors.append((c, r, MIX[k], res and res.message))
        except Exception as e:  # reported below, in the main thread
            errors.append((c, repr(e)))

    ts = [threading.Thread(target=client, args=(c,)) for c in range(clients)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(300)
    assert not any(t.is_alive() for t in ts), "a client did not finish"
    assert not errors, errors
    for c in range(clients):
        for r in range(rounds):
            for k, (x, job) in enumerate(zip(xs, MIX)):
                y = api.download_ptr(outs[c][r][k].ptr, x.size, job[2], L)
                assert np.array_equal(y, ref[k]), (c, r, job)
    st = srv.stats()
    assert st["completed"] == clients * rounds * len(MIX) and st["failed"] == 0
    srv.close()


def test_concurrent_clients(emu_lib):
    concurrent_clients(emu_lib)


@pytest.mark.gpu
def test_concurrent_clients_gpu(gpu_lib):
    concurrent_clients(None, clients=8, rounds=4)
