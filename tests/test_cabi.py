"""The product library loads without a GPU and exports every entry point
declared in include/shafft_amd.h.  No compute call is made.  Without the
library every entry point raises."""
import ctypes as ct
import os
import re

import pytest

from shafft_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "shafft_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(shafft\w*)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    L = ct.CDLL(api.library_path())
    names = declared_functions()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    names = set(declared_functions())
    assert names == set(api._SIGS), names ^ set(api._SIGS)


def test_host_only_calls():
    L = api.bind(api.library_path())
    maj, mi, pa = ct.c_int(), ct.c_int(), ct.c_int()
    L.shafftGetVersion(ct.byref(maj), ct.byref(mi), ct.byref(pa))
    assert (maj.value, mi.value) == (1, 1)
    assert L.shafftGetBackendName() == b"gfx950"
    ln, ls, la, lnt, lst = (ct.c_size_t() for _ in range(5))
    assert L.shafftAmd1DQueryLayout(12345, ct.byref(ln), ct.byref(ls), ct.byref(la), ct.byref(lnt), ct.byref(lst)) == 0
    assert (ln.value, ls.value, la.value, lnt.value, lst.value) == (12345, 0, 12345, 12345, 0)
    assert L.shafftAmd1DQueryLayout(0, ct.byref(ln), ct.byref(ls), ct.byref(la), None, None) == 5
    h = ct.c_void_p()
    assert L.shafftNDCreate(ct.byref(h)) == 0
    assert L.shafftExecute(h, 0) == 4
    buf = ct.create_string_buffer(256)
    assert L.shafftLastErrorMessage(buf, 256) > 0
    assert L.shafftDestroy(ct.byref(h)) == 0


def test_fortran_config_accessors():
    """The Fortran module's config path (src/f03/shafftf03.h:46-146) at one
    rank: heap config, Init through the Fortran handle, field accessors,
    plan from the config, and the handle check.  Host-only calls."""
    L = api.bind(api.library_path())
    cfg = L.shafftConfigNDAllocf03()
    assert cfg
    shape = (ct.c_size_t * 3)(40, 30, 20)
    zero = ct.c_int(0)
    assert L.shafftConfigNDInitf03(cfg, 3, shape, api.Z2Z, None, 0, 0, 1, 0, ct.byref(zero)) == 0
    assert L.shafftConfigNDGetNdim(cfg) == 3 and L.shafftConfigNDGetPrecision(cfg) == api.Z2Z
    assert L.shafftConfigNDGetAllocElements(cfg) == 40 * 30 * 20
    assert L.shafftConfigNDGetNda(cfg) == 0 and L.shafftConfigNDGetCommSize(cfg) == 1
    assert L.shafftConfigNDGetIsActive(cfg) == 1 and L.shafftConfigNDGetActiveRank(cfg) == 0
    assert L.shafftConfigNDGetOutputPolicy(cfg) == 1 and L.shafftConfigNDGetWorldComm(cfg) == 0
    assert L.shafftConfigNDGetFlags(cfg) & 8  # RESOLVED
    sub, off, grid = (ct.c_size_t * 3)(), (ct.c_size_t * 3)(9, 9, 9), (ct.c_int * 3)()
    L.shafftConfigNDGetInitialSubsize(cfg, sub)
    L.shafftConfigNDGetOutputOffset(cfg, off)
    L.shafftConfigNDGetCommDims(cfg, grid)
    assert list(sub) == [40, 30, 20] and list(off) == [0, 0, 0] and list(grid) == [1, 1, 1]
    n = L.shafftConfigNDGetHostnameLen(cfg)
    buf = ct.create_string_buffer(n + 1)
    L.shafftConfigNDGetHostname(cfg, buf, n + 1)
    assert n > 0 and len(buf.value) == n
    L.shafftConfigNDSetHintNda(cfg, 1)
    assert L.shafftConfigNDResolvef03(cfg) == 0 and L.shafftConfigNDGetFlags(cfg) & 2  # CHANGED_NDA
    plan = ct.c_void_p()
    assert L.shafftNDCreate(ct.byref(plan)) == 0
    assert L.shafftNDInitFromConfigf03(plan, cfg) == 0
    alloc = ct.c_size_t()
    assert L.shafftGetAllocSize(plan, ct.byref(alloc)) == 0 and alloc.value == 40 * 30 * 20
    fcomm = ct.c_int(5)
    assert L.shafftGetCommunicatorf03(plan, ct.byref(fcomm)) == 0 and fcomm.value == 0
    assert L.shafftDestroy(ct.byref(plan)) == 0
    L.shafftConfigNDRelease(cfg)
    assert L.shafftConfigNDGetNdim(cfg) == 0
    bad = ct.c_int(3)
    assert L.shafftConfigNDInitf03(cfg, 3, shape, api.C2C, None, 0, 0, 0, 0, ct.byref(bad)) == 2  # INVALID_COMM
    L.shafftConfigNDFreef03(cfg)

    c1 = L.shafftConfig1DAllocf03()
    assert L.shafftConfig1DInitf03(c1, 1000003, api.C2C, ct.byref(zero)) == 0
    assert L.shafftConfig1DGetGlobalSize(c1) == 1000003 and L.shafftConfig1DGetAllocElements(c1) == 1000003
    assert L.shafftConfig1DGetInitialLocalSize(c1) == 1000003 and L.shafftConfig1DGetOutputLocalStart(c1) == 0
    p1 = ct.c_void_p()
    assert L.shafft1DCreate(ct.byref(p1)) == 0 and L.shafft1DInitFromConfigf03(p1, c1) == 0
    assert L.shafftDestroy(ct.byref(p1)) == 0
    L.shafftConfig1DRelease(c1)
    L.shafftConfig1DFreef03(c1)
    assert L.shafftConfigNDGetNdim(None) == 0  # null config reads as 0
    # ... except activeRank, -1 as in the reference's src/f03/shafftf03.cpp
    assert L.shafftConfigNDGetActiveRank(None) == -1 and L.shafftConfig1DGetActiveRank(None) == -1


def test_every_fortran_bound_symbol_is_exported():
    """Every C symbol the reference's Fortran module binds (bind(C, name=...)
    in src/f03/shafft.f03; list in tests/golden/f03_bound_symbols.txt, made
    by tests/golden/make_f03_symbols.py) is exported, so the module links
    against this library."""
    L = ct.CDLL(api.library_path())
    names = open(os.path.join(ROOT, "tests", "golden", "f03_bound_symbols.txt")).read().split()
    assert len(names) > 90
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_synthetic_input_matches_oracle_recipe(emu_lib):
    """The library's benchmark input generator (std::mt19937_64) gives the
    same values as the oracle's independent generator, both recipes and
    precisions, bit for bit."""
    import numpy as np

    import oracle
    from shafft_amd import api

    for prec, dt in ((api.C2C, np.complex64), (api.Z2Z, np.complex128)):
        u = api.synthetic_input(1000, 43, prec, api.UNIFORM, L=emu_lib)
        assert np.array_equal(u, oracle.uniform_input((1000,), 43, dt))
        g = api.synthetic_input(1000, 44, prec, api.GAUSSIAN, L=emu_lib)
        assert np.array_equal(g, oracle.gaussian_input((1000,), 44, dt))
    assert emu_lib.shafftAmdSyntheticInput(None, 4, 1, 0, 0) != 0
    assert emu_lib.shafftAmdSyntheticInput(None, 0, 1, 0, 0) == 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """No fallback: without libshafft_amd.so every entry point raises instead
    of computing anything elsewhere."""
    monkeypatch.setattr(api, "_LIB", None)
    monkeypatch.setattr(api, "_CURRENT", None)
    monkeypatch.setattr(api, "library_path", lambda: str(tmp_path / "libshafft_amd.so"))
    with pytest.raises(ImportError):
        api.lib()
    with pytest.raises(ImportError):
        api.FFTND()
    with pytest.raises(ImportError):
        api.OneDBackendPlan(1024, api.C2C)
    with pytest.raises(ImportError):
        api.Server()


def test_bench_refuses_several_gpus():
    """bench.py runs one process on one GPU: --gpus above 1 exits with a
    message (before touching a device)."""
    import json
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
