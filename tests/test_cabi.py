"""The product library loads without a GPU and exports every entry point
declared in include/shafft_amd.h.  No compute call is made."""
import ctypes as ct
import os
import re

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
