"""ctypes binding of ``include/shafft_amd.h`` and a thin mirror of the
reference's C++ classes ``shafft::FFTND`` / ``shafft::FFT1D``
(include/shafft/shafft.hpp:51-627) over its C API (include/shafft/shafft.h).

Status codes, enum values, the plan state machine and the buffer-swap
contract are the reference's; every method raises :class:`ShafftError` with
the library's last-error record on a non-zero status.
"""

from __future__ import annotations

import ctypes as ct
import os
from typing import Optional, Sequence

import numpy as np

C2C, Z2Z = 0, 1
FORWARD, BACKWARD = 0, 1
LAYOUT_REDISTRIBUTED, LAYOUT_INITIAL = 0, 1

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_LIB_PATH = None


def library_path() -> str:
    return os.path.join(_HERE, "libshafft_amd.so")


class ShafftError(RuntimeError):
    def __init__(self, status: int, where: str, message: str = ""):
        super().__init__(f"{where}: status {status} {message}".strip())
        self.status = status


_SIGS = {
    # seam
    "shafftAmdNDPlan": (ct.c_int, [ct.POINTER(ct.c_void_p), ct.c_int, ct.POINTER(ct.c_int), ct.c_int,
                                   ct.POINTER(ct.c_int), ct.c_int, ct.c_void_p, ct.c_void_p]),
    "shafftAmdNDSetStream": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "shafftAmdNDExecute": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_void_p), ct.c_int]),
    "shafftAmdNDDestroy": (ct.c_int, [ct.c_void_p]),
    "shafftAmdNDFinalize": (ct.c_int, []),
    "shafftAmdNormalizeComplexFloat": (ct.c_int, [ct.c_float, ct.c_size_t, ct.c_void_p, ct.c_void_p]),
    "shafftAmdNormalizeComplexDouble": (ct.c_int, [ct.c_double, ct.c_size_t, ct.c_void_p, ct.c_void_p]),
    "shafftAmd1DQueryLayout": (ct.c_int, [ct.c_size_t] + [ct.POINTER(ct.c_size_t)] * 5),
    "shafftAmd1DPlan": (ct.c_int, [ct.POINTER(ct.c_void_p), ct.c_size_t, ct.c_int, ct.c_void_p, ct.c_void_p]),
    "shafftAmd1DExecute": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int]),
    "shafftAmd1DExecuteNoAlloc": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int]),
    "shafftAmd1DSetStream": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "shafftAmd1DSynchronize": (ct.c_int, [ct.c_void_p]),
    "shafftAmd1DDestroy": (ct.c_int, [ct.c_void_p]),
    "shafftAmd1DNormalize": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_size_t]),
    "shafftAmdNDNumPasses": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int)]),
    "shafftAmdNDPassBytes": (ct.c_int, [ct.c_void_p, ct.c_int, ct.POINTER(ct.c_int64)]),
    "shafftAmdNDExecutePass": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_int64,
                                          ct.c_int]),
    "shafftAmd1DExecutePass": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_int64,
                                          ct.c_int]),
    "shafftAmdNDPassExtents": (ct.c_int, [ct.c_void_p, ct.c_int, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64)]),
    "shafftAmd1DPassExtents": (ct.c_int, [ct.c_void_p, ct.c_int, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64)]),
    "shafftAmd1DNumPasses": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int)]),
    "shafftAmd1DPassBytes": (ct.c_int, [ct.c_void_p, ct.c_int, ct.POINTER(ct.c_int64)]),
    "shafftAmdNDDescribe": (ct.c_int, [ct.c_void_p, ct.c_char_p, ct.c_int]),
    "shafftAmd1DDescribe": (ct.c_int, [ct.c_void_p, ct.c_char_p, ct.c_int]),
    "shafftAmdSyntheticInput": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_uint64, ct.c_int, ct.c_int]),
    "shafftAmdBatchPlan": (ct.c_int, [ct.POINTER(ct.c_void_p), ct.c_int, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int),
                                      ct.c_int]),
    "shafftAmdBatchLayout": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64)]),
    "shafftAmdBatchSetBuffers": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "shafftAmdBatchSetStream": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "shafftAmdBatchExecute": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int]),
    "shafftAmdBatchSynchronize": (ct.c_int, [ct.c_void_p]),
    "shafftAmdBatchSetPrecisionMode": (ct.c_int, [ct.c_void_p, ct.c_int]),
    "shafftAmdBatchAutotune": (ct.c_int, [ct.c_void_p, ct.c_char_p, ct.c_int]),
    "shafftAmdBatchInfo": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int),
                                      ct.POINTER(ct.c_int64)]),
    "shafftAmdBatchDestroy": (ct.c_int, [ct.c_void_p]),
    "shafftAmdServerCreate": (ct.c_int, [ct.POINTER(ct.c_void_p), ct.c_int, ct.c_int64, ct.c_int, ct.c_int]),
    "shafftAmdServerSubmitND": (ct.c_int, [ct.c_void_p, ct.c_int, ct.POINTER(ct.c_int), ct.c_int, ct.c_int, ct.c_int,
                                           ct.c_void_p, ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_int64)]),
    "shafftAmdServerSubmit1D": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int, ct.c_int, ct.c_int,
                                           ct.c_void_p, ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_int64)]),
    "shafftAmdServerWait": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int64, ct.POINTER(ct.c_int),
                                       ct.POINTER(ct.c_int), ct.POINTER(ct.c_int64)]),
    "shafftAmdServerQuery": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_int)]),
    "shafftAmdServerDrain": (ct.c_int, [ct.c_void_p]),
    "shafftAmdServerStats": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_double)]),
    "shafftAmdServerResetStats": (ct.c_int, [ct.c_void_p]),
    "shafftAmdServerSetBatching": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int, ct.c_int]),
    "shafftAmdServerBatchStats": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64)]),
    "shafftAmdServerNow": (ct.c_int, [ct.POINTER(ct.c_int64)]),
    "shafftAmdServerDestroy": (ct.c_int, [ct.c_void_p]),
    # reference C API
    "shafftSetStream": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "shafftLastErrorStatus": (ct.c_int, []),
    "shafftLastErrorSource": (ct.c_int, []),
    "shafftLastErrorDomainCode": (ct.c_int, []),
    "shafftLastErrorMessage": (ct.c_int, [ct.c_char_p, ct.c_int]),
    "shafftClearLastError": (None, []),
    "shafftErrorSourceName": (ct.c_char_p, [ct.c_int]),
    "shafftNDCreate": (ct.c_int, [ct.POINTER(ct.c_void_p)]),
    "shafftNDInit": (ct.c_int, [ct.c_void_p, ct.c_int, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.c_int,
                                ct.c_ssize_t, ct.c_int]),
    "shafftPlan": (ct.c_int, [ct.c_void_p]),
    "shafftDestroy": (ct.c_int, [ct.POINTER(ct.c_void_p)]),
    "shafftGetLayout": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_size_t), ct.POINTER(ct.c_size_t), ct.c_int]),
    "shafftGetAxes": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.c_int]),
    "shafftGetAllocSize": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_size_t)]),
    "shafftGetGlobalSize": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_size_t)]),
    "shafftIsConfigured": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int)]),
    "shafftIsActive": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int)]),
    "shafftSetBuffers": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "shafftGetBuffers": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_void_p)]),
    "shafftExecute": (ct.c_int, [ct.c_void_p, ct.c_int]),
    "shafftNormalize": (ct.c_int, [ct.c_void_p]),
    "shafftConfiguration1D": (ct.c_int, [ct.c_size_t, ct.POINTER(ct.c_size_t), ct.POINTER(ct.c_size_t),
                                         ct.POINTER(ct.c_size_t), ct.c_int, ct.c_ssize_t]),
    "shafft1DCreate": (ct.c_int, [ct.POINTER(ct.c_void_p)]),
    "shafft1DInit": (ct.c_int, [ct.c_void_p, ct.c_size_t, ct.c_size_t, ct.c_size_t, ct.c_int, ct.c_ssize_t]),
    "shafftConfigurationND": (ct.c_int, [ct.c_int, ct.POINTER(ct.c_int), ct.c_int, ct.POINTER(ct.c_int),
                                         ct.POINTER(ct.c_int), ct.POINTER(ct.c_size_t), ct.POINTER(ct.c_size_t),
                                         ct.POINTER(ct.c_int), ct.c_int, ct.c_size_t, ct.c_ssize_t]),
    "shafftConfigurationNDf03": (ct.c_int, [ct.c_int, ct.POINTER(ct.c_int), ct.c_int, ct.POINTER(ct.c_int),
                                            ct.POINTER(ct.c_int), ct.POINTER(ct.c_size_t),
                                            ct.POINTER(ct.c_size_t), ct.POINTER(ct.c_int), ct.c_int,
                                            ct.c_size_t, ct.POINTER(ct.c_int)]),
    "shafftNDInitf03": (ct.c_int, [ct.c_void_p, ct.c_int, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.c_int,
                                   ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]),
    "shafftConfiguration1Df03": (ct.c_int, [ct.c_size_t, ct.POINTER(ct.c_size_t), ct.POINTER(ct.c_size_t),
                                            ct.POINTER(ct.c_size_t), ct.c_int, ct.POINTER(ct.c_int)]),
    "shafft1DInitf03": (ct.c_int, [ct.c_void_p, ct.c_size_t, ct.c_size_t, ct.c_size_t, ct.c_int,
                                   ct.POINTER(ct.c_int)]),
    "shafftAllocBufferF": (ct.c_int, [ct.c_size_t, ct.POINTER(ct.c_void_p)]),
    "shafftAllocBufferD": (ct.c_int, [ct.c_size_t, ct.POINTER(ct.c_void_p)]),
    "shafftFreeBufferF": (ct.c_int, [ct.c_void_p]),
    "shafftFreeBufferD": (ct.c_int, [ct.c_void_p]),
    "shafftCopyToBufferF": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_size_t]),
    "shafftCopyToBufferD": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_size_t]),
    "shafftCopyFromBufferF": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_size_t]),
    "shafftCopyFromBufferD": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_size_t]),
    "shafftGetBackendName": (ct.c_char_p, []),
    "shafftGetVersion": (None, [ct.POINTER(ct.c_int)] * 3),
    "shafftGetVersionString": (ct.c_char_p, []),
    "shafftFinalize": (ct.c_int, []),
    # configuration objects (struct pointers passed as void*)
    "shafftConfigNDInit": (ct.c_int, [ct.c_void_p, ct.c_int, ct.POINTER(ct.c_size_t), ct.c_int,
                                      ct.POINTER(ct.c_int), ct.c_int, ct.c_int, ct.c_int, ct.c_size_t,
                                      ct.c_ssize_t]),
    "shafftConfigNDRelease": (None, [ct.c_void_p]),
    "shafftConfigNDResolve": (ct.c_int, [ct.c_void_p]),
    "shafftNDInitFromConfig": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "shafftConfig1DInit": (ct.c_int, [ct.c_void_p, ct.c_size_t, ct.c_int, ct.c_ssize_t]),
    "shafftConfig1DRelease": (None, [ct.c_void_p]),
    "shafftConfig1DResolve": (ct.c_int, [ct.c_void_p]),
    "shafft1DInitFromConfig": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "shafftGetCommunicator": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_ssize_t)]),
}

# Fortran-side config companions and per-field accessors (config pointers as
# void*): name -> (restype, argtypes).
_VP, _SZ = ct.c_void_p, ct.c_size_t
_SIGS.update({
    "shafftConfigNDAllocf03": (_VP, []),
    "shafftConfigNDFreef03": (None, [_VP]),
    "shafftConfig1DAllocf03": (_VP, []),
    "shafftConfig1DFreef03": (None, [_VP]),
    "shafftConfigNDInitf03": (ct.c_int, [_VP, ct.c_int, ct.POINTER(_SZ), ct.c_int, ct.POINTER(ct.c_int),
                                         ct.c_int, ct.c_int, ct.c_int, _SZ, ct.POINTER(ct.c_int)]),
    "shafftConfigNDResolvef03": (ct.c_int, [_VP]),
    "shafftNDInitFromConfigf03": (ct.c_int, [_VP, _VP]),
    "shafftConfig1DInitf03": (ct.c_int, [_VP, _SZ, ct.c_int, ct.POINTER(ct.c_int)]),
    "shafftConfig1DResolvef03": (ct.c_int, [_VP]),
    "shafft1DInitFromConfigf03": (ct.c_int, [_VP, _VP]),
    "shafftGetCommunicatorf03": (ct.c_int, [_VP, ct.POINTER(ct.c_int)]),
})
for _n in ("Ndim", "Precision", "IsActive", "ActiveRank", "ActiveSize", "Nda", "CommSize", "Status",
           "OutputPolicy", "Strategy", "HintNda", "Flags", "WorldComm", "ActiveComm", "NodeId", "NodeCount"):
    _SIGS["shafftConfigNDGet" + _n] = (ct.c_int, [_VP])
for _n in ("AllocElements", "MemLimit", "HostnameLen", "DeviceNameLen"):
    _SIGS["shafftConfigNDGet" + _n] = (_SZ, [_VP])
for _n in ("Hostname", "DeviceName"):
    _SIGS["shafftConfigNDGet" + _n] = (None, [_VP, ct.c_char_p, _SZ])
    _SIGS["shafftConfig1DGet" + _n] = (None, [_VP, ct.c_char_p, _SZ])
for _n in ("GlobalShape", "InitialSubsize", "InitialOffset", "OutputSubsize", "OutputOffset"):
    _SIGS["shafftConfigNDGet" + _n] = (None, [_VP, ct.POINTER(_SZ)])
for _n in ("CommDims", "HintCommDims"):
    _SIGS["shafftConfigNDGet" + _n] = (None, [_VP, ct.POINTER(ct.c_int)])
for _n, _t in (("OutputPolicy", ct.c_int), ("Strategy", ct.c_int), ("MemLimit", _SZ), ("HintNda", ct.c_int),
               ("HintCommDims", ct.POINTER(ct.c_int))):
    _SIGS["shafftConfigNDSet" + _n] = (None, [_VP, _t])
for _n in ("Precision", "IsActive", "ActiveRank", "ActiveSize", "Status", "Flags", "WorldComm",
           "ActiveComm", "NodeId", "NodeCount"):
    _SIGS["shafftConfig1DGet" + _n] = (ct.c_int, [_VP])
for _n in ("GlobalSize", "AllocElements", "HostnameLen", "DeviceNameLen", "InitialLocalSize",
           "InitialLocalStart", "OutputLocalSize", "OutputLocalStart"):
    _SIGS["shafftConfig1DGet" + _n] = (_SZ, [_VP])


def bind(path: str):
    """Load a library exposing the shafft_amd C ABI and type its symbols."""
    L = ct.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def lib():
    """The gfx950 library.  Raises if it has not been built."""
    global _LIB, _LIB_PATH
    if _LIB is None:
        p = library_path()
        if not os.path.exists(p):
            raise ImportError(f"{p} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        _LIB = bind(p)
        _LIB_PATH = p
    return _LIB


_CURRENT = None


def use(L):
    """Route the Python mirror through library handle L (tests bind the CPU
    emulation build this way); returns the previous handle."""
    global _CURRENT
    prev = _CURRENT
    _CURRENT = L
    return prev


def _L():
    return _CURRENT if _CURRENT is not None else lib()


def last_error_message(L=None) -> str:
    L = L or _L()
    buf = ct.create_string_buffer(1024)
    L.shafftLastErrorMessage(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int, where: str, L=None):
    if rc != 0:
        raise ShafftError(rc, where, last_error_message(L))


def _np_dtype(prec: int):
    return np.complex64 if prec == C2C else np.complex128


class DeviceBuffer:
    """Device buffer from shafftAllocBufferF/D (include/shafft/shafft.h:738-766)."""

    def __init__(self, count: int, prec: int, L=None):
        self.L = L or _L()
        self.count = int(count)
        self.prec = prec
        p = ct.c_void_p()
        fn = self.L.shafftAllocBufferF if prec == C2C else self.L.shafftAllocBufferD
        check(fn(self.count, ct.byref(p)), "shafftAllocBuffer", self.L)
        self.ptr = p.value

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=_np_dtype(self.prec))
        assert a.size <= self.count
        fn = self.L.shafftCopyToBufferF if self.prec == C2C else self.L.shafftCopyToBufferD
        check(fn(self.ptr, a.ctypes.data, a.size), "shafftCopyToBuffer", self.L)

    def download(self, count: Optional[int] = None, ptr: Optional[int] = None) -> np.ndarray:
        n = self.count if count is None else int(count)
        out = np.empty(n, dtype=_np_dtype(self.prec))
        fn = self.L.shafftCopyFromBufferF if self.prec == C2C else self.L.shafftCopyFromBufferD
        check(fn(out.ctypes.data, ptr if ptr is not None else self.ptr, n), "shafftCopyFromBuffer", self.L)
        return out

    def free(self):
        if self.ptr:
            fn = self.L.shafftFreeBufferF if self.prec == C2C else self.L.shafftFreeBufferD
            fn(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


UNIFORM, GAUSSIAN = 0, 1


def synthetic_input(count: int, seed: int, prec: int, distribution: int = UNIFORM, L=None) -> np.ndarray:
    """Benchmark input from std::mt19937_64(seed) (shafftAmdSyntheticInput):
    uniform [-1, 1) or Box-Muller Gaussian re/im per element."""
    L = L or _L()
    out = np.empty(int(count), dtype=_np_dtype(prec))
    check(L.shafftAmdSyntheticInput(out.ctypes.data, int(count), int(seed), distribution, prec),
          "shafftAmdSyntheticInput", L)
    return out


def download_ptr(ptr: int, count: int, prec: int, L=None) -> np.ndarray:
    L = L or _L()
    out = np.empty(count, dtype=_np_dtype(prec))
    fn = L.shafftCopyFromBufferF if prec == C2C else L.shafftCopyFromBufferD
    check(fn(out.ctypes.data, ptr, count), "shafftCopyFromBuffer", L)
    return out


class _PlanBase:
    def __init__(self, L=None):
        self.L = L or _L()
        self.h = ct.c_void_p()

    def plan(self):
        check(self.L.shafftPlan(self.h), "shafftPlan", self.L)

    def set_buffers(self, data: int, work: int):
        check(self.L.shafftSetBuffers(self.h, ct.c_void_p(data), ct.c_void_p(work)), "shafftSetBuffers", self.L)

    def get_buffers(self):
        d, w = ct.c_void_p(), ct.c_void_p()
        check(self.L.shafftGetBuffers(self.h, ct.byref(d), ct.byref(w)), "shafftGetBuffers", self.L)
        return d.value, w.value

    def execute(self, direction: int):
        check(self.L.shafftExecute(self.h, direction), "shafftExecute", self.L)

    def normalize(self):
        check(self.L.shafftNormalize(self.h), "shafftNormalize", self.L)

    def set_stream(self, stream: int):
        check(self.L.shafftSetStream(self.h, ct.c_void_p(stream)), "shafftSetStream", self.L)

    def alloc_size(self) -> int:
        v = ct.c_size_t()
        check(self.L.shafftGetAllocSize(self.h, ct.byref(v)), "shafftGetAllocSize", self.L)
        return v.value

    def global_size(self) -> int:
        v = ct.c_size_t()
        check(self.L.shafftGetGlobalSize(self.h, ct.byref(v)), "shafftGetGlobalSize", self.L)
        return v.value

    def release(self):
        if self.h.value:
            self.L.shafftDestroy(ct.byref(self.h))

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class FFTND(_PlanBase):
    """Mirror of shafft::FFTND (include/shafft/shafft.hpp:51-324) at one rank."""

    def __init__(self, L=None):
        super().__init__(L)
        check(self.L.shafftNDCreate(ct.byref(self.h)), "shafftNDCreate", self.L)

    def init(self, comm_dims: Sequence[int], dims: Sequence[int], fft_type: int,
             output: int = LAYOUT_REDISTRIBUTED):
        nd = len(dims)
        if len(comm_dims) != nd:
            raise ShafftError(6, "FFTND.init", "commDims/dimensions length mismatch")
        cd = (ct.c_int * nd)(*comm_dims)
        dd = (ct.c_int * nd)(*dims)
        check(self.L.shafftNDInit(self.h, nd, cd, dd, fft_type, 0, output), "shafftNDInit", self.L)
        self.dims = tuple(int(d) for d in dims)
        self.fft_type = fft_type


class FFT1D(_PlanBase):
    """Mirror of shafft::FFT1D (include/shafft/shafft.hpp:341-627) at one rank."""

    def __init__(self, L=None):
        super().__init__(L)
        check(self.L.shafft1DCreate(ct.byref(self.h)), "shafft1DCreate", self.L)

    def init(self, n: int, local_n: int, local_start: int, fft_type: int):
        check(self.L.shafft1DInit(self.h, n, local_n, local_start, fft_type, 0), "shafft1DInit", self.L)
        self.n = int(n)
        self.fft_type = fft_type


class NDBackendPlan:
    """Direct handle on the backend seam (fftndPlan/fftndExecute/...)."""

    def __init__(self, dims: Sequence[int], axes: Sequence[int], prec: int, L=None):
        self.L = L or _L()
        self.h = ct.c_void_p()
        nd = len(dims)
        ta = (ct.c_int * len(axes))(*axes)
        dd = (ct.c_int * nd)(*dims)
        check(self.L.shafftAmdNDPlan(ct.byref(self.h), len(axes), ta, nd, dd, prec, None, None),
              "shafftAmdNDPlan", self.L)
        self.prec = prec

    def execute(self, data: int, work: int, direction: int):
        d, w = ct.c_void_p(data), ct.c_void_p(work)
        check(self.L.shafftAmdNDExecute(self.h, ct.byref(d), ct.byref(w), direction), "shafftAmdNDExecute", self.L)
        return d.value, w.value

    def set_stream(self, stream: int):
        check(self.L.shafftAmdNDSetStream(self.h, ct.c_void_p(stream)), "shafftAmdNDSetStream", self.L)

    def num_passes(self) -> int:
        v = ct.c_int()
        check(self.L.shafftAmdNDNumPasses(self.h, ct.byref(v)), "shafftAmdNDNumPasses", self.L)
        return v.value

    def pass_bytes(self, i: int) -> int:
        v = ct.c_int64()
        check(self.L.shafftAmdNDPassBytes(self.h, i, ct.byref(v)), "shafftAmdNDPassBytes", self.L)
        return v.value

    def pass_extents(self, i: int):
        """(elements read, elements written) by pass i."""
        a, b = ct.c_int64(), ct.c_int64()
        check(self.L.shafftAmdNDPassExtents(self.h, i, ct.byref(a), ct.byref(b)), "shafftAmdNDPassExtents", self.L)
        return a.value, b.value

    def execute_pass(self, i: int, src: int, src_elems: int, dst: int, dst_elems: int, direction: int):
        check(self.L.shafftAmdNDExecutePass(self.h, i, ct.c_void_p(src), src_elems, ct.c_void_p(dst), dst_elems,
                                              direction), "shafftAmdNDExecutePass", self.L)

    def describe(self) -> str:
        buf = ct.create_string_buffer(1 << 16)
        check(self.L.shafftAmdNDDescribe(self.h, buf, 1 << 16), "shafftAmdNDDescribe", self.L)
        return buf.value.decode()

    def __del__(self):
        try:
            if self.h.value:
                self.L.shafftAmdNDDestroy(self.h)
                self.h = ct.c_void_p()
        except Exception:
            pass


class OneDBackendPlan:
    """Direct handle on the FFT1D backend seam (fft1dPlan/fft1dExecute/...)."""

    def __init__(self, n: int, prec: int, L=None):
        self.L = L or _L()
        self.h = ct.c_void_p()
        check(self.L.shafftAmd1DPlan(ct.byref(self.h), n, prec, None, None), "shafftAmd1DPlan", self.L)
        self.n, self.prec = n, prec

    def execute(self, src: int, dst: int, direction: int):
        check(self.L.shafftAmd1DExecute(self.h, ct.c_void_p(src), ct.c_void_p(dst), direction),
              "shafftAmd1DExecute", self.L)

    def execute_no_alloc(self, src: int, dst: int, direction: int):
        """shafftAmd1DExecuteNoAlloc: the same transform with the
        intermediates in the library's fixed device array (no allocation)."""
        check(self.L.shafftAmd1DExecuteNoAlloc(self.h, ct.c_void_p(src), ct.c_void_p(dst), direction),
              "shafftAmd1DExecuteNoAlloc", self.L)

    def synchronize(self):
        check(self.L.shafftAmd1DSynchronize(self.h), "shafftAmd1DSynchronize", self.L)

    def set_stream(self, stream: int):
        check(self.L.shafftAmd1DSetStream(self.h, ct.c_void_p(stream)), "shafftAmd1DSetStream", self.L)

    def num_passes(self) -> int:
        v = ct.c_int()
        check(self.L.shafftAmd1DNumPasses(self.h, ct.byref(v)), "shafftAmd1DNumPasses", self.L)
        return v.value

    def pass_bytes(self, i: int) -> int:
        v = ct.c_int64()
        check(self.L.shafftAmd1DPassBytes(self.h, i, ct.byref(v)), "shafftAmd1DPassBytes", self.L)
        return v.value

    def pass_extents(self, i: int):
        """(elements read, elements written) by pass i."""
        a, b = ct.c_int64(), ct.c_int64()
        check(self.L.shafftAmd1DPassExtents(self.h, i, ct.byref(a), ct.byref(b)), "shafftAmd1DPassExtents", self.L)
        return a.value, b.value

    def execute_pass(self, i: int, src: int, src_elems: int, dst: int, dst_elems: int, direction: int):
        check(self.L.shafftAmd1DExecutePass(self.h, i, ct.c_void_p(src), src_elems, ct.c_void_p(dst), dst_elems,
                                              direction), "shafftAmd1DExecutePass", self.L)

    def describe(self) -> str:
        buf = ct.create_string_buffer(1 << 16)
        check(self.L.shafftAmd1DDescribe(self.h, buf, 1 << 16), "shafftAmd1DDescribe", self.L)
        return buf.value.decode()

    def __del__(self):
        try:
            if self.h.value:
                self.L.shafftAmd1DDestroy(self.h)
                self.h = ct.c_void_p()
        except Exception:
            pass


FP32, SPLIT_BF16, SPLIT_FP16 = 0, 1, 2


class BatchPlan:
    """Many transforms of uneven shape in one persistent launch
    (shafftAmdBatch* in include/shafft_amd.h)."""

    def __init__(self, shapes, prec: int, L=None):
        self.L = L or _L()
        self.h = ct.c_void_p()
        self.shapes = [tuple(int(v) for v in sh) for sh in shapes]
        nd = (ct.c_int * len(self.shapes))(*[len(sh) for sh in self.shapes])
        flat = [v for sh in self.shapes for v in sh]
        dd = (ct.c_int * len(flat))(*flat)
        check(self.L.shafftAmdBatchPlan(ct.byref(self.h), len(self.shapes), nd, dd, prec), "shafftAmdBatchPlan", self.L)
        self.prec = prec
        tot = ct.c_int64()
        offs = (ct.c_int64 * len(self.shapes))()
        check(self.L.shafftAmdBatchLayout(self.h, ct.byref(tot), offs), "shafftAmdBatchLayout", self.L)
        self.total = tot.value
        self.offsets = list(offs)

    def set_buffers(self, src: int, dst: int, work: int):
        check(self.L.shafftAmdBatchSetBuffers(self.h, ct.c_void_p(src), ct.c_void_p(dst), ct.c_void_p(work)),
              "shafftAmdBatchSetBuffers", self.L)

    def set_stream(self, stream: int):
        check(self.L.shafftAmdBatchSetStream(self.h, ct.c_void_p(stream)), "shafftAmdBatchSetStream", self.L)

    def execute(self, direction: int, use_graph: bool = True):
        check(self.L.shafftAmdBatchExecute(self.h, direction, 1 if use_graph else 0), "shafftAmdBatchExecute", self.L)

    def synchronize(self):
        check(self.L.shafftAmdBatchSynchronize(self.h), "shafftAmdBatchSynchronize", self.L)

    def set_precision_mode(self, mode: int):
        """0 FP32, 1 bfloat16 hi/lo split (matrix cores), 2 float16 hi/lo split."""
        check(self.L.shafftAmdBatchSetPrecisionMode(self.h, mode), "shafftAmdBatchSetPrecisionMode", self.L)

    def autotune(self) -> str:
        buf = ct.create_string_buffer(4096)
        check(self.L.shafftAmdBatchAutotune(self.h, buf, 4096), "shafftAmdBatchAutotune", self.L)
        return buf.value.decode()

    def info(self):
        np_, grid = ct.c_int(), ct.c_int()
        items, nbytes = ct.c_int64(), ct.c_int64()
        check(self.L.shafftAmdBatchInfo(self.h, ct.byref(np_), ct.byref(items), ct.byref(grid), ct.byref(nbytes)),
              "shafftAmdBatchInfo", self.L)
        return {"passes": np_.value, "items": items.value, "grid": grid.value, "bytes": nbytes.value}

    def __del__(self):
        try:
            if self.h.value:
                self.L.shafftAmdBatchDestroy(self.h)
                self.h = ct.c_void_p()
        except Exception:
            pass


class ServedJob:
    """Outcome of one served job: status (0 = success), error message, and
    host steady-clock times in ns (arrival, admit, done) plus device ns."""

    def __init__(self, status, message, times):
        self.status = status
        self.message = message
        self.arrival, self.admit, self.done, self.gpu_ns = (int(t) for t in times)

    @property
    def latency_ms(self) -> float:
        return (self.done - self.arrival) * 1e-6


class Server:
    """Continuous-batching transform server (shafftAmdServer* in
    include/shafft_amd.h): jobs of mixed shape and precision are admitted
    onto `slots` device streams as slots and arena scratch free up."""

    def __init__(self, slots: int = 4, arena_bytes: int = 16 << 30, planners: int = 2, max_bypass: int = 64,
                 L=None):
        self.L = L or _L()
        self.h = ct.c_void_p()
        check(self.L.shafftAmdServerCreate(ct.byref(self.h), slots, arena_bytes, planners, max_bypass),
              "shafftAmdServerCreate", self.L)

    @staticmethod
    def now_ns(L=None) -> int:
        L = L or _L()
        v = ct.c_int64()
        check(L.shafftAmdServerNow(ct.byref(v)), "shafftAmdServerNow", L)
        return v.value

    def submit_nd(self, dims, prec: int, direction: int, src: int, dst: int, normalize: bool = False,
                  arrival_ns: int = -1) -> int:
        dd = (ct.c_int * len(dims))(*[int(d) for d in dims])
        job = ct.c_int64()
        check(self.L.shafftAmdServerSubmitND(self.h, len(dims), dd, prec, direction, 1 if normalize else 0,
                                             ct.c_void_p(src), ct.c_void_p(dst), arrival_ns, ct.byref(job)),
              "shafftAmdServerSubmitND", self.L)
        return job.value

    def submit_1d(self, n: int, prec: int, direction: int, src: int, dst: int, normalize: bool = False,
                  arrival_ns: int = -1) -> int:
        job = ct.c_int64()
        check(self.L.shafftAmdServerSubmit1D(self.h, int(n), prec, direction, 1 if normalize else 0,
                                             ct.c_void_p(src), ct.c_void_p(dst), arrival_ns, ct.byref(job)),
              "shafftAmdServerSubmit1D", self.L)
        return job.value

    def wait(self, job: int, timeout_s: float = -1.0):
        """ServedJob once the job has ended, None on timeout."""
        fin, st = ct.c_int(), ct.c_int()
        times = (ct.c_int64 * 4)()
        tns = -1 if timeout_s < 0 else int(timeout_s * 1e9)
        check(self.L.shafftAmdServerWait(self.h, job, tns, ct.byref(fin), ct.byref(st), times),
              "shafftAmdServerWait", self.L)
        if not fin.value:
            return None
        msg = last_error_message(self.L) if st.value else ""
        return ServedJob(st.value, msg, list(times))

    def finished(self, job: int) -> bool:
        fin = ct.c_int()
        check(self.L.shafftAmdServerQuery(self.h, job, ct.byref(fin)), "shafftAmdServerQuery", self.L)
        return bool(fin.value)

    def drain(self):
        check(self.L.shafftAmdServerDrain(self.h), "shafftAmdServerDrain", self.L)

    def stats(self) -> dict:
        c = (ct.c_int64 * 8)()
        v = (ct.c_double * 6)()
        check(self.L.shafftAmdServerStats(self.h, c, v), "shafftAmdServerStats", self.L)
        keys_c = ("submitted", "completed", "failed", "plans", "plan_builds", "arena_bytes", "arena_peak",
                  "table_bytes")
        keys_v = ("p50_ms", "p99_ms", "mean_ms", "max_ms", "span_s", "gflop")
        out = {k: int(x) for k, x in zip(keys_c, c)}
        out.update({k: float(x) for k, x in zip(keys_v, v)})
        return out

    def reset_stats(self):
        check(self.L.shafftAmdServerResetStats(self.h), "shafftAmdServerResetStats", self.L)

    def set_batching(self, max_elems: int, max_jobs: int = 64, max_inflight: int = 2):
        """Micro-batch jobs of at most max_elems elements (0: off), up to
        max_jobs per group and max_inflight groups running per precision
        (jobs admitted from now on)."""
        check(self.L.shafftAmdServerSetBatching(self.h, int(max_elems), int(max_jobs), int(max_inflight)),
              "shafftAmdServerSetBatching", self.L)

    def batch_stats(self):
        """(micro-batches formed, jobs they held) since the last reset."""
        b, j = ct.c_int64(), ct.c_int64()
        check(self.L.shafftAmdServerBatchStats(self.h, ct.byref(b), ct.byref(j)), "shafftAmdServerBatchStats",
              self.L)
        return b.value, j.value

    def close(self):
        if self.h.value:
            check(self.L.shafftAmdServerDestroy(self.h), "shafftAmdServerDestroy", self.L)
            self.h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
