"""MI355X (gfx950) backend for SHAFFT's slab-decomposed C2C FFT path.

The compute path is ``libshafft_amd.so`` (hand-written HIP kernels + C++ host
plan code); this package only binds its C ABI (``include/shafft_amd.h``) with
ctypes and mirrors the reference's ``FFTND`` / ``FFT1D`` call sequence
(init -> plan -> setBuffers -> execute -> normalize) for Python callers,
tests and the benchmark.  There is no fallback: if the shared library is
missing, importing :func:`lib` raises.
"""

from .api import (  # noqa: F401
    BACKWARD,
    BatchPlan,
    C2C,
    FORWARD,
    Z2Z,
    DeviceBuffer,
    FFT1D,
    FFTND,
    NDBackendPlan,
    OneDBackendPlan,
    Server,
    ServedJob,
    ShafftError,
    lib,
    library_path,
)
