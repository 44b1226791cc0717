"""FP64 oracle for the slab C2C FFT path -- TEST INFRASTRUCTURE ONLY.

Used by tests/, by ``__graft_entry__.smoke()`` to check its result, and by the
benchmark's CPU-baseline leg as the timed CPU port.  Nothing in the product
(``shafft_amd``) imports it.  See dft_oracle.c for what each function
restates and which reference lines it follows.
"""
from .oracle import (  # noqa: F401
    dftn, fftn_pow2, uniform_input, gaussian_input, norm_factor_nd, norm_factor_1d, rel_l2,
    check_rel_error, threads, load,
)
