"""Closed-form signals with spectra known at every bin, for full-size GPU
parity checks where the direct DFT is out of reach.  Each identity is
checked against the pinned oracle on small sizes in tests/test_oracle.py.

* chirp: for even n the DFT of exp(i pi m^2/n) is
  sqrt(n) e^{i pi/4} exp(-i pi k^2/n) (quadratic Gauss sum);
* separable terms: the DFT of u (x) v (x) w is DFT(u) (x) DFT(v) (x) DFT(w),
  so a sum of such terms has a spectrum computable one plane at a time.
"""
import numpy as np


def unit_root(q: np.ndarray, L: int) -> np.ndarray:
    """exp(2 pi i q / L) for integers q, L divisible by 8: the angle is
    folded into the first octant in integers before cos/sin are taken, so the
    rounding of pi does not bias every element the same way (an unfolded
    exp(i pi m^2/n) over 2^27 elements carries a coherent error that sums to
    4.5e-13 relative at bin 0 of its spectrum)."""
    q = np.mod(q, L)
    quarter, eighth = L // 4, L // 8
    qd, r = q // quarter, q % quarter
    comp = r > eighth
    a = np.where(comp, quarter - r, r).astype(np.float64) * (2 * np.pi / L)
    c, s = np.cos(a), np.sin(a)
    cb, sb = np.where(comp, s, c), np.where(comp, c, s)
    quads = [qd == 0, qd == 1, qd == 2, qd == 3]
    return np.select(quads, [cb, -sb, -cb, sb]) + 1j * np.select(quads, [sb, cb, -sb, -cb])


def chirp(n: int, start: int = 0, count: int | None = None) -> np.ndarray:
    """x[m] = exp(i pi m^2 / n) for m in [start, start+count), n divisible
    by 4, with the phase m^2 mod 2n reduced exactly in integers (m < 2^31
    keeps m^2 in int64)."""
    count = n - start if count is None else count
    m = np.arange(start, start + count, dtype=np.int64)
    return unit_root((m * m) % (2 * n), 2 * n)


def chirp_spectrum(n: int, start: int = 0, count: int | None = None) -> np.ndarray:
    """Forward DFT of chirp(n) at bins [start, start+count)."""
    return np.sqrt(n) * np.exp(1j * np.pi / 4) * np.conj(chirp(n, start, count))


class ChunkedError:
    """Relative L2 and L-inf error accumulated over chunks, with the
    tolerance rule of the reference's check_rel_error (both within
    tol * sqrt(ln N))."""

    def __init__(self):
        self.num = self.den = 0.0
        self.max_err = self.max_ref = 0.0
        self.count = 0

    def add(self, got: np.ndarray, ref: np.ndarray):
        d = got.astype(np.complex128) - ref
        self.num += float(np.vdot(d, d).real)
        self.den += float(np.vdot(ref, ref).real)
        self.max_err = max(self.max_err, float(np.abs(d).max()))
        self.max_ref = max(self.max_ref, float(np.abs(ref).max()))
        self.count += got.size

    @property
    def rel_l2(self) -> float:
        return float(np.sqrt(self.num / self.den))

    @property
    def rel_linf(self) -> float:
        return self.max_err / self.max_ref

    def ok(self, tol: float) -> bool:
        s = np.sqrt(np.log(max(self.count, 2)))
        return self.rel_l2 <= tol * s and self.rel_linf <= tol * s
