"""Oracle (test infrastructure only): emulation of the split-BF16 muffin-tin contraction.

Restates in numpy the arithmetic of the precision variant of the setup (no reference
counterpart: the reference computes these products in FP64 with zgemm,
hamiltonsetup.f90:90-103, overlapsetup.f90:95-108):

* ``split(X, natm, naa)`` — per (row, atom block) exponent e = exponent of the block's
  largest |Re|, |Im|, and three fixed-point slices x = 2^(e-8) (s0 + s1 + s2) with
  s0 = rint(x 2^(8-e)), s1 = q2/2^8, s2 = q3/2^16 (each exactly representable in BF16).
* ``product(A, B)`` — conj(A) B^T from the six slice pairs with s + t <= 2, summed per
  atom block and scaled back by 2^(ea + eb - 16); slice products are exact, so the only
  difference to the GPU kernel is the FP32 rounding of the second accumulation level.

Only ``tests/`` may import this module.
"""
from __future__ import annotations

import numpy as np

PAIRS = [(0, 0), (0, 1), (1, 0), (0, 2), (1, 1), (2, 0)]


def split(X: np.ndarray, natm: int, naa: int):
    """X (n, natm*naa) complex -> slices (3, n, natm*naa) complex (re, im sliced separately), exps (n, natm)."""
    n = X.shape[0]
    Xb = X.reshape(n, natm, naa)
    m = np.maximum(np.abs(Xb.real), np.abs(Xb.imag)).max(axis=2)
    _, e = np.frexp(m)
    e = np.where(m > 0, e, 0).astype(np.int64)
    sc = np.ldexp(1.0, 8 - e)[:, :, None]
    out = []
    for part in (Xb.real, Xb.imag):
        x = part * sc
        q1 = np.rint(x)
        r1 = (x - q1) * 256.0
        q2 = np.rint(r1)
        q3 = np.rint((r1 - q2) * 256.0)
        out.append((q1, q2 / 256.0, q3 / 65536.0))
    slices = np.stack([out[0][s] + 1j * out[1][s] for s in range(3)]).reshape(3, n, natm * naa)
    return slices, e


def product(A: np.ndarray, B: np.ndarray, natm: int, naa: int, pairs=PAIRS) -> np.ndarray:
    """conj(A) B^T through the slice pairs; A (n, K), B (m, K), K = natm*naa."""
    sa, ea = split(A, natm, naa)
    sb, eb = split(B, natm, naa)
    out = np.zeros((A.shape[0], B.shape[0]), dtype=np.complex128)
    for a in range(natm):
        cols = slice(a * naa, (a + 1) * naa)
        blk = np.zeros_like(out)
        for s, t in pairs:
            blk += sa[s][:, cols].conj() @ sb[t][:, cols].T
        out += blk * np.ldexp(1.0, (ea[:, a][:, None] + eb[:, a][None, :] - 16))
    return out
