"""Oracle (test infrastructure only): exact emulation of the modular split of the muffin-tin
contraction (EXC_BATCH_SPLIT_CRT).

There is no reference counterpart: the reference forms these products in FP64 with zgemm
(hamiltonsetup.f90:90-103, overlapsetup.f90:95-108).  The library replaces each of them by
integer products that are exact and therefore independent of summation order, and this module
restates that arithmetic step by step so that a GPU result can be compared with it:

* ``row_scale`` — per row i of an operand X (N x K complex) the exponent e_i with
  |X_i|_2 < 2^e_i (frexp), and the scale exponent s_i = beta - e_i;
* ``integers`` — X' = rint(Re X 2^s_i) + i rint(Im X 2^s_i), exact integers below 2^beta;
* ``residues`` — symmetric residues of integers modulo m: [-128, 127] for m = 256,
  [-(m-1)/2, (m-1)/2] for odd m;
* ``product`` — conj(A) B^T from the integers: for each modulus the three real residue
  products of Gauss's form, P1 = ar br^T, P2 = ai bi^T, P3 = (ar - ai)(br + bi)^T (exact:
  residues <= 128 in magnitude, sums below 2^53, so float64 GEMMs are exact), combined to
  Re = P1 + P2, Im = P3 - P1 + P2 and reduced modulo m; the integers Re', Im' in [-P/2, P/2)
  reconstructed from their residues by Garner's mixed-radix algorithm; the FP64 value the
  correctly rounded Re' (Im'), times 2^-(s_i + t_j).

With 2^(2 beta + 2) <= P = prod m and every row's 2-norm below 2^beta, |Re'|, |Im'| <= |A'_i|
|B'_j| < P / 2, so the reconstruction is exact.  ``big_product`` forms Re', Im' with Python
integers instead (the pin of the modular route).

Only ``tests/`` may import this module.
"""
from __future__ import annotations

import math

import numpy as np

# pairwise coprime moduli, 256 first (its residues are two's-complement low bytes)
MODULI = (256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199)
LO_DIGITS = 6          # digits of the low part of the reconstruction: prod(MODULI[:6]) < 2^53


def moduli(n):
    if not 2 <= n <= len(MODULI):
        raise ValueError(f"{n} moduli not supported")
    return MODULI[:n]


def beta(n):
    """The largest beta with 2^(2 beta + 2) <= prod of the first n moduli."""
    P = math.prod(moduli(n))
    b = 0
    while 1 << (2 * (b + 1) + 2) <= P:
        b += 1
    return b


def row_scale(X, n):
    """Scale exponents s_i = beta - e_i, |X_i|_2 < 2^e_i (e_i = 0 for a zero row)."""
    nrm = np.sqrt((X.real ** 2 + X.imag ** 2).sum(axis=1))
    _, e = np.frexp(nrm)
    e = np.where(nrm > 0, e, 0).astype(np.int64)
    return beta(n) - e


def integers(X, s):
    sc = np.ldexp(1.0, s)[:, None]
    return np.rint(X.real * sc), np.rint(X.imag * sc)


def residues(x, m):
    """Symmetric residues of the integer-valued float64 array x modulo m."""
    if m == 256:
        return x - 256.0 * np.floor((x + 128.0) / 256.0)
    return x - m * np.rint(x / m)


def _garner_tables(mods):
    n = len(mods)
    M = [1] * n
    for j in range(1, n):
        M[j] = M[j - 1] * mods[j - 1]
    w = np.zeros((n, n), dtype=np.int64)       # w[l, j] = M_j mod m_l
    v = np.zeros(n, dtype=np.int64)            # v[l] = M_l^-1 mod m_l
    for l in range(1, n):
        for j in range(l):
            w[l, j] = M[j] % mods[l]
        v[l] = pow(M[l] % mods[l], -1, mods[l])
    return M, w, v


def _sym(x, m):
    """Symmetric residue of int64 x: [-128, 127] for m = 256, [-(m-1)/2, (m-1)/2] for odd m."""
    r = np.mod(x, m)
    return np.where(r > (m - 1) // 2, r - m, r)


def reconstruct(res, mods):
    """Integers x in [-P/2, P/2) from their residues res[l] (int64 arrays) modulo mods[l], as
    (x_hi, x_lo) with x = x_hi * M_6 + x_lo: Garner's mixed-radix digits d_l (symmetric),
    d_l = (r_l - sum_{j<l} d_j (M_j mod m_l)) (M_l^-1 mod m_l) mod m_l."""
    M, w, v = _garner_tables(mods)
    n = len(mods)
    d = []
    for l in range(n):
        acc = np.asarray(res[l], dtype=np.int64).copy()
        for j in range(l):
            acc = acc - d[j] * w[l, j]
        d.append(_sym(_sym(acc, mods[l]) * v[l], mods[l]) if l else _sym(acc, mods[l]))
    k = min(LO_DIGITS, n)
    lo = np.zeros_like(d[0])
    for j in range(k - 1, -1, -1):
        lo = lo * mods[j] + d[j]
    hi = np.zeros_like(d[0])
    for j in range(n - 1, k - 1, -1):
        hi = hi * mods[j] + d[j]
    return hi, lo, M[k] if k < n else math.prod(mods)


def to_float(hi, lo, Mk):
    """Correctly rounded float64 of hi * Mk + lo (Python integers, round half to even)."""
    ex = hi.astype(object) * Mk + lo.astype(object)
    return ex.astype(np.float64)


def _int_products(ar, ai, br, bi, mods, conj):
    """Residues modulo each m of Re and Im of op(a) b^T, op = conj or identity (Gauss's form)."""
    out_re, out_im = [], []
    for m in mods:
        rar, rai, rbr, rbi = (residues(x, m) for x in (ar, ai, br, bi))
        ras = residues(ar - ai if conj else ar + ai, m)
        rbs = residues(br + bi, m)
        p1 = rar @ rbr.T
        p2 = rai @ rbi.T
        p3 = ras @ rbs.T
        re = p1 + p2 if conj else p1 - p2
        im = p3 - p1 + p2 if conj else p3 - p1 - p2
        out_re.append(_sym(re.astype(np.int64), m))
        out_im.append(_sym(im.astype(np.int64), m))
    return out_re, out_im


def product(A, B, n, conj=True, sa=None, sb=None):
    """op(A) B^T through n moduli (op = conj by default), A (NA x K), B (NB x K) complex.
    sa, sb: scale exponents (default: row_scale).  Returns the complex128 result and the
    exponents used."""
    mods = moduli(n)
    sa = row_scale(A, n) if sa is None else sa
    sb = row_scale(B, n) if sb is None else sb
    ar, ai = integers(A, sa)
    br, bi = integers(B, sb)
    rre, rim = _int_products(ar, ai, br, bi, mods, conj)
    hr, lr, Mk = reconstruct(rre, mods)
    hi, li, _ = reconstruct(rim, mods)
    sc = -(sa[:, None] + sb[None, :])
    re = np.ldexp(to_float(hr, lr, Mk), sc)
    im = np.ldexp(to_float(hi, li, Mk), sc)
    return re + 1j * im, sa, sb


def big_product(ar, ai, br, bi, conj=True):
    """Re', Im' of op(a') b'^T for integer-valued float64 arrays, in exact Python integers."""
    to = lambda x: np.array(x.astype(np.int64), dtype=object)
    ar, ai, br, bi = map(to, (ar, ai, br, bi))
    if conj:
        return ar.dot(br.T) + ai.dot(bi.T), ar.dot(bi.T) - ai.dot(br.T)
    return ar.dot(br.T) - ai.dot(bi.T), ar.dot(bi.T) + ai.dot(br.T)
