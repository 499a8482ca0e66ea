"""Oracle (test infrastructure only): exact emulation of the modular split of the muffin-tin
contraction (EXC_BATCH_SPLIT_CRT).

There is no reference counterpart: the reference forms these products in FP64 with zgemm
(hamiltonsetup.f90:90-103, overlapsetup.f90:95-108).  The library replaces each of them by
integer products that are exact and therefore independent of summation order, and this module
restates that arithmetic step by step so that a GPU result can be compared with it:

* ``row_scale`` — per row i of an operand X (N x K complex) the exponent e_i with
  |X_i|_2 < 2^e_i (frexp), and the scale exponent s_i = beta - e_i;
* ``integers`` — X' = rint(Re X 2^s_i) + i rint(Im X 2^s_i), exact integers below 2^beta;
* ``residues`` — symmetric residues of integers modulo an odd m, in [-(m-1)/2, (m-1)/2];
* ``product`` — conj(A) B^T from the integers.  Every modulus m has a square root iota of -1
  (its prime factors are all 1 mod 4), so Z_m[i] splits into two copies of Z_m through
  x + iy -> x + iota y and x - iota y, both ring maps.  Per modulus the planes
  u_a = ar - iota ai, v_a = ar + iota ai (the maps of conj(a)) and u_b = br + iota bi,
  v_b = br - iota bi give two real residue products U = u_a u_b^T, V = v_a v_b^T, the maps
  of conj(A) B^T, and Re = (U + V)/2, Im = (U - V)/(2 iota) modulo m; the integers Re', Im' in
  [-P/2, P/2) reconstructed from their residues by Garner's mixed-radix algorithm; the FP64
  value the correctly rounded Re' (Im'), times 2^-(s_i + t_j).

With 2^(2 beta + 2) <= P = prod m and every row's 2-norm below 2^beta, |Re'|, |Im'| <= |A'_i|
|B'_j| < P / 2, so the reconstruction is exact.  ``big_product`` forms Re', Im' with Python
integers instead (the pin of the modular route).

Only ``tests/`` may import this module.
"""
from __future__ import annotations

import math

import numpy as np

# pairwise coprime odd moduli whose prime factors are all 1 mod 4 (so -1 is a square modulo
# each), the largest first: 221 = 13 17, 205 = 5 41
MODULI = (241, 233, 229, 221, 205, 197, 193, 181, 173, 157, 149, 137, 113, 109)
LO_DIGITS = 6          # digits of the low part of the reconstruction: prod(MODULI[:6]) < 2^53


def iota(m):
    """The smallest square root of -1 modulo m."""
    for x in range(1, m):
        if (x * x + 1) % m == 0:
            return x
    raise ValueError(f"-1 is not a square modulo {m}")


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
    """Symmetric residues of the integer-valued float64 array x modulo the odd m."""
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
    """Symmetric residue of int64 x modulo the odd m: [-(m-1)/2, (m-1)/2]."""
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


def planes(xr, xi, m, left):
    """The two residue planes of an operand modulo m: for the left operand conj(a) the maps
    ar - iota ai and ar + iota ai, for the right operand b the maps br + iota bi and br - iota bi."""
    t = iota(m)
    s = -1.0 if left else 1.0
    ri = residues(xi, m)                       # keeps xr +- t ri small: exact in float64
    return residues(xr + s * t * ri, m), residues(xr - s * t * ri, m)


def _int_products(ar, ai, br, bi, mods, conj):
    """Residues modulo each m of Re and Im of op(a) b^T, op = conj or identity: two real
    products per modulus through the split of Z_m[i]."""
    out_re, out_im = [], []
    for m in mods:
        t = iota(m)
        if conj:
            ua, va = planes(ar, ai, m, True)
        else:                                   # the maps of a itself: ar + iota ai, ar - iota ai
            ua, va = planes(ar, ai, m, False)
        ub, vb = planes(br, bi, m, False)
        U = (ua @ ub.T).astype(np.int64)        # exact: |residue| <= 120, sums below 2^53
        V = (va @ vb.T).astype(np.int64)
        inv2 = pow(2, -1, m)
        inv2i = pow(2 * t, -1, m)
        out_re.append(_sym(_sym(U + V, m) * inv2, m))
        out_im.append(_sym(_sym(U - V, m) * inv2i, m))
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
