"""Oracle (test infrastructure only): spherical Bessel functions and spherical harmonics.

Restates, from the mathematics, what the reference's special-function routines
compute for ``match``:

* ``sbessel_array``  — j_l(x), l = 0..lmax (src_lib/sbessel.f90:13).
* ``sbesseld1_array`` — d/dx j_l(x) (src_lib/sbesseldm.f90:13 with m = 1).
* ``genylm``          — complex Y_lm with the Condon–Shortley phase, packed at
                        lm = l(l+1)+m (0-based) (src_lib/genylm.f90:13-94).

Only ``tests/``, ``__graft_entry__.smoke`` and ``bench.py``'s CPU baseline may use
this module.  Pinned by tests/test_oracle_specfun.py against scipy.special, the
three-term recurrences, closed forms for l <= 2, and (when built) the
reference's own Fortran routines compiled into ``oracle/_ref``.
"""
from __future__ import annotations

import math

import numpy as np


def sbessel_array(lmax: int, x: np.ndarray) -> np.ndarray:
    """j_l(x) for l = 0..lmax, x >= 0 (array), shape (lmax+1, len(x)).

    x > lmax: upward recurrence j_{l+1} = (2l+1)/x j_l - j_{l-1} (stable there).
    x <= lmax: Miller downward recurrence from l = lmax+20+x, normalised to the
    larger of j_0 = sin x / x and j_1 = sin x / x^2 - cos x / x.
    x < 1e-8: leading term x^l / (2l+1)!!.
    """
    x = np.asarray(x, dtype=float)
    out = np.zeros((lmax + 1, x.size))
    tiny = x < 1e-8
    up = (~tiny) & (x > lmax)
    down = (~tiny) & (~up)
    if tiny.any():
        xt = x[tiny]
        t = np.ones_like(xt)
        dfac = 1.0
        out[0, tiny] = 1.0
        for l in range(1, lmax + 1):
            dfac *= 2 * l + 1
            t = t * xt
            out[l, tiny] = t / dfac
    if up.any():
        xu = x[up]
        j0 = np.sin(xu) / xu
        out[0, up] = j0
        if lmax >= 1:
            j1 = (j0 - np.cos(xu)) / xu
            out[1, up] = j1
            for l in range(1, lmax):
                j2 = (2 * l + 1) / xu * j1 - j0
                out[l + 1, up] = j2
                j0, j1 = j1, j2
    if down.any():
        xd = x[down]
        lstart = lmax + 20 + int(math.ceil(xd.max()))
        jp1 = np.zeros_like(xd)
        jl = np.full_like(xd, 1e-300)
        vals = np.zeros((lmax + 2, xd.size))
        for l in range(lstart, -1, -1):
            # j_{l-1} = (2l+1)/x j_l - j_{l+1}
            if l <= lmax + 1:
                vals[l] = jl
            jm1 = (2 * l + 1) / xd * jl - jp1
            jp1, jl = jl, jm1
            big = np.abs(jl) > 1e250
            if big.any():
                jl[big] *= 1e-250
                jp1[big] *= 1e-250
                vals[:, big] *= 1e-250
        # after the loop vals[0..lmax+1] hold the unnormalised sequence
        j0t = np.sin(xd) / xd
        j1t = np.sin(xd) / xd ** 2 - np.cos(xd) / xd
        use0 = np.abs(j0t) >= np.abs(j1t)
        scale = np.where(use0, j0t, j1t) / np.where(use0, vals[0], vals[1])
        out[:, down] = vals[:lmax + 1] * scale
    return out


def sbesseld1_array(lmax: int, x: np.ndarray) -> np.ndarray:
    """d/dx j_l(x), l = 0..lmax: j_l' = (l/x) j_l - j_{l+1}; at x = 0, 1/3 for l = 1 else 0."""
    x = np.asarray(x, dtype=float)
    j = sbessel_array(lmax + 1, x)
    out = np.zeros((lmax + 1, x.size))
    nz = x > 0.0
    for l in range(lmax + 1):
        out[l, nz] = l / x[nz] * j[l, nz] - j[l + 1, nz]
    if (~nz).any() and lmax >= 1:
        out[1, ~nz] = 1.0 / 3.0
    return out


def genylm(lmax: int, theta: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """Complex Y_lm(theta, phi) with Condon–Shortley phase, shape ((lmax+1)^2, n).

    Normalised associated Legendre functions by the standard three-term
    recurrence in l at fixed m (p_mm from p_{m-1,m-1}; p_{m+1,m}; then upward).
    Y_{l,-m} = (-1)^m conj(Y_lm).
    """
    theta = np.atleast_1d(np.asarray(theta, dtype=float))
    phi = np.atleast_1d(np.asarray(phi, dtype=float))
    n = theta.size
    ct = np.cos(theta)
    st = np.sin(theta)
    y = np.zeros(((lmax + 1) ** 2, n), dtype=np.complex128)
    pmm = np.full(n, math.sqrt(1.0 / (4.0 * math.pi)))
    for m in range(lmax + 1):
        if m > 0:
            pmm = -math.sqrt((2 * m + 1) / (2.0 * m)) * st * pmm
        eimp = np.exp(1j * m * phi)
        p_lm2 = None
        p_lm1 = pmm
        for l in range(m, lmax + 1):
            if l == m:
                p = pmm
            elif l == m + 1:
                p = math.sqrt(2 * m + 3) * ct * pmm
            else:
                a = math.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
                b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
                p = a * (ct * p_lm1 - b * p_lm2)
            if l > m:
                p_lm2, p_lm1 = p_lm1, p
            else:
                p_lm2, p_lm1 = None, p
            ylm = p * eimp
            y[l * (l + 1) + m] = ylm
            if m > 0:
                y[l * (l + 1) - m] = ((-1) ** m) * np.conj(ylm)
    return y
