"""Oracle (test infrastructure only): FP64 CPU restatement of the G+k set of a k-point.

* ``gvectors``  — the G vectors of an integer box in Cartesian form, G = i1 b1 + i2 b2 + i3 b3
                  (gengvec.f90:105-107), for the given ivg order
* ``gengpvec``  — gengpvec.f90:50-92: the G (of the first ngvec, in their order) with
                  |G + k|^2 < gkmax^2, igpig (1-based), vgpl = ivg + vpl, vgpc = G + k, and
                  their spherical coordinates (``sphcrd``)
* ``sphcrd``    — sphcrd.f90:27-49: length, theta = acos(z / r) clamped to [-1, 1], phi =
                  atan2(y, x) moved to [0, 2 pi); zero angles below r = 1e-14, phi zero when
                  |x|, |y| <= 1e-14
* ``gensfacgp`` — gensfacgp.f90:38-49: S_a(G + k) = exp(i (G + k).r_a)

Every function here is plain numpy, one vector operation per step of the reference loop
(no reordering of the sums), so its rounding is the reference's without FMA contraction.

Only ``tests/``, ``__graft_entry__.smoke`` and ``bench.py``'s CPU-baseline leg may import
this module.  The library never calls it.
"""
from __future__ import annotations

import numpy as np

TWOPI = 6.2831853071795864769
EPS = 1e-14


def gvectors(ivg: np.ndarray, bvec: np.ndarray) -> np.ndarray:
    """vgc(3, n) = dble(i1) bvec(:,1) + dble(i2) bvec(:,2) + dble(i3) bvec(:,3), summed left to
    right (gengvec.f90:106-107)."""
    iv = ivg.astype(np.float64)
    return (iv[0][None, :] * bvec[:, 0][:, None] + iv[1][None, :] * bvec[:, 1][:, None]) + \
        iv[2][None, :] * bvec[:, 2][:, None]


def sphcrd(v: np.ndarray):
    """(r, theta, phi) of the columns of v (sphcrd.f90:27-49)."""
    r = np.sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])
    th = np.zeros_like(r)
    ph = np.zeros_like(r)
    ok = r > EPS
    t1 = np.zeros_like(r)
    t1[ok] = v[2, ok] / r[ok]
    t1 = np.where(t1 > 0.0, np.minimum(t1, 1.0), np.maximum(t1, -1.0))
    th[ok] = np.arccos(t1[ok])
    okp = ok & ((np.abs(v[0]) > EPS) | (np.abs(v[1]) > EPS))
    ph[okp] = np.arctan2(v[1, okp], v[0, okp])
    ph = np.where(okp & (ph < 0.0), ph + TWOPI, ph)
    return r, th, ph


def gengpvec(ivg: np.ndarray, bvec: np.ndarray, ngvec: int, vpl, vpc, gkmax: float):
    """G+k set of one k-point: (ngp, igpig(ngp) 1-based, vgpl(3,ngp), vgpc(3,ngp), gpc(ngp),
    tpgpc(2,ngp))."""
    vpl = np.asarray(vpl, dtype=np.float64)
    vpc = np.asarray(vpc, dtype=np.float64)
    vgc = gvectors(ivg[:, :ngvec], bvec)
    v = vgc + vpc[:, None]
    t2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]
    sel = np.nonzero(t2 < gkmax * gkmax)[0]
    vgpc = v[:, sel]
    vgpl = ivg[:, sel].astype(np.float64) + vpl[:, None]
    r, th, ph = sphcrd(vgpc)
    return sel.size, (sel + 1).astype(np.int32), vgpl, vgpc, r, np.stack([th, ph])


def gensfacgp(vgpc: np.ndarray, atposc: np.ndarray) -> np.ndarray:
    """sfacgp(ngp, natm) = exp(i vgpc . atposc), the dot product summed left to right."""
    t = (vgpc[0][:, None] * atposc[0][None, :] + vgpc[1][:, None] * atposc[1][None, :]) + \
        vgpc[2][:, None] * atposc[2][None, :]
    return np.cos(t) + 1j * np.sin(t)
