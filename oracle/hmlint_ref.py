"""Oracle (test infrastructure only): FP64 CPU restatement of the muffin-tin integrals of one
SCF iteration, the producers of the setup's h_a, oalo and ololo:

* ``spline``        — src_lib/spline.f90:46-96 (cubic spline with the reference's end conditions)
* ``fderiv_int``    — src_lib/fderiv.f90:64-72 (m = -1: the integral over the whole mesh, g(n))
* ``wigner3j``      — the Racah formula (src_lib/wigner3j.f90 evaluates the same symbol)
* ``gaunt``         — src_lib/gaunt.f90: int conj(Y_l1m1) Y_l2m2 Y_l3m3 dOmega
* ``gauntyry``      — src_lib/gauntyry.f90:30-56: int conj(Y_l1m1) R_l2m2 Y_l3m3 dOmega
* ``rlm``           — src_lib/genrlm.f90:58-74 (real spherical harmonics from the complex ones)
* ``hmlint``        — src_eigensystem/hmlint.f90:33-34 = mt_kin.f90:96-221 + mt_pot.f90:111-307
                      (ValenceRelativity 'none', spinless: no rmtable factor, no bxcmt, no SO)
* ``olprad``        — src_eigensystem/olprad.f90:90-125 (oalo, ololo; no IORA terms)
* ``genrhomt``      — genrhomt.f90:34-150 (the muffin-tin density rho_lm(r) in real harmonics
                      from the density matrix of gendmatmt, listgnt as init1.f90:604-621)
* ``polynom``       — src_lib/polynom.f90 (np = 4: the cubic through four points and its
                      derivatives at x), and ``dmat_from_radial`` — match.f90:85, 139-148
                      (the boundary values zd of the radial functions at R)

Arrays follow the reference shapes (Fortran order, 0-based here): apwfr(nrmtmax, 2, apwordmax,
0:lmaxapw, natm) with [:, 0] = u and [:, 1] = du/dr, lofr(nrmtmax, 2, nlomax, natm),
veffmt(lmmaxvr, nrmtmax, natm) (coefficients of the real harmonics R_lm), spr(nrmtmax, nspecies).

Only ``tests/``, ``__graft_entry__.smoke`` and ``bench.py``'s CPU-baseline leg may import this
module.  The library never calls it.
"""
from __future__ import annotations

import functools
import math
from fractions import Fraction

import numpy as np

from . import specfun

Y00 = 0.28209479177387814347   # 1 / sqrt(4 pi)


def spline(x, f):
    """spline.f90:46-96.  x: (n,), f: (n,) or (n, m).  Returns (c1, c2, c3) shaped like f, the
    coefficients cf(1:3, i) of f(x_i + t) = f_i + c1 t + c2 t^2 + c3 t^3."""
    x = np.asarray(x, dtype=np.float64)
    f = np.asarray(f, dtype=np.float64)
    n = x.shape[0]
    w = 1.0 / (x[1:] - x[:-1])                                   # w(i), i = 1 .. n-1
    ex = (slice(None),) + (None,) * (f.ndim - 1)
    c1 = np.zeros_like(f)
    c3 = np.zeros_like(f)
    c2s = np.zeros(n)                                            # cf(2, :) before :86 (f-independent)
    c1[:-1] = w[ex] * (f[1:] - f[:-1])
    c2s[0] = 1.0
    c3[0] = (c1[1] - c1[0]) / (x[2] - x[0])
    t1 = (x[1] - x[0]) * w[1]
    t2 = t1 * c2s[0]
    t3 = 1.0 / (2.0 * (t1 + 1.0))
    c2s[1] = t3
    t4 = 3.0 * (c1[1] - c1[0]) * w[1] - t2 * c3[0]
    c3[1] = t4
    for i in range(2, n - 1):
        t1 = (x[i] - x[i - 1]) * w[i]
        t2 = t1 * t3
        t3 = 1.0 / (2.0 * t1 + 2.0 - t2)
        c2s[i] = t3
        t4 = 3.0 * (c1[i] - c1[i - 1]) * w[i] - t2 * t4
        c3[i] = t4
    c2s[n - 1] = 1.0
    t1 = (c1[n - 2] - c1[n - 3]) / (x[n - 1] - x[n - 3])
    c3[n - 1] = t1
    for i in range(n - 2, 0, -1):
        t1 = c3[i] - t1 * c2s[i + 1]
        c3[i] = t1
    c2 = c2s[ex] * c3
    for i in range(n - 1):
        t1 = 0.3333333333333333333 * (c2[i + 1] - c2[i]) * w[i]
        c3[i] = t1
        t2 = x[i + 1] - x[i]
        c1[i] = c1[i] - (c2[i] + t1 * t2) * t2
    t1 = x[n - 1] - x[n - 2]
    c1[n - 1] = c1[n - 2] + (2.0 * c2[n - 2] + 3.0 * c3[n - 2] * t1) * t1
    c3[n - 1] = c3[n - 2]
    return c1, c2, c3


def fderiv_int(x, f):
    """fderiv.f90:64-72 with m = -1, the last entry g(n): the integral of the spline over the
    mesh, accumulated interval by interval.  f: (n,) or (n, m)."""
    x = np.asarray(x, dtype=np.float64)
    f = np.asarray(f, dtype=np.float64)
    c1, c2, c3 = spline(x, f)
    g = np.zeros(f.shape[1:])
    for i in range(x.shape[0] - 1):
        dx = x[i + 1] - x[i]
        g = g + (((0.25 * c3[i] * dx + 0.3333333333333333333 * c2[i]) * dx + 0.5 * c1[i]) * dx + f[i]) * dx
    return g


@functools.lru_cache(maxsize=None)
def wigner3j(j1, j2, j3, m1, m2, m3):
    """(j1 j2 j3; m1 m2 m3) by the Racah formula in exact rational arithmetic."""
    if m1 + m2 + m3 != 0 or abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    if j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0.0
    fa = math.factorial
    tri = Fraction(fa(j1 + j2 - j3) * fa(j1 - j2 + j3) * fa(-j1 + j2 + j3), fa(j1 + j2 + j3 + 1))
    pre = tri * fa(j1 + m1) * fa(j1 - m1) * fa(j2 + m2) * fa(j2 - m2) * fa(j3 + m3) * fa(j3 - m3)
    s = Fraction(0)
    for k in range(0, j1 + j2 + j3 + 1):
        den = [k, j3 - j2 + k + m1, j3 - j1 + k - m2, j1 + j2 - j3 - k, j1 - k - m1, j2 - k + m2]
        if min(den) < 0:
            continue
        d = 1
        for v in den:
            d *= fa(v)
        s += Fraction((-1) ** k, d)
    sign = -1.0 if (j1 - j2 - m3) % 2 else 1.0
    # sqrt of the rational prefactor times the rational sum, rounded once
    val = s * s * pre
    return sign * (1.0 if s >= 0 else -1.0) * math.sqrt(float(val.numerator) / float(val.denominator)) \
        if val != 0 else 0.0


@functools.lru_cache(maxsize=None)
def gaunt(l1, l2, l3, m1, m2, m3):
    """int conj(Y_l1m1) Y_l2m2 Y_l3m3 dOmega (nonzero for m1 = m2 + m3, l1+l2+l3 even)."""
    if m1 - m2 - m3 != 0:
        return 0.0
    if (l1 + l2 + l3) % 2:
        return 0.0
    t = math.sqrt((2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1) / (4.0 * math.pi))
    sign = -1.0 if m1 % 2 else 1.0
    return sign * t * wigner3j(l1, l2, l3, 0, 0, 0) * wigner3j(l1, l2, l3, -m1, m2, m3)


def gauntyry(l1, l2, l3, m1, m2, m3):
    """gauntyry.f90:30-56: the Gaunt coefficient with the real harmonic R_l2m2 in the middle,
    R_lm = sqrt2 Re Y_lm (m > 0), Y_l0 (m = 0), sqrt2 Im Y_lm (m < 0)."""
    c1 = 0.7071067811865475244
    if m2 > 0:
        t1 = 1.0 if m2 % 2 == 0 else -1.0
        return complex(c1 * (gaunt(l1, l2, l3, m1, m2, m3) + t1 * gaunt(l1, l2, l3, m1, -m2, m3)), 0.0)
    if m2 < 0:
        t1 = 1.0 if m2 % 2 == 0 else -1.0
        return complex(0.0, -c1 * (gaunt(l1, l2, l3, m1, m2, m3) - t1 * gaunt(l1, l2, l3, m1, -m2, m3)))
    return complex(gaunt(l1, l2, l3, m1, m2, m3), 0.0)


def rlm(lmax, theta, phi):
    """genrlm.f90:58-74 from the oracle's complex Y_lm: (npoints, (lmax+1)^2)."""
    y = specfun.genylm(lmax, np.atleast_1d(theta), np.atleast_1d(phi)).T
    r = np.empty(y.shape)
    s2 = math.sqrt(2.0)
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            lm = l * l + l + m
            r[..., lm] = s2 * y[..., lm].imag if m < 0 else (y[..., lm].real if m == 0 else s2 * y[..., lm].real)
    return r


@functools.lru_cache(maxsize=None)
def gaunt_table(lmaxmat, lmaxvr):
    """G[lm1, lm2, lm3] = gauntyry(l1, l2, l3, m1, m2, m3) for l1, l3 <= lmaxmat, l2 <= lmaxvr
    (the dense form of gntnonz, init1.f90:560-600)."""
    n1, n2 = (lmaxmat + 1) ** 2, (lmaxvr + 1) ** 2
    G = np.zeros((n1, n2, n1), dtype=np.complex128)
    for l1 in range(lmaxmat + 1):
        for m1 in range(-l1, l1 + 1):
            for l3 in range(lmaxmat + 1):
                for m3 in range(-l3, l3 + 1):
                    mabs = abs(m1 - m3)
                    for l2 in range(abs(l1 - l3), min(l1 + l3, lmaxvr) + 1, 2):
                        if l2 < mabs:
                            continue
                        for m2 in {mabs, -mabs}:
                            g = gauntyry(l1, l2, l3, m1, m2, m3)
                            if g != 0:
                                G[l1 * l1 + l1 + m1, l2 * l2 + l2 + m2, l3 * l3 + l3 + m3] = g
    return G


def _functions(cr, rd, ias):
    """The radial functions of atom ias: APW (io, l) for l <= lmaxmat, then the lo functions.
    Returns (u (nr, nf), du (nr, nf), l of each function, nr, x)."""
    is_ = cr.spidx[ias]
    nr = int(rd.nrmt[is_])
    x = rd.spr[:nr, is_]
    us, dus, ls = [], [], []
    for l in range(cr.lmaxmat + 1):
        for io in range(cr.apwordmax):
            us.append(rd.apwfr[:nr, 0, io, l, ias])
            dus.append(rd.apwfr[:nr, 1, io, l, ias])
            ls.append(l)
    for ilo in range(cr.nlorb[is_]):
        us.append(rd.lofr[:nr, 0, ilo, ias])
        dus.append(rd.lofr[:nr, 1, ilo, ias])
        ls.append(int(cr.lorbl[ilo, is_]))
    return np.array(us).T, np.array(dus).T, np.array(ls), nr, x


def radial_integrals(cr, rd, ias):
    """Per atom: P[lm2, a, b] = int u_a u_b V_lm2 r^2 (mt_pot.f90:142-146, 177-181, 208-212),
    K[a, b] = int (1/2 u_a' u_b' + 1/2 l(l+1) u_a u_b / r^2) r^2 for l_a = l_b (mt_kin.f90:
    120-127, 139-146, 160-167; the /y00 * y00 of the reference cancels), S[a, b] = int u_a u_b
    r^2 (olprad.f90:94-98, 121-125).  Functions a: APW (l, io) io fastest, then lo."""
    u, du, ls, nr, x = _functions(cr, rd, ias)
    nf = u.shape[1]
    r2 = x * x
    V = rd.veffmt[:, :nr, ias]                                   # (lmmaxvr, nr)
    prod = u[:, :, None] * u[:, None, :]                         # (nr, nf, nf)
    P = fderiv_int(x, (prod[:, None, :, :] * (V.T * r2[:, None])[:, :, None, None]).reshape(nr, -1))
    P = P.reshape(V.shape[0], nf, nf)
    ang = (ls * (ls + 1)).astype(np.float64)
    kin = (0.5 * du[:, :, None] * du[:, None, :] * r2[:, None, None]
           + 0.5 * ang[None, :, None] * prod)                    # l_a (l_a+1) u_a u_b / r^2 * r^2
    K = fderiv_int(x, kin.reshape(nr, -1)).reshape(nf, nf)
    K[ls[:, None] != ls[None, :]] = 0.0
    S = fderiv_int(x, (prod * r2[:, None, None]).reshape(nr, -1)).reshape(nf, nf)
    return P, K, S, ls


def hmlint(cr, rd):
    """haa(maxaa, maxaa, natm), loa(maxnlo, maxaa, natm), lolo(maxnlo, maxnlo, natm) of
    mt_kin + mt_pot (hmlint.f90:33-34).  Rows/columns: APW f = (l, m, io) io fastest
    (mt_kin.f90:171-184, mt_pot.f90:224-253), lo (ilo, m) m fastest (:258-307)."""
    ao, L = cr.apwordmax, cr.lmaxmat
    naa = ao * (L + 1) ** 2
    mn = cr.maxnlo
    G = gaunt_table(L, rd.lmaxvr)                                # (lm1, lm2, lm3)
    haa = np.zeros((naa, naa, cr.natm), dtype=np.complex128, order="F")
    loa = np.zeros((max(mn, 1), naa, cr.natm), dtype=np.complex128, order="F")
    lolo = np.zeros((max(mn, 1), max(mn, 1), cr.natm), dtype=np.complex128, order="F")
    lmv = (rd.lmaxvr + 1) ** 2
    for ias in range(cr.natm):
        is_ = cr.spidx[ias]
        P, K, _, ls = radial_integrals(cr, rd, ias)
        fa = lambda l, io: io + ao * l                           # APW function index
        # APW-APW: sum_lm2 G(lm1, lm2, lm3) P(lm2, (l1,io1), (l3,io2)) + delta_lm1lm3 K
        for l1 in range(L + 1):
            for m1 in range(-l1, l1 + 1):
                lm1 = l1 * l1 + l1 + m1
                for io1 in range(ao):
                    f1 = io1 + ao * lm1
                    for l3 in range(L + 1):
                        for m3 in range(-l3, l3 + 1):
                            lm3 = l3 * l3 + l3 + m3
                            g = G[lm1, :lmv, lm3]
                            for io2 in range(ao):
                                f3 = io2 + ao * lm3
                                z = np.dot(g, P[:lmv, fa(l1, io1), fa(l3, io2)])
                                if lm1 == lm3:
                                    z += K[fa(l1, io1), fa(l3, io2)]
                                haa[f1, f3, ias] = z
        # lo rows
        nlo = cr.nlorb[is_]
        r = 0
        for ilo in range(nlo):
            l1 = int(cr.lorbl[ilo, is_])
            b1 = ao * (L + 1) + ilo
            for m1 in range(-l1, l1 + 1):
                lm1 = l1 * l1 + l1 + m1
                for l3 in range(L + 1):
                    for m3 in range(-l3, l3 + 1):
                        lm3 = l3 * l3 + l3 + m3
                        g = G[lm1, :lmv, lm3]
                        for io in range(ao):
                            z = np.dot(g, P[:lmv, b1, fa(l3, io)])
                            if lm1 == lm3:
                                z += K[b1, fa(l3, io)]
                            loa[r, io + ao * lm3, ias] = z
                r3 = 0
                for ilo2 in range(nlo):
                    l3 = int(cr.lorbl[ilo2, is_])
                    b3 = ao * (L + 1) + ilo2
                    for m3 in range(-l3, l3 + 1):
                        lm3 = l3 * l3 + l3 + m3
                        z = np.dot(G[lm1, :lmv, lm3], P[:lmv, b1, b3])
                        if lm1 == lm3:
                            z += K[b1, b3]
                        lolo[r, r3, ias] = z
                        r3 += 1
                r += 1
    return haa, loa, lolo


def olprad(cr, rd):
    """oalo(apwordmax, nlomax, natm), ololo(nlomax, nlomax, natm) (olprad.f90:90-125): the
    APW-lo overlap integrals for the lo's own l, and the lo-lo ones for equal l (else 0)."""
    ao, L = cr.apwordmax, cr.lmaxmat
    nlm = max(cr.nlomax, 1)
    oalo = np.zeros((ao, nlm, cr.natm), order="F")
    ololo = np.zeros((nlm, nlm, cr.natm), order="F")
    for ias in range(cr.natm):
        is_ = cr.spidx[ias]
        _, _, S, ls = radial_integrals(cr, rd, ias)
        for ilo in range(cr.nlorb[is_]):
            l = int(cr.lorbl[ilo, is_])
            b = ao * (L + 1) + ilo
            for io in range(ao):
                oalo[io, ilo, ias] = S[io + ao * l, b]
            for ilo2 in range(cr.nlorb[is_]):
                if int(cr.lorbl[ilo2, is_]) == l:
                    ololo[ilo, ilo2, ias] = S[b, ao * (L + 1) + ilo2]
    return oalo, ololo


def polynom(m, xa, ya, x):
    """polynom.f90, np = 4 (the case match.f90:85 uses for apwordmax <= 3): the m-th derivative
    at x (m = 0..3) of the cubic through (xa_i, ya_i), in the reference's form: the coefficients
    of the cubic in t = x - xa_1 from the divided differences of the three offsets."""
    x1, x2, x3 = xa[1] - xa[0], xa[2] - xa[0], xa[3] - xa[0]
    y1, y2, y3 = ya[1] - ya[0], ya[2] - ya[0], ya[3] - ya[0]
    t1 = 1.0 / (x1 * x2 * x3 * (x1 - x2) * (x1 - x3) * (x2 - x3))
    t2, t3, t4 = x1 * x2 * y3, x2 * x3 * y1, x3 * x1 * y2
    t5, t6, t7 = x1 ** 2, x2 ** 2, x3 ** 2
    c2 = t1 * (t3 * (x3 * t6 - x2 * t7) + t4 * (x1 * t7 - x3 * t5) + t2 * (x2 * t5 - x1 * t6))
    c3 = t1 * (t3 * (t7 - t6) + t4 * (t5 - t7) + t2 * (t6 - t5))
    c4 = t1 * (t3 * (x2 - x3) + t4 * (x3 - x1) + t2 * (x1 - x2))
    t = x - xa[0]
    if m == 0:
        return ya[0] + t * (c2 + t * (c3 + c4 * t))
    if m == 1:
        return c2 + t * (2.0 * c3 + 3.0 * c4 * t)
    if m == 2:
        return 6.0 * c4 * t + 2.0 * c3
    if m == 3:
        return 6.0 * c4
    return 0.0


def dmat_from_radial(cr, rd):
    """dmat(io1, io2, l, ias) = d^(io1-1) u_(io2,l) / dr^(io1-1) at R from the last four mesh
    points of apwfr (match.f90:139-148 with np = max(apwordmax+1, 4) = 4)."""
    ao = cr.apwordmax
    assert max(ao + 1, 4) == 4
    d = np.zeros((ao, ao, cr.lmaxapw + 1, cr.natm), order="F")
    for ias in range(cr.natm):
        is_ = cr.spidx[ias]
        n = int(rd.nrmt[is_])
        xa = rd.spr[n - 4:n, is_]
        for l in range(cr.lmaxapw + 1):
            for io2 in range(ao):
                ya = rd.apwfr[n - 4:n, 0, io2, l, ias]
                for io1 in range(ao):
                    d[io1, io2, l, ias] = polynom(io1, xa, ya, float(cr.rmt[is_]))
    return d


def genrhomt(cr, rd, dm):
    """genrhomt.f90:34-150: densitymt(lmmaxvr, nrmtmax, natm) from dm(wfsize, wfsize, natm)
    (rows f = (l, io, m) m fastest, then the atom's lo functions (ilo, m)):
        rho_lm2(r) = sum_{a, b} u_a(r) u_b(r) sum_{m1, m3} Re(dm(f1, f3) conj(G(lm1, lm2, lm3)))
    over APW-APW pairs, twice the APW-lo ones, and the lo-lo ones, with G(lm1, lm2, lm3) =
    int conj(Y_lm1) R_lm2 Y_lm3 (listgnt(i, lm3, lm1), init1.f90:569, 619)."""
    ao, L = cr.apwordmax, cr.lmaxmat
    maxaa = ao * cr.lmmaxapw
    lmv = (rd.lmaxvr + 1) ** 2
    G = gaunt_table(max(L, int(cr.lorbl.max()) if cr.nlotot else 0), rd.lmaxvr)
    nrmax = rd.spr.shape[0]
    out = np.zeros((lmv, nrmax, cr.natm), order="F")
    for ias in range(cr.natm):
        is_ = cr.spidx[ias]
        n = int(rd.nrmt[is_])
        rho = np.zeros((lmv, n))
        # functions: (row offset in dm, l, u) — APW in dm order (l, io, m), then lo (ilo, m)
        funcs = []
        off = 0
        for l in range(L + 1):
            for io in range(ao):
                funcs.append((off, l, rd.apwfr[:n, 0, io, l, ias], False))
                off += 2 * l + 1
        off = maxaa
        for ilo in range(cr.nlorb[is_]):
            l = int(cr.lorbl[ilo, is_])
            funcs.append((off, l, rd.lofr[:n, 0, ilo, ias], True))
            off += 2 * l + 1
        for (o1, l1, u1, lo1) in funcs:
            for (o3, l3, u3, lo3) in funcs:
                if lo1 and not lo3:
                    continue                      # the lo-APW block enters as twice APW-lo
                w = 2.0 if (lo3 and not lo1) else 1.0
                fac = np.zeros(lmv)
                for m1 in range(-l1, l1 + 1):
                    for m3 in range(-l3, l3 + 1):
                        z = dm[o1 + m1 + l1, o3 + m3 + l3, ias]
                        g = G[l1 * l1 + l1 + m1, :lmv, l3 * l3 + l3 + m3]
                        fac += w * (z * np.conj(g)).real
                rho += fac[:, None] * (u1 * u3)[None, :]
        out[:, :n, ias] = rho
    return out
