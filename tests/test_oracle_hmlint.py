"""Pins the oracle's muffin-tin integrals (oracle/hmlint_ref.py: hmlint = mt_kin + mt_pot,
olprad).

* The cubic-spline integral (fderiv.f90, m = -1), the Gaunt coefficients (gaunt.f90) and the
  real spherical harmonics (genrlm.f90) against the reference's own routines compiled from
  its sources by oracle/Makefile (oracle/_ref), and against the committed vectors they
  produced (tests/golden/hmlint_ref.npz).
* Gaunt coefficients against a product quadrature of conj(Y) Y Y on the sphere, with the
  oracle's Y_lm (pinned in test_oracle_specfun.py).
* The spline rule against a closed-form radial integral (convergence on refined meshes).
* The assembled h_a, loa, lolo against a direct evaluation of <u1 Y_lm1 | V | u3 Y_lm3>:
  V(r, rhat) = sum_lm veffmt(lm, r) R_lm(rhat) summed on an angular quadrature grid at every
  mesh point, then the same radial rule, plus the radial kinetic form for lm1 = lm3.  This
  fixes the Gaunt combination for real harmonics, the conjugated side and the row/column
  roles independently of the Gaunt tables.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from conftest import small
from exciting_amd import synth
from oracle import hmlint_ref as M
from oracle import specfun

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFLIB = os.path.join(ROOT, "oracle", "_ref", "libexcref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden", "hmlint_ref.npz")


def _reflib():
    if not os.path.exists(REFLIB):
        return None
    lib = C.CDLL(REFLIB)
    if not hasattr(lib, "fderiv_") or not hasattr(lib, "gaunt_"):
        return None
    lib.fderiv_.restype = None
    lib.genrlm_.restype = None
    lib.gaunt_.restype = C.c_double
    lib.polynom_.restype = C.c_double
    return lib


def ref_polynom(lib, m, xa, ya, x):
    xa = np.ascontiguousarray(xa, dtype=np.float64)
    ya = np.ascontiguousarray(ya, dtype=np.float64)
    c = np.zeros(len(xa))
    return lib.polynom_(C.byref(C.c_int(m)), C.byref(C.c_int(len(xa))), xa.ctypes.data_as(C.c_void_p),
                        ya.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p), C.byref(C.c_double(x)))


def ref_fderiv_int(lib, x, f):
    n = len(x)
    x = np.ascontiguousarray(x, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    g = np.zeros(n)
    cf = np.zeros((n, 3))
    lib.fderiv_(C.byref(C.c_int(-1)), C.byref(C.c_int(n)), x.ctypes.data_as(C.c_void_p),
                f.ctypes.data_as(C.c_void_p), g.ctypes.data_as(C.c_void_p), cf.ctypes.data_as(C.c_void_p))
    return g[-1]


def ref_gaunt(lib, l1, l2, l3, m1, m2, m3):
    a = [C.c_int(v) for v in (l1, l2, l3, m1, m2, m3)]
    return lib.gaunt_(*[C.byref(v) for v in a])


def ref_genrlm(lib, lmax, theta, phi):
    tp = np.array([theta, phi])
    out = np.zeros((lmax + 1) ** 2)
    lib.genrlm_(C.byref(C.c_int(lmax)), tp.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return out


def fixture_inputs():
    """The fixed inputs of the reference comparison (also used by tests/golden/make_golden.py)."""
    rng = np.random.default_rng(17)
    meshes, funcs = [], []
    for n in (7, 50, 200, 401):
        x = 1e-5 * np.exp(np.arange(n) * math.log(2.0 / 1e-5) / (n - 1))
        meshes.append(x)
        a, b, c = rng.uniform(-1, 1), rng.uniform(0.5, 2.0), rng.uniform(-3, 3)
        funcs.append(x * x * (1 + a * x) * np.exp(-b * x) + c / (1.0 + x))
    args = []
    for l1 in range(0, 11, 3):
        for l2 in range(0, 7):
            for l3 in range(0, 11, 2):
                for m1 in range(-l1, l1 + 1, max(1, l1)):
                    for m3 in range(-l3, l3 + 1, max(1, l3 // 2)):
                        m2 = m1 - m3
                        if abs(m2) <= l2:
                            args.append((l1, l2, l3, m1, m2, m3))
    tps = np.stack([rng.uniform(0, math.pi, 20), rng.uniform(0, 2 * math.pi, 20)], axis=1)
    return meshes, funcs, np.array(args, dtype=np.int64), tps


def polynom_inputs():
    """Four-point stencils at the end of logarithmic meshes, as match uses them."""
    rng = np.random.default_rng(23)
    out = []
    for n in (60, 300, 700):
        x = 1e-5 * np.exp(np.arange(n) * math.log(2.2 / 1e-5) / (n - 1))
        xa = x[-4:]
        ya = np.sin(rng.uniform(0.5, 2.0) * xa) * np.exp(-xa) + rng.uniform(-1, 1)
        out.append((xa, ya, x[-1]))
    return out


def _check(ints, gaunts, rlms, polys=None):
    meshes, funcs, args, tps = fixture_inputs()
    if polys is not None:
        for (xa, ya, x), ref in zip(polynom_inputs(), polys):
            for m in range(4):
                assert abs(M.polynom(m, xa, ya, x) - ref[m]) <= 1e-12 * max(1.0, abs(ref[m]))
    for x, f, g in zip(meshes, funcs, ints):
        assert abs(M.fderiv_int(x, f) - g) <= 1e-13 * max(1.0, abs(g))
    ours = np.array([M.gaunt(*map(int, a)) for a in args])
    np.testing.assert_allclose(ours, gaunts, rtol=1e-12, atol=1e-15)
    for (t, p), r in zip(tps, rlms):
        np.testing.assert_allclose(M.rlm(8, t, p)[0], r, rtol=1e-12, atol=1e-14)


def test_against_reference_fortran():
    lib = _reflib()
    if lib is None:
        pytest.skip("oracle/_ref not built (needs the reference checkout; see oracle/Makefile)")
    meshes, funcs, args, tps = fixture_inputs()
    ints = [ref_fderiv_int(lib, x, f) for x, f in zip(meshes, funcs)]
    gaunts = np.array([ref_gaunt(lib, *map(int, a)) for a in args])
    rlms = np.array([ref_genrlm(lib, 8, t, p) for t, p in tps])
    assert np.count_nonzero(gaunts) > 50
    polys = [[ref_polynom(lib, m, xa, ya, x) for m in range(4)] for xa, ya, x in polynom_inputs()]
    _check(ints, gaunts, rlms, polys)


def test_against_reference_fixtures():
    if not os.path.exists(GOLDEN):
        pytest.skip("golden fixture not generated")
    g = np.load(GOLDEN)
    _check(list(g["ints"]), g["gaunt"], g["rlm"], g["polynom"] if "polynom" in g else None)


def test_polynom_exact_on_cubics():
    """The cubic through four points reproduces any cubic and its derivatives."""
    rng = np.random.default_rng(2)
    for _ in range(20):
        a = rng.normal(size=4)
        xa = np.sort(rng.uniform(1.0, 2.0, 4))
        x = rng.uniform(1.0, 2.2)
        p = np.polynomial.Polynomial(a)
        for m in range(4):
            ref = p.deriv(m)(x) if m else p(x)
            assert abs(M.polynom(m, xa, p(xa), x) - ref) <= 1e-9 * max(1.0, abs(ref))


def _sphere_grid(nt=24, nphi=52):
    xg, wg = np.polynomial.legendre.leggauss(nt)
    T, P = np.meshgrid(np.arccos(xg), np.arange(nphi) * 2 * math.pi / nphi, indexing="ij")
    W = np.repeat(wg[:, None], nphi, axis=1) * (2 * math.pi / nphi)
    return T.ravel(), P.ravel(), W.ravel()


def test_gaunt_by_quadrature():
    T, P, W = _sphere_grid()
    Y = specfun.genylm(10, T, P)                                  # (lm, npt)
    R = M.rlm(6, T, P).T
    rng = np.random.default_rng(4)
    for _ in range(60):
        l1, l3 = rng.integers(0, 11, 2)
        l2 = rng.integers(0, 7)
        m1, m3 = rng.integers(-l1, l1 + 1), rng.integers(-l3, l3 + 1)
        m2 = m1 - m3 if abs(m1 - m3) <= l2 else 0
        i1, i2, i3 = l1 * l1 + l1 + m1, l2 * l2 + l2 + m2, l3 * l3 + l3 + m3
        q = np.sum(W * np.conj(Y[i1]) * Y[i2] * Y[i3])
        assert abs(q - M.gaunt(int(l1), int(l2), int(l3), int(m1), int(m2), int(m3))) <= 1e-13
        for mm in (m2, -m2, abs(m2)):
            j2 = l2 * l2 + l2 + mm
            qr = np.sum(W * np.conj(Y[i1]) * R[j2] * Y[i3])
            assert abs(qr - M.gauntyry(int(l1), int(l2), int(l3), int(m1), int(mm), int(m3))) <= 1e-13


def test_spline_rule_converges_to_closed_form():
    # int_0^2 r^2 e^-r dr = 2 - 10 e^-2
    exact = 2.0 - 10.0 * math.exp(-2.0)
    errs = []
    for n in (200, 400, 800, 1600):
        x = 1e-6 * np.exp(np.arange(n) * math.log(2.0 / 1e-6) / (n - 1))
        errs.append(abs(M.fderiv_int(x, x * x * np.exp(-x)) - exact))
    # fourth order: the error falls by ~16 per doubling of the mesh (measured 14.8, 15.5, 15.7)
    assert all(a / b > 12.0 for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-9


def test_hmlint_elements_by_quadrature():
    w = small("d")
    cr = w.crystal
    rd = synth.make_radial(cr, nr=(150, 120), lmaxvr=4)
    haa, loa, lolo = M.hmlint(cr, rd)
    T, P, W = _sphere_grid()
    ao = cr.apwordmax
    for ias in (0, cr.natm - 1):
        is_ = cr.spidx[ias]
        n = int(rd.nrmt[is_])
        x = rd.spr[:n, is_]
        Y = specfun.genylm(cr.lmaxmat, T, P)
        Rl = M.rlm(rd.lmaxvr, T, P)                                # (npt, lmv)
        Vang = rd.veffmt[:, :n, ias].T @ Rl.T                       # (nr, npt): V(r, rhat)

        def elem(u1, du1, l1, m1, u3, du3, l3, m3):
            i1, i3 = l1 * l1 + l1 + m1, l3 * l3 + l3 + m3
            A = Vang @ (W * np.conj(Y[i1]) * Y[i3])                # angular integral at each r
            pot = M.fderiv_int(x, (u1 * u3 * x * x)[:, None] * np.stack([A.real, A.imag], 1))
            z = pot[0] + 1j * pot[1]
            if i1 == i3:
                z += M.fderiv_int(x, (0.5 * du1 * du3 + 0.5 * l1 * (l1 + 1) * u1 * u3 / (x * x)) * x * x)
            return z

        def apw(io, l):
            return rd.apwfr[:n, 0, io, l, ias], rd.apwfr[:n, 1, io, l, ias]

        def lo(ilo):
            return rd.lofr[:n, 0, ilo, ias], rd.lofr[:n, 1, ilo, ias]

        scale = np.abs(haa[:, :, ias]).max()
        for (l1, m1, io1, l3, m3, io3) in [(0, 0, 0, 0, 0, 1), (1, -1, 1, 2, 1, 0), (2, 2, 0, 2, 2, 0),
                                           (3, -2, 1, 1, 0, 1), (cr.lmaxmat, 1, 0, cr.lmaxmat - 2, -1, 1)]:
            f1, f3 = io1 + ao * (l1 * l1 + l1 + m1), io3 + ao * (l3 * l3 + l3 + m3)
            z = elem(*apw(io1, l1), l1, m1, *apw(io3, l3), l3, m3)
            assert abs(haa[f1, f3, ias] - z) <= 1e-11 * scale, (l1, m1, io1, l3, m3, io3)
        # lo rows: (ilo, m) m fastest
        r = 0
        for ilo in range(cr.nlorb[is_]):
            l1 = int(cr.lorbl[ilo, is_])
            for m1 in range(-l1, l1 + 1):
                for (l3, m3, io3) in [(l1, m1, 0), (1, 0, 1), (2, -1, 0)]:
                    f3 = io3 + ao * (l3 * l3 + l3 + m3)
                    z = elem(*lo(ilo), l1, m1, *apw(io3, l3), l3, m3)
                    assert abs(loa[r, f3, ias] - z) <= 1e-11 * scale
                r3 = 0
                for ilo2 in range(cr.nlorb[is_]):
                    l3 = int(cr.lorbl[ilo2, is_])
                    for m3 in range(-l3, l3 + 1):
                        z = elem(*lo(ilo), l1, m1, *lo(ilo2), l3, m3)
                        assert abs(lolo[r, r3, ias] - z) <= 1e-11 * scale
                        r3 += 1
                r += 1


def test_genrhomt_reproduces_potential_energy():
    """The muffin-tin density from the density matrix (genrhomt) integrated against the
    potential equals tr(dm h_pot), with h_pot the potential part of the oracle's hmlint (its
    matrix minus the one at V = 0) in the density matrix's order: sum_lm int rho_lm V_lm r^2 dr
    = sum_ff' dm(f, f') h_pot(f', f) for any coefficients, by the same radial rule."""
    import dataclasses
    from oracle import lapw_ref as R
    w = small("d")
    cr, kp = w.crystal, w.kpoints[0]
    rd = synth.make_radial(cr, nr=(120, 100), lmaxvr=4)
    apw = R.match(cr, kp)
    rng = np.random.default_rng(31)
    nmat, nst = kp.ngp + cr.nlotot, 7
    C = rng.standard_normal((nmat, nst)) + 1j * rng.standard_normal((nmat, nst))
    dm = R.gendmatmt(cr, apw, kp.ngp, C, rng.uniform(0, 2, nst), 0.7)
    rho = M.genrhomt(cr, rd, dm)
    h1 = M.hmlint(cr, rd)
    h0 = M.hmlint(cr, dataclasses.replace(rd, veffmt=np.zeros_like(rd.veffmt)))
    order = R.dmat_order(cr)
    naa, maxaa, ws = len(order), cr.apwordmax * cr.lmmaxapw, R.dmat_wfsize(cr)
    e_rho = e_dm = 0.0
    for a in range(cr.natm):
        is_ = cr.spidx[a]
        n = int(rd.nrmt[is_])
        x = rd.spr[:n, is_]
        e_rho += sum(M.fderiv_int(x, rho[q, :n, a] * rd.veffmt[q, :n, a] * x * x) for q in range(rho.shape[0]))
        hp = [h1[i][..., a] - h0[i][..., a] for i in range(3)]
        Hd = np.zeros((ws, ws), dtype=np.complex128)
        Hd[:naa, :naa] = hp[0][np.ix_(order, order)]
        ns = cr.losize[is_] if cr.nlorb[is_] else 0
        Hd[maxaa:maxaa + ns, :naa] = hp[1][:ns][:, order]
        Hd[:naa, maxaa:maxaa + ns] = hp[1][:ns][:, order].conj().T
        Hd[maxaa:maxaa + ns, maxaa:maxaa + ns] = hp[2][:ns, :ns]
        e_dm += np.trace(dm[:, :, a] @ Hd).real
    assert abs(e_rho - e_dm) <= 1e-11 * abs(e_dm), (e_rho, e_dm)
