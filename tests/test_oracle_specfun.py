"""Pins the oracle's special functions (oracle/specfun.py).

* against scipy.special (an independent implementation),
* against closed forms (j_0, j_1, j_2; Y_00, Y_10, Y_11),
* against the three-term recurrence and the derivative identity,
* against the reference's own Fortran routines (src/src_lib/sbesseldm.f90, genylm.f90,
  sphcrd.f90) compiled by oracle/Makefile into oracle/_ref, when that library exists, and
  against the committed fixtures they produced (tests/golden/specfun_ref.npz).
"""
import ctypes as C
import math
import os

import numpy as np
import pytest
import scipy.special as sp

from oracle import specfun

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFLIB = os.path.join(ROOT, "oracle", "_ref", "libexcref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden", "specfun_ref.npz")

XS = np.concatenate([[0.0, 1e-12, 1e-9, 1e-8, 2e-8, 1e-5, 0.01, 0.3, 0.999, 1.0, 1.001],
                     np.linspace(0.05, 40.0, 397), [math.pi, 2 * math.pi, 3 * math.pi, 9.9999, 10.0, 10.0001]])


def _close(a, b, rtol=1e-13, atol=1e-15):
    np.testing.assert_allclose(a, b, rtol=rtol, atol=atol)


@pytest.mark.parametrize("lmax", [0, 1, 8, 10, 14])
def test_sbessel_vs_scipy(lmax):
    j = specfun.sbessel_array(lmax, XS)
    for l in range(lmax + 1):
        _close(j[l], sp.spherical_jn(l, XS), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("lmax", [1, 8, 10, 14])
def test_sbessel_derivative_vs_scipy(lmax):
    d = specfun.sbesseld1_array(lmax, XS)
    for l in range(lmax + 1):
        _close(d[l], sp.spherical_jn(l, XS, derivative=True), rtol=1e-11, atol=1e-14)


def test_sbessel_closed_forms_and_recurrence():
    x = np.linspace(0.2, 30.0, 101)
    j = specfun.sbessel_array(12, x)
    s, c = np.sin(x), np.cos(x)
    _close(j[0], s / x)
    _close(j[1], s / x ** 2 - c / x, rtol=1e-12)
    _close(j[2], (3 / x ** 2 - 1) * s / x - 3 * c / x ** 2, rtol=1e-11, atol=1e-14)
    for l in range(1, 12):
        _close(j[l - 1] + j[l + 1], (2 * l + 1) / x * j[l], rtol=1e-10, atol=1e-14)


def test_ylm_vs_scipy_and_closed_forms():
    rng = np.random.default_rng(3)
    th = np.concatenate([[0.0, math.pi, math.pi / 2], rng.uniform(0, math.pi, 200)])
    ph = np.concatenate([[0.0, 1.0, 2 * math.pi - 1e-3], rng.uniform(0, 2 * math.pi, 200)])
    L = 14
    y = specfun.genylm(L, th, ph)
    for l in range(L + 1):
        for m in range(-l, l + 1):
            _close(y[l * (l + 1) + m], sp.sph_harm_y(l, m, th, ph), rtol=1e-11, atol=1e-13)
    # closed forms with the Condon–Shortley phase
    _close(y[0], np.full(th.size, 0.5 / math.sqrt(math.pi)))
    _close(y[2], math.sqrt(3 / (4 * math.pi)) * np.cos(th))
    _close(y[3], -math.sqrt(3 / (8 * math.pi)) * np.sin(th) * np.exp(1j * ph))
    # orthonormality on a product quadrature
    xg, wg = np.polynomial.legendre.leggauss(24)
    nphi = 48
    T, P = np.meshgrid(np.arccos(xg), np.arange(nphi) * 2 * math.pi / nphi, indexing="ij")
    W = np.repeat(wg[:, None], nphi, axis=1) * (2 * math.pi / nphi)
    Y = specfun.genylm(8, T.ravel(), P.ravel())
    G = (Y.conj() * W.ravel()) @ Y.T
    _close(G, np.eye(81), atol=1e-12)


def _reflib():
    if not os.path.exists(REFLIB):
        return None
    lib = C.CDLL(REFLIB)
    lib.sbesseldm_.restype = None
    lib.genylm_.restype = None
    return lib


def ref_sbesseldm(lib, m, lmax, x):
    out = np.zeros(lmax + 1)
    lib.sbesseldm_(C.byref(C.c_int(m)), C.byref(C.c_int(lmax)), C.byref(C.c_double(x)),
                   out.ctypes.data_as(C.c_void_p))
    return out


def ref_genylm(lib, lmax, theta, phi):
    tp = np.array([theta, phi])
    out = np.zeros((lmax + 1) ** 2, dtype=np.complex128)
    lib.genylm_(C.byref(C.c_int(lmax)), tp.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return out


def _check_against(jd, jd1, ylm, xs, tps, lmax):
    j = specfun.sbessel_array(lmax, xs)
    d = specfun.sbesseld1_array(lmax, xs)
    _close(j.T, jd, rtol=1e-12, atol=1e-15)
    _close(d.T, jd1, rtol=1e-11, atol=1e-14)
    y = specfun.genylm(lmax, tps[:, 0], tps[:, 1])
    _close(y.T, ylm, rtol=1e-12, atol=1e-14)


def test_against_reference_fortran():
    """The restatement agrees with the reference's compiled sbesseldm/genylm (oracle/_ref)."""
    lib = _reflib()
    if lib is None:
        pytest.skip("oracle/_ref not built (needs the reference checkout; see oracle/Makefile)")
    lmax = 10
    xs = XS[XS <= 1e5]
    jd = np.array([ref_sbesseldm(lib, 0, lmax, x) for x in xs])
    jd1 = np.array([ref_sbesseldm(lib, 1, lmax, x) for x in xs])
    rng = np.random.default_rng(5)
    tps = np.stack([rng.uniform(0, math.pi, 50), rng.uniform(0, 2 * math.pi, 50)], axis=1)
    ylm = np.array([ref_genylm(lib, lmax, t, p) for t, p in tps])
    _check_against(jd, jd1, ylm, xs, tps, lmax)


def test_against_reference_fixtures():
    """Same check against the committed vectors that oracle/_ref produced."""
    if not os.path.exists(GOLDEN):
        pytest.skip("golden fixture not generated")
    g = np.load(GOLDEN)
    _check_against(g["jl"], g["djl"], g["ylm"], g["x"], g["tp"], int(g["lmax"]))
