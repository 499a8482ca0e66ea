"""Pins the oracle's per-SCF producers of Theta(G) and V(G) (oracle.lapw_ref: igfft, gencfun,
genveffig, dft3).

* dft3 against the defining sum written out element by element (tiny grid), and the numpy
  transform the oracle uses against dft3, both directions (zfftifc.f90: forward sign -1 scaled
  by 1/N, backward +1 unscaled).
* igfft: a bijection of the G list onto the grid, G = 0 at 0, negative components wrapped.
* gencfun: Theta(G) equals the configuration's (pinned against a real-space integral in
  test_oracle_setup.py); the grid mean of Theta(r) is Theta(0); forward(backward) = identity
  for the G whose -G is in the box (Theta(r) is the real part, gencfun.f90).
* genveffig: a plane wave cos(G0.r) with Theta(r) = 1 gives 1/2 at +-G0 and nothing else.
* genrhoir: Parseval, (Omega/N) sum_r rho(r) = w sum_j occ_j sum_G |C_Gj|^2; psi_j at grid
  points by direct summation over G; states with |w occ| <= epsocc skipped.
"""
import math

import numpy as np

from conftest import small
from oracle import lapw_ref as R


def test_dft3_against_the_sum():
    n = [3, 4, 5]
    rng = np.random.default_rng(1)
    z = rng.standard_normal(60) + 1j * rng.standard_normal(60)
    for sgn in (-1, 1):
        out = np.zeros(60, dtype=np.complex128)
        for k3 in range(5):
            for k2 in range(4):
                for k1 in range(3):
                    s = 0j
                    for r3 in range(5):
                        for r2 in range(4):
                            for r1 in range(3):
                                ph = sgn * 2 * math.pi * (k1 * r1 / 3 + k2 * r2 / 4 + k3 * r3 / 5)
                                s += z[r1 + 3 * (r2 + 4 * r3)] * complex(math.cos(ph), math.sin(ph))
                    out[k1 + 3 * (k2 + 4 * k3)] = s / 60 if sgn < 0 else s
        assert np.abs(R.dft3(z, n, sgn) - out).max() <= 1e-12
        assert np.abs(R._fft3(z, n, sgn) - out).max() <= 1e-12


def test_igfft_is_a_bijection():
    cr = small("a").crystal
    ix = R.igfft(cr)
    n = R.fft_grid(cr)
    assert ix[0] == 0 and (cr.ivg[:, 0] == 0).all()
    assert len(set(ix.tolist())) == cr.ngrtot == n[0] * n[1] * n[2]
    g = np.nonzero((cr.ivg[0] == -1) & (cr.ivg[1] == 2) & (cr.ivg[2] == -3))[0][0]
    assert ix[g] == (n[0] - 1) + n[0] * (2 + n[1] * (n[2] - 3))


def test_gencfun_properties():
    cr = small("d").crystal
    cf, cir = R.gencfun(cr)
    assert np.abs(cf - cr.cfunig).max() <= 1e-14
    assert abs(cir.mean() - cf[0].real) <= 1e-13
    back = R._fft3(cir, R.fft_grid(cr), -1)[R.igfft(cr)]
    # Theta(r) keeps the real part of the backward transform (gencfun.f90); on an even grid the
    # Nyquist plane G_c = n_c/2 has no -G partner in the box, so the identity holds for the
    # G whose -G is in the box
    paired = np.all(-cr.ivg >= cr.intgv[:, :1], axis=0)
    assert (~paired).any() and paired.sum() > cr.ngvec
    assert np.abs(back - cf)[paired].max() <= 1e-13


def test_gencfun_damping():
    """cfdamp > 0 (gencfun.f90:108-114): Theta(G) is multiplied by a Gaussian in |G| whose
    width is gmaxvr / cfdamp, after Theta(r) is taken: Theta(r) and Theta(0) unchanged, and
    log(damped / undamped) against |G|^2 from the integer vectors (bvec ivg, not the crystal's
    gc) is a line through the origin of slope -(cfdamp / gmaxvr)^2 — the ratio at |G| = gmaxvr
    is exp(-cfdamp^2)."""
    cr = small("d").crystal
    cf0, cir0 = R.gencfun(cr)
    cfdamp = 0.7
    cf, cir = R.gencfun(cr, cfdamp=cfdamp, gmaxvr=cr.gmaxvr)
    assert np.array_equal(cir, cir0) and cf[0] == cf0[0]
    g2 = np.sum((cr.bvec @ cr.ivg.astype(float)) ** 2, axis=0)
    ok = np.abs(cf0) > 1e-8
    lr = np.log(np.abs(cf[ok] / cf0[ok]))
    slope = np.sum(lr * g2[ok]) / np.sum(g2[ok] ** 2)
    assert abs(slope + (cfdamp / cr.gmaxvr) ** 2) <= 1e-12
    assert np.abs(lr - slope * g2[ok]).max() <= 1e-10
    assert np.abs(np.angle(cf[ok] / cf0[ok])).max() <= 1e-14       # a real positive factor
    # the ratio where |G| is nearest gmaxvr
    i = np.argmin(np.abs(np.sqrt(g2) - cr.gmaxvr) + ~ok)
    assert abs(abs(cf[i] / cf0[i]) - np.exp(-(cfdamp * np.sqrt(g2[i]) / cr.gmaxvr) ** 2)) <= 1e-13


def test_genveffig_plane_wave():
    cr = small("d").crystal
    n = R.fft_grid(cr)
    g0 = np.nonzero((cr.ivg[0] == 1) & (cr.ivg[1] == -2) & (cr.ivg[2] == 0))[0][0]
    assert g0 < cr.ngvec
    r1, r2, r3 = np.meshgrid(np.arange(n[0]) / n[0], np.arange(n[1]) / n[1], np.arange(n[2]) / n[2], indexing="ij")
    v = np.cos(2 * math.pi * (1 * r1 - 2 * r2 + 0 * r3)).ravel(order="F")
    vg = R.genveffig(cr, v, np.ones(cr.ngrtot))
    gm = np.nonzero((cr.ivg[0] == -1) & (cr.ivg[1] == 2) & (cr.ivg[2] == 0))[0][0]
    expect = np.zeros(cr.ngvec, dtype=np.complex128)
    expect[g0] = expect[gm] = 0.5
    assert np.abs(vg - expect).max() <= 1e-13


def test_genrhoir_parseval_and_direct_sum():
    w = small("d")
    cr, kp = w.crystal, w.kpoints[0]
    rng = np.random.default_rng(8)
    nmat, nst = kp.ngp + cr.nlotot, 5
    C = rng.standard_normal((nmat, nst)) + 1j * rng.standard_normal((nmat, nst))
    occ = np.array([2.0, 1.5, 0.0, 1e-12, 0.7])
    wk = 0.4
    rho = R.genrhoir(cr, kp, C, occ, wk)
    n = R.fft_grid(cr)
    N = n[0] * n[1] * n[2]
    kept = np.abs(wk * occ) > 1e-8
    par = wk * sum(occ[j] * np.sum(np.abs(C[:kp.ngp, j]) ** 2) for j in range(nst) if kept[j])
    assert abs(cr.omega / N * rho.sum() - par) <= 1e-12 * par
    # direct sum at a few grid points: psi(r) = sum_G C_G e^{i G.r}, r = sum_i (j_i / n_i) a_i
    G = cr.ivg[:, kp.igpig[:kp.ngp] - 1]
    for (j1, j2, j3) in [(0, 0, 0), (3, 5, 7), (n[0] - 1, 1, n[2] - 2)]:
        frac = np.array([j1 / n[0], j2 / n[1], j3 / n[2]])
        ph = np.exp(2j * np.pi * (frac @ G))
        val = sum((wk * occ[j] / cr.omega) * abs(np.dot(C[:kp.ngp, j], ph)) ** 2 for j in range(nst) if kept[j])
        assert abs(rho[j1 + n[0] * (j2 + n[1] * j3)] - val) <= 1e-12 * max(val, 1e-300) + 1e-14
