"""Pins the oracle's vector-Jacobian product of (H, O) (oracle/lapw_ref.setup_vjp).

* central finite differences (step 1e-4 bohr) of L(tau) = Re sum_ij (D_ji H_ij + W_ji O_ij)
  with H and O rebuilt from moved atoms (structure factors and Theta(G));
* Hellmann-Feynman: with D = C f C^H and W = -C (f e) C^H from the eigenvectors, the same
  numbers are dE/dtau of E = sum_n f_n e_n, checked by finite differences of the
  eigenvalues (the contract's 1e-6 Ha/bohr);
* the restated reference forcek (forcek.f90:93-383) gives F = -dE/dtau;
* L is linear in V(G): a Hermitian perturbation of V(g), V(-g) changes L by exactly
  2 Re(dL/dV(g) delta).
"""
import numpy as np
import pytest

from conftest import small
from exciting_amd import synth
from oracle import lapw_ref as R

H_STEP = 1e-4


def _L(cr, kp, D, W):
    H, O, _ = R.setup(cr, kp)
    return float(np.real(np.sum(D.T * H) + np.sum(W.T * O)))


def _eig_cotangents(cr, kp, nst=12, seed=0):
    H, O, _ = R.setup(cr, kp)
    e, C = R.eigen(H, O, nst)
    f = np.random.default_rng(seed).uniform(0.0, 2.0, nst)
    D = (C * f) @ C.conj().T
    W = -(C * (f * e)) @ C.conj().T
    return e, C, f, D, W


def _moved(cr, kp, a, l, h):
    pos = cr.atposc.copy()
    pos[l, a] += h
    c2 = synth.with_positions(cr, pos)
    return c2, synth.kpoint_at(kp, c2)


@pytest.mark.parametrize("key", ["a", "b"])
def test_vjp_tau_finite_differences(key):
    w = small(key)
    cr, kp = w.crystal, w.kpoints[0]
    e, C, f, D, W = _eig_cotangents(cr, kp)
    g_tau, _ = R.setup_vjp(cr, kp, D, W)
    for a in range(cr.natm):
        for l in range(3):
            cp, kpp = _moved(cr, kp, a, l, H_STEP)
            cm, kpm = _moved(cr, kp, a, l, -H_STEP)
            fd = (_L(cp, kpp, D, W) - _L(cm, kpm, D, W)) / (2 * H_STEP)
            assert abs(fd - g_tau[l, a]) < 1e-7 * max(1.0, abs(fd)), (a, l, fd, g_tau[l, a])


def test_hellmann_feynman_eigenvalue_sum():
    w = small("a")
    cr, kp = w.crystal, w.kpoints[0]
    nst = 12
    e, C, f, D, W = _eig_cotangents(cr, kp, nst)
    g_tau, _ = R.setup_vjp(cr, kp, D, W)
    for a in range(cr.natm):
        for l in range(3):
            ev = []
            for h in (H_STEP, -H_STEP):
                c2, k2 = _moved(cr, kp, a, l, h)
                H2, O2, _ = R.setup(c2, k2)
                ev.append(np.sum(f * R.eigen(H2, O2, nst)[0]))
            fd = (ev[0] - ev[1]) / (2 * H_STEP)
            assert abs(fd - g_tau[l, a]) < 1e-6, (a, l, fd, g_tau[l, a])


@pytest.mark.parametrize("key", ["a", "b"])
def test_forcek_is_minus_gradient(key):
    w = small(key)
    cr, kp = w.crystal, w.kpoints[0]
    e, C, f, D, W = _eig_cotangents(cr, kp)
    g_tau, _ = R.setup_vjp(cr, kp, D, W)
    F = R.forcek_ibs(cr, kp, C, e, f)
    np.testing.assert_allclose(F, -g_tau, atol=1e-9 * max(1.0, np.abs(F).max()))


def test_vjp_potential_linearity():
    w = small("a")
    cr, kp = w.crystal, w.kpoints[0]
    e, C, f, D, W = _eig_cotangents(cr, kp)
    _, g_v = R.setup_vjp(cr, kp, D, W)
    L0 = _L(cr, kp, D, W)
    key = {tuple(cr.ivg[:, g]): g for g in range(cr.ngvec)}
    import copy
    for g in [0, 1, 5, 17, 40]:
        mg = key[tuple(-cr.ivg[:, g])]
        for delta in (1e-3, 1e-3j):
            c2 = copy.copy(cr)
            c2.veffig = cr.veffig.copy()
            c2.veffig[g] += delta
            if mg != g:
                c2.veffig[mg] += np.conj(delta)
            dL = _L(c2, kp, D, W) - L0
            pred = np.real(g_v[g] * delta + (g_v[mg] * np.conj(delta) if mg != g else 0.0))
            assert abs(dL - pred) < 1e-10, (g, delta, dL, pred)
