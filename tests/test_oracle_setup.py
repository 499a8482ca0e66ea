"""Pins the oracle's setup restatement (oracle/lapw_ref.py) by routes independent of it.

* match: the matching conditions themselves — the lm-projection of the plane wave
  exp(i(G+k).r)/sqrt(Omega) on the muffin-tin sphere (numerical sphere quadrature with
  scipy's Y_lm) and its radial derivative must equal sum_io A u_io(R), sum_io A u_io'(R).
* H/O muffin-tin + lo parts: quadratic forms c^H M c evaluated through matrix-vector
  products of the definitions, with explicit loops over the local orbitals.
* interstitial: a dictionary lookup of G_i - G_j (not the ivgig table), the free-electron
  limit (V = 0, no spheres: H = |G+k|^2/2, O = 1), and Theta(G) against a real-space
  integral of the step function on a grid.
* assembly: exact Hermiticity from the upper triangle, positive-definite O.
* eigen: zhegv against scipy.linalg.eigh.
"""
import math

import numpy as np
import pytest
import scipy.linalg
import scipy.special as sp

from conftest import small, workload
from oracle import lapw_ref as R


def _sphere_quadrature(n=40):
    xg, wg = np.polynomial.legendre.leggauss(n)
    nphi = 2 * n
    th = np.arccos(xg)
    ph = np.arange(nphi) * 2 * math.pi / nphi
    T, P = np.meshgrid(th, ph, indexing="ij")
    W = np.repeat(wg[:, None], nphi, axis=1) * (2 * math.pi / nphi)
    rhat = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], axis=0).reshape(3, -1)
    return T.ravel(), P.ravel(), W.ravel(), rhat


@pytest.mark.parametrize("key", ["a", "b"])
def test_match_satisfies_matching_conditions(key):
    w = small(key)
    cr, kp = w.crystal, w.kpoints[0]
    A = R.match(cr, kp)
    T, P, W, rhat = _sphere_quadrature(48)
    L = cr.lmaxapw
    rng = np.random.default_rng(0)
    for ias in range(cr.natm):
        Rmt = cr.rmt[cr.spidx[ias]]
        for ig in rng.choice(kp.ngp, 6, replace=False):
            q = kp.vgpc[:, ig]
            pw = np.exp(1j * (q @ (cr.atposc[:, ias][:, None] + Rmt * rhat))) / math.sqrt(cr.omega)
            dpw = 1j * (q @ rhat) * pw                      # d/dR
            for l in range(L + 1):
                for m in range(-l, l + 1):
                    y = sp.sph_harm_y(l, m, T, P)
                    val = np.sum(W * np.conj(y) * pw)
                    der = np.sum(W * np.conj(y) * dpw)
                    lm = l * (l + 1) + m
                    a = A[ig, :, lm, ias]
                    d = cr.dmat[:, :, l, ias]
                    np.testing.assert_allclose(d[0] @ a, val, atol=1e-12, rtol=1e-10)
                    np.testing.assert_allclose(d[1] @ a, der, atol=1e-12, rtol=1e-10)


def _gdict(cr):
    return {tuple(cr.ivg[:, g]): g for g in range(cr.ngvec)}


def _interstitial_dense(cr, kp):
    """Theta(G_i-G_j), V(G_i-G_j) and (G_i+k).(G_j+k)/2 by dictionary lookup, full matrix."""
    n = kp.ngp
    gd = _gdict(cr)
    th = np.zeros((n, n), dtype=np.complex128)
    vv = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        gi = cr.ivg[:, kp.igpig[i] - 1]
        for j in range(n):
            g = gd.get(tuple(gi - cr.ivg[:, kp.igpig[j] - 1]))
            if g is not None:
                th[i, j] = cr.cfunig[g]
                vv[i, j] = cr.veffig[g]
    dp = 0.5 * (kp.vgpc[:, :n].T @ kp.vgpc[:, :n])
    return th, vv, dp


def _lo_list(cr):
    out = []
    for ias in range(cr.natm):
        is_ = cr.spidx[ias]
        for ilo in range(cr.nlorb[is_]):
            l = cr.lorbl[ilo, is_]
            for m in range(-l, l + 1):
                out.append((ias, ilo, l, m))
    return out


@pytest.mark.parametrize("key", ["a", "b", "d"])
def test_quadratic_forms(key):
    w = small(key)
    cr = w.crystal
    for kp in w.kpoints:
        H, O, A = R.setup(cr, kp)
        n = kp.ngp
        nmat = n + cr.nlotot
        rng = np.random.default_rng(1)
        c = rng.standard_normal(nmat) + 1j * rng.standard_normal(nmat)
        ca, cl = c[:n], c[n:]
        th, vv, dp = _interstitial_dense(cr, kp)
        # interstitial uses the upper triangle then mirrors: build that Hermitian matrix
        HI = np.triu(vv + dp * th)
        HI = HI + np.triu(HI, 1).conj().T
        OI = np.triu(th)
        OI = OI + np.triu(OI, 1).conj().T
        qo = np.vdot(ca, OI @ ca)
        qh = np.vdot(ca, HI @ ca)
        los = _lo_list(cr)
        for ias in range(cr.natm):
            Aa = A[:n, :, :, ias].reshape(n, -1, order="F")          # (G, f) with f = (io, lm)
            u = Aa.T @ ca                                             # A_a^T c
            qo += np.vdot(u, u)
            qh += np.vdot(u, cr.haa[:, :, ias] @ u)
        # APW-lo and lo-lo, explicit loops over lo functions
        for p, (ias, ilo, l, m) in enumerate(los):
            row = sum(1 for q in los[:p] if q[0] == ias)
            lm = l * (l + 1) + m
            hcol = np.array([np.sum(cr.loa[row, :, ias] * A[i, :, :, ias].reshape(-1, order="F"))
                             for i in range(n)])                     # H(lo, i)
            ocol = np.array([np.sum(np.conj(A[i, :, lm, ias]) * cr.oalo[:, ilo, ias]) for i in range(n)])  # O(i, lo)
            qh += 2 * np.real(np.conj(cl[p]) * np.sum(hcol * ca))
            qo += 2 * np.real(np.sum(np.conj(ca) * ocol) * cl[p])
            for p2, (ias2, ilo2, l2, m2) in enumerate(los):
                if ias2 != ias:
                    continue
                row2 = sum(1 for q in los[:p2] if q[0] == ias2)
                hv = cr.lolo[row, row2, ias] if p <= p2 else np.conj(cr.lolo[row2, row, ias])
                qh += np.conj(cl[p]) * hv * cl[p2]
                if l2 == l and m2 == m:
                    ov = cr.ololo[ilo, ilo2, ias] if p <= p2 else cr.ololo[ilo2, ilo, ias]
                    qo += np.conj(cl[p]) * ov * cl[p2]
        np.testing.assert_allclose(np.vdot(c, H @ c), qh, rtol=1e-11)
        np.testing.assert_allclose(np.vdot(c, O @ c), qo, rtol=1e-11)


def test_hermitian_and_positive_definite():
    for key in "abcd":
        w = small(key)
        for kp in w.kpoints:
            H, O, _ = R.setup(w.crystal, kp)
            il = np.tril_indices(H.shape[0], -1)
            assert np.array_equal(H[il], np.conj(H.T[il])) and np.array_equal(O[il], np.conj(O.T[il]))
            # the diagonal keeps the rounding-level imaginary part of the MT product
            assert np.abs(np.diag(H).imag).max() < 1e-12 * np.abs(H).max()
            np.linalg.cholesky(O)


def test_free_electron_limit():
    """V = 0 and Theta = delta_G0 (no spheres): the interstitial blocks are |G+k|^2/2 and 1."""
    w = small("a")
    cr, kp = w.crystal, w.kpoints[0]
    import copy
    c2 = copy.copy(cr)
    c2.cfunig = np.zeros_like(cr.cfunig)
    c2.cfunig[0] = 1.0
    c2.veffig = np.zeros_like(cr.veffig)
    n = kp.ngp
    H = np.zeros((n, n), dtype=np.complex128)
    O = np.zeros((n, n), dtype=np.complex128)
    R.hmlistln(c2, kp, H)
    R.olpistln(c2, kp, O)
    np.testing.assert_allclose(H, np.diag(0.5 * (kp.vgpc[:, :n] ** 2).sum(axis=0)), atol=1e-15)
    np.testing.assert_allclose(O, np.eye(n), atol=0)


def test_meffig_kinetic_term():
    """ValenceRelativity /= 'none' (hmlistln.f90:61-72): meffig takes Theta's place in H's
    kinetic term only.  meffig = delta_G0 with Theta kept gives the free-electron kinetic
    diagonal whatever Theta is, and meffig = cfunig gives the 'none' matrix bit for bit."""
    w = small("a")
    cr, kp = w.crystal, w.kpoints[0]
    import copy
    c2 = copy.copy(cr)
    c2.veffig = np.zeros_like(cr.veffig)
    n = kp.ngp
    m = np.zeros(cr.ngvec, dtype=np.complex128)
    m[0] = 1.0
    H = np.zeros((n, n), dtype=np.complex128)
    R.hmlistln(c2, kp, H, meffig=m)
    np.testing.assert_allclose(H, np.diag(0.5 * (kp.vgpc[:, :n] ** 2).sum(axis=0)), atol=1e-15)
    H0 = np.zeros((n, n), dtype=np.complex128)
    H1 = np.zeros((n, n), dtype=np.complex128)
    R.hmlistln(cr, kp, H0)
    R.hmlistln(cr, kp, H1, meffig=cr.cfunig[:cr.ngvec].copy())
    assert np.array_equal(H0, H1)


def test_iora_overlap_terms():
    """The oracle's 'iora*' overlap (overlapsetup.f90:81-94, 126, 144; olpistln.f90:53-62),
    pinned term by term against forms written independently of it:
    * all terms zero: the 'none' O bit for bit;
    * h1aa with only (io1, io2) = (0, 1) set to e: O gains e sum_{a,lm} conj(A(i, io 1, lm, a))
      A(j, io 0, lm, a) — an einsum over apwalm (a transposed h1aa would pair io 0 with io 1
      the other way round);
    * h1aa = (c - 1) I: the muffin-tin APW part times c;
    * h1loa = -oalo, h1lolo = -ololo: the APW-lo and lo-lo blocks vanish;
    * m2effig alone: O - O_none = a2 times the kinetic part hmlistln forms with meffig = m2effig
      and V = 0, with a2 = alpha^2 / 2, alpha = 1/137.03599911."""
    import copy
    from conftest import iora_terms
    w = small("b")
    cr, kp = w.crystal, w.kpoints[0]
    n, nmat = kp.ngp, kp.ngp + cr.nlotot
    t = iora_terms(cr, 5)
    _, O0, A = R.setup(cr, kp)
    zero = {k: np.zeros_like(v) for k, v in t.items()}
    _, Oz, _ = R.setup(cr, kp, A, iora=zero)
    assert np.array_equal(Oz, O0)
    # h1aa(0, 1, l, a) = e
    e = 1e-3
    h = np.zeros_like(t["h1aa"])
    h[0, 1] = e
    _, O1, _ = R.setup(cr, kp, A, iora=dict(h1aa=h))
    nl2 = (cr.lmaxmat + 1) ** 2
    ref = e * np.einsum("ila,jla->ij", np.conj(A[:n, 1, :nl2, :]), A[:n, 0, :nl2, :])
    dO = (O1 - O0)[:n, :n]
    iu = np.triu_indices(n)
    assert np.abs(dO[iu] - ref[iu]).max() <= 1e-14 * np.abs(O0).max()
    # h1aa = (c - 1) I
    c = 1.25
    h = np.zeros_like(t["h1aa"])
    h[0, 0] = h[1, 1] = c - 1.0
    _, Oc, _ = R.setup(cr, kp, A, iora=dict(h1aa=h))
    mt = np.zeros((n, n), dtype=complex)
    mt += sum(R._apwi(cr, A, n, a).conj().T @ R._apwi(cr, A, n, a) for a in range(cr.natm))
    d = (Oc - O0)[:n, :n]
    assert np.abs(d[iu] - (c - 1.0) * mt[iu]).max() <= 1e-13 * np.abs(mt).max()
    # h1loa = -oalo, h1lolo = -ololo
    _, Ol, _ = R.setup(cr, kp, A, iora=dict(h1loa=-cr.oalo, h1lolo=-cr.ololo))
    assert np.abs(Ol[:n, n:]).max() == 0.0 and np.abs(Ol[n:, n:]).max() == 0.0
    assert np.array_equal(Ol[:n, :n], O0[:n, :n])
    assert np.abs(O0[:n, n:]).max() > 1e-2 and np.abs(O0[n:, n:]).max() > 0.1
    # m2effig
    m = t["m2effig"]
    _, Om, _ = R.setup(cr, kp, A, iora=dict(m2effig=m))
    c2 = copy.copy(cr)
    c2.veffig = np.zeros_like(cr.veffig)
    K = np.zeros((n, n), dtype=np.complex128)
    R.hmlistln(c2, kp, K, meffig=m)
    a2 = 0.5 / 137.03599911 ** 2
    d = (Om - O0)[:n, :n]
    # the difference of two O of order 1: rounding of O itself, far below the term (~1e-4)
    assert np.abs(d[iu] - a2 * K[iu]).max() <= 1e-15 * np.abs(O0).max()
    assert np.abs(K).max() > 1.0


def test_step_function_real_space():
    """Theta(G) = (1/Omega) int_cell Theta(r) exp(-i G.r) dr, by a grid sum over the cell."""
    w = small("a")
    cr = w.crystal
    ng = 72
    f = (np.arange(ng) + 0.5) / ng
    F = np.stack(np.meshgrid(f, f, f, indexing="ij"), axis=0).reshape(3, -1)
    r = cr.avec @ F
    theta = np.ones(r.shape[1])
    for a in range(cr.natm):
        d = F - np.linalg.solve(cr.avec, cr.atposc[:, a])[:, None]
        d -= np.round(d)
        dist = np.linalg.norm(cr.avec @ d, axis=0)
        theta[dist < cr.rmt[cr.spidx[a]]] = 0.0
    for g in range(12):
        val = np.mean(theta * np.exp(-1j * (cr.vgc[:, g] @ r)))
        assert abs(val - cr.cfunig[g]) < 4e-3, (g, val, cr.cfunig[g])
    # the phase convention matters: the conjugate would fail for some G
    errs = [abs(np.mean(theta * np.exp(+1j * (cr.vgc[:, g] @ r))) - cr.cfunig[g]) for g in range(1, 12)]
    assert max(errs) > 1e-2


def test_eigen_matches_eigh():
    w = small("d")
    H, O, _ = R.setup(w.crystal, w.kpoints[0])
    e, v = R.eigen(H, O)
    e2 = scipy.linalg.eigh(H, O, eigvals_only=True)
    np.testing.assert_allclose(e, e2, rtol=1e-12, atol=1e-10)
    # eigenvectors are O-orthonormal
    np.testing.assert_allclose(v.conj().T @ O @ v, np.eye(len(e)), atol=1e-9)


def test_config_sizes():
    """The workloads have the stated shapes (N ranges, nmat = N + natm*nlo)."""
    w = workload(0)
    assert w.crystal.natm == 2 and w.crystal.lmaxapw == 8 and w.crystal.nlotot == 16
    assert 80 <= w.kpoints[0].ngp <= 120
    w1 = workload(1)
    assert w1.crystal.natm == 8 and w1.crystal.nlotot == 32 and len(w1.kpoints) == 16
    assert all(1150 <= k.ngp <= 1250 for k in w1.kpoints)
    assert len({k.ngp for k in w1.kpoints}) > 4


def _radial_factors(cr, kp):
    """rf(atom, G, l, io) = (4 pi/sqrt(Omega)) sqrt((2l+1)/(4 pi)) r_{io,l}(|G+k|), r = D^-1
    [j_l(|G+k| R), |G+k| j_l'(|G+k| R)]: the real radial part of apwalm (match.f90:83-187)."""
    from oracle import specfun
    ngp, nl = kp.ngp, cr.lmaxmat + 1
    fpso = 4 * math.pi / math.sqrt(cr.omega)
    g = kp.gpc[:ngp]
    rf = np.zeros((cr.natm, ngp, nl, cr.apwordmax))
    for a in range(cr.natm):
        x = g * cr.rmt[cr.spidx[a]]
        j = specfun.sbessel_array(cr.lmaxapw, x)
        dj = specfun.sbesseld1_array(cr.lmaxapw, x) * g[None, :]
        for l in range(nl):
            r = np.linalg.solve(cr.dmat[:, :, l, a], np.stack([j[l], dj[l]]))
            rf[a, :, l, :] = fpso * math.sqrt((2 * l + 1) / (4 * math.pi)) * r.T
    return rf


@pytest.mark.parametrize("key", ["a", "d"])
def test_overlap_addition_theorem(key):
    """The identity the GPU's EXC_BATCH_OLP_LEGENDRE overlap rests on: with apwalm = i^l r S
    conj(Y_lm), the reference's muffin-tin overlap conj(A) A^T equals
    sum_a conj(S_a(G)) S_a(G') sum_l P_l(u.u') sum_io rf rf' (addition theorem), here with
    P_l from scipy rather than the kernel's Clenshaw recurrence.  A dropped (2l+1) weight,
    a missing conjugate or a wrong l range fails it by O(1)."""
    w = small(key)
    cr = w.crystal
    for kp in w.kpoints:
        n = kp.ngp
        A = R.match(cr, kp)
        apwi = [R._apwi(cr, A, n, a) for a in range(cr.natm)]
        O_mt = R._mt_sum(cr, apwi, apwi)
        rf = _radial_factors(cr, kp)
        th, ph = kp.tpgpc[0, :n], kp.tpgpc[1, :n]
        u = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
        x = np.clip(u @ u.T, -1.0, 1.0)
        O = np.zeros((n, n), complex)
        for a in range(cr.natm):
            t = sum(sp.eval_legendre(l, x) * (rf[a, :, l, :] @ rf[a, :, l, :].T) for l in range(cr.lmaxmat + 1))
            S = kp.sfacgp[:n, a]
            O += np.conj(S)[:, None] * S[None, :] * t
        err = np.linalg.norm(O - O_mt) / np.linalg.norm(O_mt)
        assert err < 1e-13, err


@pytest.mark.parametrize("key", ["b", "d"])
def test_per_atom_order_equals_l_blocks(key):
    """The CPU baseline's per-atom zgemm order (the reference's, hamiltonsetup.f90:90-103) and
    the checker's l-block order sum the same terms: H and O agree to rounding."""
    w = small(key)
    cr, kp = w.crystal, w.kpoints[0]
    A = R.match(cr, kp)
    H0, O0, _ = R.setup(cr, kp, apwalm=A)
    H1, O1, _ = R.setup(cr, kp, apwalm=A, per_atom=True)
    assert np.linalg.norm(H1 - H0) <= 1e-13 * np.linalg.norm(H0)
    assert np.linalg.norm(O1 - O0) <= 1e-13 * np.linalg.norm(O0)
    assert not np.array_equal(H1, H0) or cr.natm == 1 or cr.lmaxmat == 0   # the two orders are different code
