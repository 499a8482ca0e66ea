"""Pins the oracle's muffin-tin density matrix (oracle.lapw_ref.gendmatmt, gendmatmt.f90:59-172).

* Row order: a known-answer list written from gendmatmt.f90:93-98 (Do l / Do io / Do m,
  m fastest) against apwalm's storage (io, lm), lm = l(l+1)+m.
* Contraction, conjugation, lo rows, weights: for any coefficient vectors C_j (not only
  eigenvectors), the density matrix reproduces the occupied-state expectation values of the
  setup's muffin-tin Hamiltonian, computed from the oracle's H minus its interstitial part
  (hamiltonsetup.f90 / hmlistln.f90):
      w sum_j occ_j C_j^H (H - H_int) C_j
        = sum_a [ tr(h_a dm_aa) + 2 Re tr(loa_a dm_a,lo) + tr(lolo_a dm_lo,lo) ],
  and the APW trace reproduces the muffin-tin overlap, w sum_j occ_j c_j^H (O - O_int) c_j
  = sum_a tr(dm_aa) for vectors c_j without lo components.
"""
import numpy as np

from conftest import small
from oracle import lapw_ref as R


def test_dmat_row_order_known_answer():
    w = small("a")
    cr = w.crystal
    assert cr.apwordmax == 2
    o = R.dmat_order(cr)
    # l = 0: (io 0), (io 1); l = 1: io 0 m -1..1 -> lm 1..3 -> storage 2, 4, 6; io 1 -> 3, 5, 7
    assert list(o[:8]) == [0, 1, 2, 4, 6, 3, 5, 7]
    # l = 2 starts at f = 8: io 0, m = -2..2 -> lm 4..8 -> storage 8, 10, ..., 16
    assert list(o[8:13]) == [8, 10, 12, 14, 16]
    assert len(o) == cr.apwordmax * (cr.lmaxmat + 1) ** 2
    assert sorted(o) == list(range(len(o)))


def _interstitial(cr, kp, nmat):
    Hi = np.zeros((nmat, nmat), dtype=np.complex128)
    Oi = np.zeros((nmat, nmat), dtype=np.complex128)
    R.hmlistln(cr, kp, Hi)
    R.olpistln(cr, kp, Oi)
    return R.copy_ul(Hi), R.copy_ul(Oi)


def test_dmat_reproduces_muffin_tin_expectation_values():
    w = small("d")
    cr, kp = w.crystal, w.kpoints[0]
    ngp = kp.ngp
    apw = R.match(cr, kp)
    H, O, _ = R.setup(cr, kp, apw)
    nmat = H.shape[0]
    Hi, Oi = _interstitial(cr, kp, nmat)
    rng = np.random.default_rng(5)
    nst, wk = 9, 0.37
    C = rng.standard_normal((nmat, nst)) + 1j * rng.standard_normal((nmat, nst))
    occ = rng.uniform(0.0, 2.0, nst)
    dm = R.gendmatmt(cr, apw, ngp, C, occ, wk)
    ws = R.dmat_wfsize(cr)
    assert dm.shape == (ws, ws, cr.natm)
    for a in range(cr.natm):
        assert np.allclose(dm[:, :, a], dm[:, :, a].conj().T, rtol=0, atol=1e-12 * np.abs(dm).max())

    order = R.dmat_order(cr)
    naa = len(order)
    maxaa = cr.apwordmax * cr.lmmaxapw
    e_dm = 0.0
    for a in range(cr.natm):
        is_ = cr.spidx[a]
        h = cr.haa[:, :, a][np.ix_(order, order)]          # haa rows are storage-ordered
        e_dm += np.trace(h @ dm[:naa, :naa, a]).real
        ns = cr.losize[is_] if cr.nlorb[is_] else 0
        if ns:
            loa = cr.loa[:ns, :, a][:, order]
            e_dm += 2.0 * np.trace(loa @ dm[:naa, maxaa:maxaa + ns, a]).real
            e_dm += np.trace(cr.lolo[:ns, :ns, a] @ dm[maxaa:maxaa + ns, maxaa:maxaa + ns, a]).real
    Hmt = H - Hi
    e_ref = wk * sum(occ[j] * (C[:, j].conj() @ Hmt @ C[:, j]).real for j in range(nst))
    assert abs(e_dm - e_ref) <= 1e-11 * abs(e_ref), (e_dm, e_ref)

    c = C.copy()
    c[ngp:] = 0.0
    dm0 = R.gendmatmt(cr, apw, ngp, c, occ, wk)
    n_dm = sum(np.trace(dm0[:naa, :naa, a]).real for a in range(cr.natm))
    Omt = O - Oi
    n_ref = wk * sum(occ[j] * (c[:, j].conj() @ Omt @ c[:, j]).real for j in range(nst))
    assert abs(n_dm - n_ref) <= 1e-11 * abs(n_ref), (n_dm, n_ref)


def test_dmat_second_variational_factor():
    """evecsv multiplies wf1 from the right (gendmatmt.f90:134-147): with a unitary evecsv
    and equal occupations the density matrix is unchanged."""
    w = small("a")
    cr, kp = w.crystal, w.kpoints[0]
    apw = R.match(cr, kp)
    nmat = kp.ngp + cr.nlotot
    rng = np.random.default_rng(9)
    nst = 6
    C = rng.standard_normal((nmat, nst)) + 1j * rng.standard_normal((nmat, nst))
    U, _ = np.linalg.qr(rng.standard_normal((nst, nst)) + 1j * rng.standard_normal((nst, nst)))
    occ = np.full(nst, 1.5)
    a = R.gendmatmt(cr, apw, kp.ngp, C, occ, 0.5)
    b = R.gendmatmt(cr, apw, kp.ngp, C, occ, 0.5, evecsv=U)
    assert np.allclose(a, b, rtol=0, atol=1e-12 * np.abs(a).max())
    occ2 = rng.uniform(0, 2, nst)
    c = R.gendmatmt(cr, apw, kp.ngp, C, occ2, 0.5, evecsv=U)
    d = R.gendmatmt(cr, apw, kp.ngp, C @ U, occ2, 0.5)
    assert np.allclose(c, d, rtol=0, atol=1e-12 * np.abs(c).max())
