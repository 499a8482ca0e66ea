"""Pins the split-BF16 emulation (oracle/split_ref.py).

* slices are exactly representable in BF16 (at most 8 significant bits, checked bitwise),
* the three slices reconstruct every value to within 2^(e-8) * 2^-17 (e = block exponent),
* with all nine slice pairs the product equals the FP64 product to the slicing error
  (about 2^-24 of the block maxima), and the six pairs kept by the kernel add at most a
  few times that (the dropped pairs are of order 2^-24 themselves),
* the leading-pair partial sums of one atom block stay below 2^25 (FP32-exact integer range
  is 2^24; the kernel's first accumulation level relies on it being nearly exact).
"""
import numpy as np

from conftest import small
from oracle import lapw_ref as R
from oracle import split_ref as S


def _bf16_exact(v):
    f = np.asarray(v, dtype=np.float32)
    assert np.array_equal(f.astype(np.float64), v)
    bits = f.view(np.uint32)
    assert np.all((bits & 0xFFFF) == 0)


def _mats(key="d"):
    w = small(key)
    cr, kp = w.crystal, w.kpoints[0]
    A = R.match(cr, kp)
    n = kp.ngp
    Am = A[:n].reshape(n, -1, order="F")
    Yt = np.concatenate([A[:n, :, :, a].reshape(n, -1, order="F") @ cr.haa[:, :, a].T for a in range(cr.natm)], axis=1)
    return cr, Am, Yt


def test_slices_exact_and_reconstruct():
    cr, Am, _ = _mats()
    sl, e = S.split(Am, cr.natm, cr.maxaa)
    for s in range(3):
        _bf16_exact(sl[s].real)
        _bf16_exact(sl[s].imag)
    assert np.abs(sl[0].real).max() <= 256 and np.abs(sl[1].real).max() <= 0.5 + 1e-12
    rec = (sl[0] + sl[1] + sl[2]).reshape(Am.shape[0], cr.natm, cr.maxaa) * np.ldexp(1.0, e - 8)[:, :, None]
    err = np.abs(rec - Am.reshape(rec.shape))
    bound = np.ldexp(1.0, e - 8 - 17)[:, :, None] * np.sqrt(2)
    assert np.all(err <= bound)


def test_product_error_bounds():
    cr, Am, Yt = _mats()
    natm, naa = cr.natm, cr.maxaa
    exact = Am.conj() @ Yt.T
    full = S.product(Am, Yt, natm, naa, pairs=[(s, t) for s in range(3) for t in range(3)])
    kept = S.product(Am, Yt, natm, naa)
    nrm = np.linalg.norm(exact)
    e_full = np.linalg.norm(full - exact) / nrm
    e_kept = np.linalg.norm(kept - exact) / nrm
    assert e_full < 2e-7
    assert e_kept < 4 * e_full
    # dropping (0,1),(1,0) as well would be visibly worse: the kept pairs matter
    worse = S.product(Am, Yt, natm, naa, pairs=[(0, 0)])
    assert np.linalg.norm(worse - exact) / nrm > 10 * np.linalg.norm(kept - exact) / nrm


def test_leading_pair_sums_in_fp32_integer_range():
    cr, Am, Yt = _mats()
    sa, _ = S.split(Am, cr.natm, cr.maxaa)
    sb, _ = S.split(Yt, cr.natm, cr.maxaa)
    naa = cr.maxaa
    worst = 0.0
    for a in range(cr.natm):
        c = slice(a * naa, (a + 1) * naa)
        re = sa[0][:, c].real @ sb[0][:, c].real.T + sa[0][:, c].imag @ sb[0][:, c].imag.T
        worst = max(worst, np.abs(re).max())
    assert worst < 2.0 ** 25
