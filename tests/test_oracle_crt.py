"""Pins the modular emulation of the split contraction (oracle/crt_ref.py) by routes independent
of it.

* moduli: pairwise coprime, odd, -1 a square modulo each (iota^2 = -1), and the split of
  Z_m[i] a ring map: the maps of a product are the products of the maps;
* residues: against Python's own modulo of the same integers, in the symmetric ranges;
* reconstruction: Garner's digits give back random integers anywhere in [-(P-1)/2, (P-1)/2],
  for every modulus count, and the float conversion equals float() of the exact integer
  (correct rounding, ties to even included);
* products: Re', Im' reconstructed from the residue products equal the exact Python-integer
  products of the same integer matrices (conjugated and plain), with row norms at the 2^beta
  bound; a product past the bound (one modulus too few) is detected as different;
* scaling: every scaled row's 2-norm stays below 2^beta (+ rounding), so |Re'|, |Im'| < P/2;
* accuracy: the emulated conj(A) B^T against the FP64 product for random operands with the
  spread of magnitudes of apwalm's columns, within the scheme's bound: a rounded row is within
  sqrt(2K)/2 units of the scaled one and a unit is below 2^(1 - beta) |X_i|, so the error is
  at most 2^(1 - beta) sqrt(2K) |A_i| |B_j| (and in practice far below it).
"""
import math

import numpy as np
import pytest

from oracle import crt_ref as Q


def test_moduli_coprime_and_beta():
    for n in range(2, len(Q.MODULI) + 1):
        mods = Q.moduli(n)
        for i in range(n):
            for j in range(i):
                assert math.gcd(mods[i], mods[j]) == 1
        P = math.prod(mods)
        b = Q.beta(n)
        assert (1 << (2 * b + 2)) <= P < (1 << (2 * b + 4))
    assert math.prod(Q.MODULI[:Q.LO_DIGITS]) < 2 ** 53
    for m in Q.MODULI:
        assert m % 2 == 1 and m <= 255
        t = Q.iota(m)
        assert (t * t + 1) % m == 0


def test_gaussian_split_is_a_ring_map():
    """(x + iy) -> x + iota y modulo m maps sums to sums and products to products, for both
    square roots iota and -iota: the basis of the two-product form."""
    rng = np.random.default_rng(2)
    for m in Q.MODULI:
        t = Q.iota(m)
        for s in (t, m - t):
            for _ in range(50):
                a, b, c, d = (int(v) for v in rng.integers(-10 ** 6, 10 ** 6, 4))
                re, im = a * c - b * d, a * d + b * c          # (a + ib)(c + id)
                assert (re + s * im - (a + s * b) * (c + s * d)) % m == 0


def test_residues_symmetric():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.integers(-2 ** 46, 2 ** 46, 5000), np.arange(-600, 600)]).astype(np.float64)
    for m in Q.MODULI:
        r = Q.residues(x, m)
        lo, hi = -(m - 1) // 2, (m - 1) // 2
        assert r.min() >= lo and r.max() <= hi
        ref = np.array([((int(v) - lo) % m) + lo for v in x], dtype=np.float64)
        assert np.array_equal(r, ref)


@pytest.mark.parametrize("n", [6, 9, 11, 12, 14])
def test_reconstruct_round_trip(n):
    mods = Q.moduli(n)
    P = math.prod(mods)
    rng = np.random.default_rng(n)
    xs = [int(v) for v in rng.integers(-2 ** 62, 2 ** 62, 300)]
    xs += [(int(a) << 40) + int(b) for a, b in zip(rng.integers(-2 ** 40, 2 ** 40, 300), rng.integers(0, 2 ** 40, 300))]
    xs = [x % P - (P - 1) // 2 for x in xs] + [-(P - 1) // 2, (P - 1) // 2, 0, 1, -1, P // 3]
    res = [np.array([x % m for x in xs], dtype=np.int64) for m in mods]
    hi, lo, Mk = Q.reconstruct(res, mods)
    got = [int(h) * Mk + int(l) for h, l in zip(hi, lo)]
    assert got == xs
    f = Q.to_float(hi, lo, Mk)
    assert np.array_equal(f, np.array([float(x) for x in xs]))


def test_to_float_rounds_ties_to_even():
    Mk = math.prod(Q.MODULI[:Q.LO_DIGITS])
    # hi * Mk + lo = 2^60 + 2^7 (a tie at 53 bits: rounds down to even) and 2^60 + 3 * 2^7 (up)
    for x in (2 ** 60 + 2 ** 7, 2 ** 60 + 3 * 2 ** 7, -(2 ** 61) - 2 ** 8, 2 ** 70 + 1):
        hi, lo = divmod(x, Mk)
        f = Q.to_float(np.array([hi]), np.array([lo]), Mk)[0]
        assert f == float(x)


@pytest.mark.parametrize("conj", [True, False])
@pytest.mark.parametrize("n", [9, 11])
def test_residue_products_equal_exact_integers(conj, n):
    rng = np.random.default_rng(3 + n)
    K = 37
    b = Q.beta(n)
    A = rng.standard_normal((13, K)) + 1j * rng.standard_normal((13, K))
    B = rng.standard_normal((11, K)) + 1j * rng.standard_normal((11, K))
    A[3] *= 1e-9                               # rows of very different size
    B[2, :30] = 0.0
    sa, sb = Q.row_scale(A, n), Q.row_scale(B, n)
    ar, ai = Q.integers(A, sa)
    br, bi = Q.integers(B, sb)
    assert (np.sqrt(ar ** 2 + ai ** 2).sum(1) > 0).all()
    assert np.sqrt((ar ** 2 + ai ** 2).sum(1)).max() < 2.0 ** b + K
    mods = Q.moduli(n)
    rre, rim = Q._int_products(ar, ai, br, bi, mods, conj)
    ere, eim = Q.big_product(ar, ai, br, bi, conj)
    for rs, ex in ((rre, ere), (rim, eim)):
        hi, lo, Mk = Q.reconstruct(rs, mods)
        got = hi.astype(object) * Mk + lo.astype(object)
        assert (got == ex).all()
    # one modulus too few cannot hold the products: the reconstruction differs
    rre2, _ = Q._int_products(ar, ai, br, bi, mods[:-2], conj)
    hi, lo, Mk = Q.reconstruct(rre2, mods[:-2])
    assert not (hi.astype(object) * Mk + lo.astype(object) == ere).all()


def test_product_accuracy_bound():
    rng = np.random.default_rng(7)
    K = 300
    scale = np.exp(-np.linspace(0, 25, K))       # columns spanning ~11 decades, as apwalm's l-blocks
    A = (rng.standard_normal((40, K)) + 1j * rng.standard_normal((40, K))) * scale
    B = (rng.standard_normal((30, K)) + 1j * rng.standard_normal((30, K))) * scale[::-1] ** 0.5
    for n in (9, 11):
        C, sa, sb = Q.product(A, B, n)
        ref = np.conj(A) @ B.T
        na = np.linalg.norm(A, axis=1)
        nb = np.linalg.norm(B, axis=1)
        bound = 2.0 ** (1 - Q.beta(n)) * np.sqrt(2 * K) * np.outer(na, nb)
        assert (np.abs(C - ref) <= bound).all()
        assert (np.abs(C - ref) <= bound / 8).mean() > 0.99        # the bound is a worst case
        P, _, _ = Q.product(A, B, n, conj=False)
        assert (np.abs(P - A @ B.T) <= bound).all()
