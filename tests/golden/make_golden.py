"""Writes tests/golden/specfun_ref.npz and hmlint_ref.npz from the reference's own Fortran routines.

The reference's src/src_lib/{sbessel,sbesseldm,factnm,factr,genylm,sphcrd}.f90 are compiled
(where they lie in the reference checkout) by oracle/Makefile into oracle/_ref/libexcref.so;
this script evaluates j_l, j_l' (sbesseldm with m = 0, 1) and Y_lm (genylm) on fixed inputs
and stores inputs and outputs; likewise the spline integral (fderiv, m = -1), Gaunt
coefficients (gaunt), real spherical harmonics (genrlm) and the four-point derivatives
(polynom) on the inputs of test_oracle_hmlint.fixture_inputs / polynom_inputs.  Run from the repository root:

    make -C oracle && python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from test_oracle_specfun import XS, _reflib, ref_genylm, ref_sbesseldm  # noqa: E402
import test_oracle_hmlint as H  # noqa: E402


def main():
    lib = _reflib()
    if lib is None:
        raise SystemExit("oracle/_ref/libexcref.so missing: run `make -C oracle` first")
    lmax = 10
    xs = XS[XS <= 1e5]
    jl = np.array([ref_sbesseldm(lib, 0, lmax, x) for x in xs])
    djl = np.array([ref_sbesseldm(lib, 1, lmax, x) for x in xs])
    rng = np.random.default_rng(5)
    tp = np.stack([rng.uniform(0, math.pi, 50), rng.uniform(0, 2 * math.pi, 50)], axis=1)
    ylm = np.array([ref_genylm(lib, lmax, t, p) for t, p in tp])
    np.savez_compressed(os.path.join(HERE, "specfun_ref.npz"), x=xs, tp=tp, lmax=lmax, jl=jl, djl=djl, ylm=ylm)
    print("wrote", os.path.join(HERE, "specfun_ref.npz"))
    hl = H._reflib()
    meshes, funcs, args, tps = H.fixture_inputs()
    ints = np.array([H.ref_fderiv_int(hl, x, f) for x, f in zip(meshes, funcs)])
    gaunts = np.array([H.ref_gaunt(hl, *map(int, a)) for a in args])
    rlms = np.array([H.ref_genrlm(hl, 8, t, p) for t, p in tps])
    polys = np.array([[H.ref_polynom(hl, m, xa, ya, x) for m in range(4)] for xa, ya, x in H.polynom_inputs()])
    np.savez_compressed(os.path.join(HERE, "hmlint_ref.npz"), ints=ints, gaunt=gaunts, rlm=rlms, polynom=polys)
    print("wrote", os.path.join(HERE, "hmlint_ref.npz"))


if __name__ == "__main__":
    main()
