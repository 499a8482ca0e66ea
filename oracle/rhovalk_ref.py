"""CPU restatement of the muffin-tin density of the direct route, rhovalk (test infrastructure:
only tests/, __graft_entry__.smoke and bench.py's cpu_baseline leg use it; nothing in the
library calls it).

The spinless part (spin unassociated; first-variational states, or with evecsv the states of
tevecsv true with one spinor):

* sphcover (src_lib/sphcover.f90:49-60): n points on the unit sphere, (theta, phi) per point;
* genshtmat (genshtmat.f90:78-108): zbshtvr(itp, lm) = Y_lm, rbshtvr(itp, lm) = R_lm at the
  lmmaxvr points of sphcover(lmmaxvr), rfshtvr = rbshtvr^-1;
* wavefmt (wavefmt.f90:86-124): the muffin-tin wavefunction of one state in complex
  harmonics on every lradstep-th mesh point,
      wfmt(lm, irc) = sum_io (sum_G evec(G) apwalm(G, io, lm)) apwfr(ir, io, l)
                    + sum_ilo evec(ngp + idxlo(lm, ilo)) lofr(ir, ilo)      (l <= lmaxvr);
* rhovalk (rhovalk.f90:99-207): per atom, for every state with |wkpt occ| > epsocc,
  rfmt(itp, irc) += wkpt occ |(zbshtvr wfmt)(itp, irc)|^2 (with evecsv, tevecsv true and one
  spinor: wfmt of state j = sum over first-variational states ist with
  |Re evecsv(ist, j)| + |Im evecsv(ist, j)| > epsocc of evecsv(ist, j) wfmt_ist,
  rhovalk.f90:108-136), then
  rhomt(:, ir, a) += rfshtvr rfmt(:, irc) at ir = 1, 1 + lradstep, ... (the other points of
  the mesh are not touched).

wavefmt_add's skip of coefficients below 1e-14 (wavefmt.f90:126-152) leaves out terms of at
most 1e-14 |u|; the restatement adds them.

Pinned in tests/test_oracle_rhovalk.py: sphcover against the reference's own routine compiled
by oracle/Makefile and its committed vectors; the transform pair (rbshtvr rfshtvr = 1; zbshtvr
rows are the oracle's Y_lm); and, on band-limited wavefunctions (radial functions zero above
l = 2, so |psi|^2 has l <= 4 <= lmaxvr and the point transform is exact), equality with the
density-matrix route gendmatmt + genrhomt, pinned on its own.
"""
import math

import numpy as np

from oracle import hmlint_ref as M
from oracle import specfun


def sphcover(n):
    """tp(2, n): (theta, phi) of the reference's spiral covering (sphcover.f90:49-60)."""
    if n <= 0:
        raise ValueError("sphcover: n <= 0")
    tp = np.zeros((2, n))
    c = math.sqrt(8.0 * math.pi / math.sqrt(3.0))
    t1 = c / math.sqrt(float(n))
    tp[0, 0] = math.pi
    for k in range(1, n - 1):                       # Fortran k = 2 .. n-1
        h = 2.0 * k / (n - 1) - 1.0
        tp[0, k] = math.acos(h)
        tp[1, k] = math.fmod(tp[1, k - 1] + t1 / math.sqrt(1.0 - h * h), 2.0 * math.pi)
    tp[0, n - 1] = 0.0
    tp[1, n - 1] = 0.0
    return tp


def genshtmat(lmaxvr):
    """(zbshtvr, rbshtvr, rfshtvr), each (lmmaxvr, lmmaxvr) with rows the points itp
    (genshtmat.f90:78-108; the inverse by LU as dgetrf/dgetri)."""
    lmv = (lmaxvr + 1) ** 2
    tp = sphcover(lmv)
    zb = specfun.genylm(lmaxvr, tp[0], tp[1]).T.copy()           # (itp, lm)
    rb = M.rlm(lmaxvr, tp[0], tp[1])                               # (itp, lm)
    rf = np.linalg.inv(rb)
    return np.asfortranarray(zb), np.asfortranarray(rb), np.asfortranarray(rf)


def coarse_points(nrmt, lradstep):
    """0-based mesh indices ir = 0, lradstep, ... < nrmt (rhovalk.f90:183, wavefmt.f90:86)."""
    return np.arange(0, int(nrmt), int(lradstep))


def wavefmt(cr, rd, apwalm, ngp, evec, ias, lmaxvr, lradstep):
    """wfmt(lmmaxvr, nrc) of one state (wavefmt.f90:86-124); evec is the state's column."""
    if lmaxvr > cr.lmaxapw:
        raise ValueError("wavefmt: lmax > lmaxapw")
    is_ = cr.spidx[ias]
    irs = coarse_points(rd.nrmt[is_], lradstep)
    lmv = (lmaxvr + 1) ** 2
    wf = np.zeros((lmv, len(irs)), dtype=np.complex128)
    for l in range(lmaxvr + 1):
        for m in range(-l, l + 1):
            lm = l * l + l + m
            for io in range(cr.apwordmax):
                z = np.dot(evec[:ngp], apwalm[:ngp, io, lm, ias])      # zdotu: no conjugate
                wf[lm] += z * rd.apwfr[irs, 0, io, l, ias]
    for ilo in range(cr.nlorb[is_]):
        l = int(cr.lorbl[ilo, is_])
        if l > lmaxvr:
            continue
        for m in range(-l, l + 1):
            lm = l * l + l + m
            i = ngp + cr.idxlo[lm, ilo, ias] - 1
            wf[lm] += evec[i] * rd.lofr[irs, 0, ilo, ias]
    return wf


def rhovalk(cr, rd, apwalm, ngp, evec, occ, wkpt, lradstep=1, epsocc=1e-8, sht=None, rhomt=None, evecsv=None):
    """rhomt(lmmaxvr, nrmtmax, natm) += the k-point's muffin-tin density (rhovalk.f90:99-207,
    spinless).  evec: (>= nmat, nst); occ: (nst,); evecsv: (nst, nst) or None for the
    identity; sht: genshtmat's (zbshtvr, rbshtvr, rfshtvr) or None to make them."""
    lmaxvr = rd.lmaxvr
    lmv = (lmaxvr + 1) ** 2
    zb, _, rf = sht if sht is not None else genshtmat(lmaxvr)
    if rhomt is None:
        rhomt = np.zeros((lmv, rd.spr.shape[0], cr.natm), order="F")
    for ias in range(cr.natm):
        is_ = cr.spidx[ias]
        irs = coarse_points(rd.nrmt[is_], lradstep)
        rfmt = np.zeros((lmv, len(irs)))
        for j in range(evec.shape[1]):
            t1 = wkpt * occ[j]
            if abs(t1) <= epsocc:
                continue
            if evecsv is None:
                w3 = zb @ wavefmt(cr, rd, apwalm, ngp, evec[:, j], ias, lmaxvr, lradstep)
            else:
                w3 = 0
                for ist in range(evec.shape[1]):
                    z = evecsv[ist, j]
                    if abs(z.real) + abs(z.imag) > epsocc:
                        w3 = w3 + z * (zb @ wavefmt(cr, rd, apwalm, ngp, evec[:, ist], ias, lmaxvr, lradstep))
            rfmt += t1 * (w3.real ** 2 + w3.imag ** 2)
        rhomt[:, irs, ias] += rf @ rfmt
    return rhomt
