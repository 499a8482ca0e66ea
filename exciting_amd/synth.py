"""Deterministic synthetic inputs for the per-k LAPW H+O setup.

This module plays the role of the reference's initialisation producers for the
setup's inputs: the G-vector set and its lookup table (gengvec.f90:56-125), the
G+k basis sets (gengpvec.f90:50-75), the structure factors (gensfacgp.f90:38-49),
the local-orbital index (genidxlo.f90:24-43), the characteristic function
(gencfun.f90:52-91) and the per-SCF muffin-tin data that hmlint/olprad would
produce (seeded here).  Every array that crosses the C-ABI is stored with the
reference's Fortran shape and column-major order, so ``arr.ctypes.data`` can be
handed to the library as the Fortran caller would hand it.

Five workloads are defined (``make_config(cfg)``):

0  cubic a=10.26, 2 atoms / 1 species, R=2.0, lmax 8, 8 lo functions, 1 k,
   |G+k| < 3.5/R (N about 97; the cutoff is followed, not the N~250 estimate)
1  orthorhombic, 8 atoms / 2 species (R 2.0, 1.8), lmax 10, 4 lo, 16 k, N in [1150,1250]
2  16 atoms / 3 species, lmax 10, 4 lo, 64 k, N in [2400,2700]
3  32 atoms / 4 species, lmax 10, 4 lo, 8 k, N about 5000
4  the cell of workload 2, 8 k, N about 2500

Seeds: ``20261202 + cfg`` with numpy's PCG64.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

FOURPI = 4.0 * math.pi

# workload table: cell, species radii, atoms per species, lmax, lo l-list, k count, N range
_CONFIGS = {
    0: dict(cell=(10.26, 10.26, 10.26), rmt=(2.0,), natoms=(2,), lmax=8,
            lorbl=(0, 0, 1, 1), nk=1, nrange=None, gkmax=3.5 / 2.0),
    1: dict(cell=(10.2, 11.8, 13.77), rmt=(2.0, 1.8), natoms=(4, 4), lmax=10,
            lorbl=(0, 1), nk=16, nrange=(1150, 1250), gkmax=None),
    2: dict(cell=(13.5, 15.2, 17.16), rmt=(2.0, 1.8, 2.2), natoms=(6, 5, 5), lmax=10,
            lorbl=(0, 1), nk=64, nrange=(2400, 2700), gkmax=None),
    3: dict(cell=(17.5, 19.0, 20.77), rmt=(2.0, 1.8, 2.2, 1.6), natoms=(8, 8, 8, 8), lmax=10,
            lorbl=(0, 1), nk=8, nrange=(4950, 5050), gkmax=None),
    4: dict(cell=(13.5, 15.2, 17.16), rmt=(2.0, 1.8, 2.2), natoms=(6, 5, 5), lmax=10,
            lorbl=(0, 1), nk=8, nrange=(2475, 2525), gkmax=None),
}

APWORD = 2          # apword(l, is) for every l and species (u and u-dot)


@dataclasses.dataclass
class Crystal:
    """SCF-constant data of the setup (what the reference keeps in module state)."""
    cfg: int
    avec: np.ndarray          # (3,3) lattice vectors as columns, bohr
    bvec: np.ndarray          # (3,3) reciprocal vectors as columns
    omega: float
    nspecies: int
    natoms: np.ndarray        # (nspecies,)
    rmt: np.ndarray           # (nspecies,)
    spidx: np.ndarray         # (natm,) species (0-based) of each atom, species-major order
    atposc: np.ndarray        # (3, natm) Cartesian positions, Fortran order
    lmaxapw: int
    lmaxmat: int
    apwordmax: int
    nlorb: np.ndarray         # (nspecies,)
    lorbl: np.ndarray         # (nlomax, nspecies) Fortran order, l of each lo
    idxlo: np.ndarray         # (lolmmax, nlomax, natm) 1-based column index of each lo function
    losize: np.ndarray        # (nspecies,)
    nlotot: int
    # radial boundary values: D(io1, io2, l, ias) = d^{io1-1} u_{io2,l}/dr^{io1-1} at R
    dmat: np.ndarray          # (apwordmax, apwordmax, lmaxapw+1, natm) real, Fortran order
    # muffin-tin matrices (mt_hscf%main and olprad outputs)
    haa: np.ndarray           # (maxaa, maxaa, natm) complex
    loa: np.ndarray           # (maxnlo, maxaa, natm) complex
    lolo: np.ndarray          # (maxnlo, maxnlo, natm) complex
    oalo: np.ndarray          # (apwordmax, nlomax, natm) real
    ololo: np.ndarray         # (nlomax, nlomax, natm) real
    # G-vector set
    intgv: np.ndarray         # (3,2) integer box bounds
    ivg: np.ndarray           # (3, ngrtot) int32 Fortran order
    ivgig: np.ndarray         # (n1, n2, n3) int32 Fortran order, 1-based, box offset by intgv[:,0]
    vgc: np.ndarray           # (3, ngrtot)
    gc: np.ndarray            # (ngrtot,)
    ngvec: int
    gmaxvr: float
    cfunig: np.ndarray        # (ngrtot,) complex: step function Theta(G)
    veffig: np.ndarray        # (ngvec,) complex: V(G), V(-G) = conj V(G)
    ffacg: np.ndarray         # (ngvec, nspecies) step-function form factors (genffacg.f90)

    @property
    def natm(self) -> int:
        return int(self.atposc.shape[1])

    @property
    def lmmaxapw(self) -> int:
        return (self.lmaxapw + 1) ** 2

    @property
    def maxaa(self) -> int:
        return self.apwordmax * (self.lmaxmat + 1) ** 2

    @property
    def maxnlo(self) -> int:
        return int(self.losize.max()) if self.nlotot else 0

    @property
    def nlomax(self) -> int:
        return int(self.lorbl.shape[0])

    @property
    def lolmmax(self) -> int:
        return int(self.idxlo.shape[0])

    @property
    def ngrtot(self) -> int:
        return int(self.ivg.shape[1])


@dataclasses.dataclass
class KPoint:
    """Per-k basis data (gengpvec/gensfacgp outputs), padded to ngkmax rows."""
    vkl: np.ndarray           # (3,) lattice coordinates
    vkc: np.ndarray           # (3,) Cartesian
    ngp: int
    gkmax: float
    igpig: np.ndarray         # (ngkmax,) int32, 1-based G index
    vgpc: np.ndarray          # (3, ngkmax) Fortran order
    gpc: np.ndarray           # (ngkmax,)
    tpgpc: np.ndarray         # (2, ngkmax) Fortran order
    sfacgp: np.ndarray        # (ngkmax, natm) complex Fortran order

    def nmat(self, cr: Crystal) -> int:
        return self.ngp + cr.nlotot


@dataclasses.dataclass
class Workload:
    cfg: int
    crystal: Crystal
    kpoints: List[KPoint]
    ngkmax: int


def sphcrd(v: np.ndarray):
    """Length and (theta, phi) of the columns of v, following sphcrd.f90:27-49."""
    r = np.sqrt((v * v).sum(axis=0))
    eps = 1e-14
    theta = np.zeros_like(r)
    phi = np.zeros_like(r)
    ok = r > eps
    t1 = np.zeros_like(r)
    t1[ok] = np.clip(v[2, ok] / r[ok], -1.0, 1.0)
    theta[ok] = np.arccos(t1[ok])
    okp = ok & ((np.abs(v[0]) > eps) | (np.abs(v[1]) > eps))
    phi[okp] = np.arctan2(v[1, okp], v[0, okp])
    phi[okp & (phi < 0)] += 2.0 * math.pi
    return r, theta, phi


def _min_image_distance(d: np.ndarray, avec: np.ndarray) -> float:
    """Minimum periodic distance of a fractional displacement d."""
    best = 1e30
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                v = avec @ (d + np.array([i, j, k], dtype=float))
                best = min(best, float(np.sqrt(v @ v)))
    return best


def _positions(rng, avec, rmt, spidx):
    natm = len(spidx)
    frac = []
    tries = 0
    while len(frac) < natm:
        tries += 1
        if tries > 200000:
            raise RuntimeError("could not place atoms")
        cand = rng.random(3)
        a = len(frac)
        ok = True
        for b in range(a):
            dmin = rmt[spidx[a]] + rmt[spidx[b]]
            d = cand - frac[b]
            d -= np.round(d)
            if _min_image_distance(d, avec) < dmin:
                ok = False
                break
        if ok:
            frac.append(cand)
    return np.array(frac).T  # (3, natm) fractional


def nfftifc(n):
    """The next n' >= n whose prime factors are 2, 3 and 5 only (nfftifc.f90)."""
    if n <= 0:
        raise ValueError("nfftifc: n <= 0")
    while True:
        i = n
        for p in (2, 3, 5):
            while i % p == 0:
                i //= p
        if i == 1:
            return n
        n += 1


def gvector_set(avec, bvec, gmaxvr):
    """G vectors of an integer box, sorted by length (stable), following gengvec.f90:56-125, on
    the box of gridsize.f90:31-52: ngrid_i = nfftifc(int(gmaxvr |a_i| / pi) + 1), intgv =
    [ngrid/2 - ngrid + 1, ngrid/2] (the FFT grid; it holds the sphere |G| < gmaxvr).

    Returns intgv (3,2), ivg (3,ngrtot), ivgig (box, 1-based), vgc, gc, ngvec."""
    ngrid = [nfftifc(int(gmaxvr * np.linalg.norm(avec[:, i]) / math.pi) + 1) for i in range(3)]
    intgv = np.array([[g // 2 - g + 1, g // 2] for g in ngrid], dtype=np.int32)
    r1 = np.arange(intgv[0, 0], intgv[0, 1] + 1)
    r2 = np.arange(intgv[1, 0], intgv[1, 1] + 1)
    r3 = np.arange(intgv[2, 0], intgv[2, 1] + 1)
    # i1 outermost, i3 innermost as in the reference's triple loop
    i1, i2, i3 = np.meshgrid(r1, r2, r3, indexing="ij")
    iv = np.stack([i1.ravel(), i2.ravel(), i3.ravel()]).astype(np.int64)
    v = bvec @ iv.astype(float)
    g = np.sqrt((v * v).sum(axis=0))
    order = np.argsort(g, kind="stable")
    ivg = np.asfortranarray(iv[:, order].astype(np.int32))
    gc = g[order]
    vgc = np.asfortranarray(bvec @ ivg.astype(float))
    shape = tuple(int(g) for g in ngrid)
    ivgig = np.zeros(shape, dtype=np.int32, order="F")
    ivgig[ivg[0] - intgv[0, 0], ivg[1] - intgv[1, 0], ivg[2] - intgv[2, 0]] = np.arange(1, ivg.shape[1] + 1,
                                                                                       dtype=np.int32)
    below = np.nonzero(gc < gmaxvr)[0]
    ngvec = int(below[-1] + 1) if below.size else 1
    return intgv, ivg, ivgig, vgc, gc, ngvec


def step_function(omega, rmt, natoms_per_species, spidx, atposc, vgc, gc, epslat=1e-6):
    """Theta(G) = delta_G0 - sum_a exp(-i G.r_a) Theta_s(G), gencfun.f90:52-91 (cfdamp = 0).

    Also returns the per-species form factors Theta_s(G) (genffacg.f90:19-33)."""
    t1 = FOURPI / omega
    nsp = len(rmt)
    ffac = np.empty((gc.shape[0], nsp))
    big = gc > epslat
    for s in range(nsp):
        t2 = gc[big] * rmt[s]
        ffac[big, s] = t1 * (np.sin(t2) - t2 * np.cos(t2)) / gc[big] ** 3
        ffac[~big, s] = (t1 / 3.0) * rmt[s] ** 3
    cf = np.zeros(gc.shape[0], dtype=np.complex128)
    cf[0] = 1.0
    for a in range(atposc.shape[1]):
        ph = -(vgc.T @ atposc[:, a])
        cf -= np.exp(1j * ph) * ffac[:, spidx[a]]
    return cf, ffac


def _hermitian_gauss(rng, n, sigma):
    m = np.zeros((n, n), dtype=np.complex128)
    iu = np.triu_indices(n, 1)
    vals = (rng.standard_normal(len(iu[0])) + 1j * rng.standard_normal(len(iu[0]))) * (sigma / math.sqrt(2.0))
    m[iu] = vals
    m = m + m.conj().T
    return m


def make_crystal(cfg: int, rng) -> Crystal:
    spec = _CONFIGS[cfg]
    a, b, c = spec["cell"]
    avec = np.diag([a, b, c]).astype(float)
    bvec = 2.0 * math.pi * np.linalg.inv(avec).T
    omega = float(abs(np.linalg.det(avec)))
    rmt = np.array(spec["rmt"], dtype=float)
    natoms = np.array(spec["natoms"], dtype=np.int32)
    nsp = len(rmt)
    spidx = np.repeat(np.arange(nsp), natoms).astype(np.int32)
    natm = len(spidx)
    frac = _positions(rng, avec, rmt, spidx)
    atposc = np.asfortranarray(avec @ frac)
    lmax = spec["lmax"]
    lmmax = (lmax + 1) ** 2
    apwordmax = APWORD
    naa = apwordmax * lmmax
    lorbl_list = spec["lorbl"]
    nlomax = len(lorbl_list)
    nlorb = np.full(nsp, nlomax, dtype=np.int32)
    lorbl = np.asfortranarray(np.tile(np.array(lorbl_list, dtype=np.int32)[:, None], (1, nsp)))
    lolmax = max(lorbl_list)
    lolmmax = (lolmax + 1) ** 2
    idxlo = np.zeros((lolmmax, nlomax, natm), dtype=np.int32, order="F")
    i = 0
    for ias in range(natm):                # species-major atom order == idxas
        for ilo in range(nlomax):
            l = lorbl_list[ilo]
            for m in range(-l, l + 1):
                i += 1
                idxlo[l * (l + 1) + m, ilo, ias] = i
    nlotot = i
    losize = np.array([sum(2 * l + 1 for l in lorbl_list)] * nsp, dtype=np.int32)
    maxnlo = int(losize.max())

    # radial boundary values, u(R) and u'(R) in [0.5,1.5]; redraw (l, atom) pairs whose
    # 2x2 matching matrix is badly conditioned (cond > 25) so that A stays O(1)
    dmat = np.zeros((apwordmax, apwordmax, lmax + 1, natm), order="F")
    for ias in range(natm):
        for l in range(lmax + 1):
            while True:
                d = rng.uniform(0.5, 1.5, size=(apwordmax, apwordmax))
                if np.linalg.cond(d) <= 25.0:
                    break
            dmat[:, :, l, ias] = d

    # muffin-tin Hamiltonian h_a: Hermitian, diagonal l(l+1)/(2R^2) + U(-1,1), off-diagonal sigma 0.05
    lof = np.repeat(np.array([l for l in range(lmax + 1) for _ in range(2 * l + 1)]), apwordmax)
    haa = np.zeros((naa, naa, natm), dtype=np.complex128, order="F")
    loa = np.zeros((maxnlo, naa, natm), dtype=np.complex128, order="F")
    lolo = np.zeros((maxnlo, maxnlo, natm), dtype=np.complex128, order="F")
    oalo = np.zeros((apwordmax, nlomax, natm), order="F")
    ololo = np.zeros((nlomax, nlomax, natm), order="F")
    for ias in range(natm):
        R = rmt[spidx[ias]]
        h = _hermitian_gauss(rng, naa, 0.05)
        h[np.diag_indices(naa)] = lof * (lof + 1) / (2.0 * R * R) + rng.uniform(-1.0, 1.0, naa)
        haa[:, :, ias] = h
        ns = losize[spidx[ias]]
        loa[:ns, :, ias] = (rng.standard_normal((ns, naa)) + 1j * rng.standard_normal((ns, naa))) * (0.05 / math.sqrt(2.0))
        ll = _hermitian_gauss(rng, ns, 0.05)
        ll[np.diag_indices(ns)] = np.abs(ll).sum(axis=1) + rng.uniform(0.5, 1.5, ns)
        lolo[:ns, :ns, ias] = ll
        oalo[:, :, ias] = rng.uniform(-0.1, 0.1, size=(apwordmax, nlomax))
        # SPD per l-group of local orbitals, diagonal U[0.5,1.5]
        oo = np.zeros((nlomax, nlomax))
        for l in sorted(set(lorbl_list)):
            grp = [k for k in range(nlomax) if lorbl_list[k] == l]
            g = len(grp)
            dg = rng.uniform(0.5, 1.5, g)
            m = rng.uniform(-0.1, 0.1, size=(g, g))
            blk = np.diag(dg) + 0.5 * (m + m.T) * (1.0 if g > 1 else 0.0)
            blk = blk + np.eye(g) * max(0.0, 0.05 - np.linalg.eigvalsh(blk).min())
            oo[np.ix_(grp, grp)] = blk
        ololo[:, :, ias] = oo

    # G set covering |G| < 2*gkmax (+ margin): the differences G_i - G_j of the basis
    gk_est = spec["gkmax"] if spec["gkmax"] is not None else _gkmax_for_n(spec["nrange"][1] + 50, omega)
    gmaxvr = 2.0 * gk_est + 0.05
    intgv, ivg, ivgig, vgc, gc, ngvec = gvector_set(avec, bvec, gmaxvr)
    cfunig, ffac = step_function(omega, rmt, natoms, spidx, atposc, vgc, gc)
    # V(G): complex Gaussian x exp(-|G|^2/8), V(0) real, V(-G) = conj V(G)
    veff = np.zeros(ngvec, dtype=np.complex128)
    key = {tuple(ivg[:, g]): g for g in range(ngvec)}
    z = (rng.standard_normal(ngvec) + 1j * rng.standard_normal(ngvec)) / math.sqrt(2.0)
    for g in range(ngvec):
        mg = key[tuple(-ivg[:, g])]
        if mg == g:
            veff[g] = z[g].real * math.exp(-gc[g] ** 2 / 8.0)
        elif mg > g:
            veff[g] = z[g] * math.exp(-gc[g] ** 2 / 8.0)
            veff[mg] = np.conj(veff[g])
    return Crystal(cfg=cfg, avec=avec, bvec=bvec, omega=omega, nspecies=nsp, natoms=natoms, rmt=rmt,
                   spidx=spidx, atposc=atposc, lmaxapw=lmax, lmaxmat=lmax, apwordmax=apwordmax,
                   nlorb=nlorb, lorbl=lorbl, idxlo=idxlo, losize=losize, nlotot=nlotot, dmat=dmat,
                   haa=haa, loa=loa, lolo=lolo, oalo=oalo, ololo=ololo, intgv=intgv, ivg=ivg,
                   ivgig=ivgig, vgc=vgc, gc=gc, ngvec=ngvec, gmaxvr=gmaxvr,
                   cfunig=cfunig, veffig=veff, ffacg=np.asfortranarray(ffac[:ngvec]))


def _gkmax_for_n(n, omega):
    return (6.0 * math.pi ** 2 * n / omega) ** (1.0 / 3.0)


def structure_factors(vgpc: np.ndarray, atposc: np.ndarray) -> np.ndarray:
    """S_a(G+k) = exp(i (G+k).r_a), gensfacgp.f90:38-49."""
    return np.exp(1j * (vgpc.T @ atposc))


def make_kpoint(cr: Crystal, rng, ntarget: Optional[int], gkmax: Optional[float], ngkmax: int = 0) -> KPoint:
    f = rng.uniform(-0.5, 0.5, 3)
    vkc = cr.bvec @ f
    v = cr.vgc[:, :cr.ngvec] + vkc[:, None]
    gk = np.sqrt((v * v).sum(axis=0))
    if gkmax is None:
        srt = np.sort(gk)
        n = ntarget
        while srt[n] - srt[n - 1] < 1e-9:      # no ties at the basis edge
            n += 1
        gkmax = 0.5 * (srt[n - 1] + srt[n])
    sel = np.nonzero(gk < gkmax)[0]             # in G order, as gengpvec.f90:50-75
    if 2.0 * gkmax > cr.gmaxvr:
        raise RuntimeError("2*gkmax exceeds gmaxvr")
    ngp = int(sel.size)
    nrow = max(ngkmax, ngp)
    igpig = np.zeros(nrow, dtype=np.int32)
    igpig[:ngp] = sel + 1
    vgpc = np.zeros((3, nrow), order="F")
    vgpc[:, :ngp] = v[:, sel]
    r, th, ph = sphcrd(vgpc[:, :ngp])
    gpc = np.zeros(nrow)
    gpc[:ngp] = r
    tpgpc = np.zeros((2, nrow), order="F")
    tpgpc[0, :ngp] = th
    tpgpc[1, :ngp] = ph
    sf = np.zeros((nrow, cr.natm), dtype=np.complex128, order="F")
    sf[:ngp] = structure_factors(vgpc[:, :ngp], cr.atposc)
    return KPoint(vkl=f, vkc=vkc, ngp=ngp, gkmax=float(gkmax), igpig=igpig, vgpc=vgpc, gpc=gpc,
                  tpgpc=tpgpc, sfacgp=sf)


def pad_kpoint(kp: KPoint, ngkmax: int) -> KPoint:
    """Copy of kp with arrays padded to ngkmax rows (the reference's ngkmax leading dimension)."""
    if kp.igpig.shape[0] == ngkmax:
        return kp
    n = kp.ngp
    out = dataclasses.replace(kp)
    out.igpig = np.zeros(ngkmax, dtype=np.int32); out.igpig[:n] = kp.igpig[:n]
    out.vgpc = np.zeros((3, ngkmax), order="F"); out.vgpc[:, :n] = kp.vgpc[:, :n]
    out.gpc = np.zeros(ngkmax); out.gpc[:n] = kp.gpc[:n]
    out.tpgpc = np.zeros((2, ngkmax), order="F"); out.tpgpc[:, :n] = kp.tpgpc[:, :n]
    out.sfacgp = np.zeros((ngkmax, kp.sfacgp.shape[1]), dtype=np.complex128, order="F")
    out.sfacgp[:n] = kp.sfacgp[:n]
    return out


def make_config(cfg: int, nk: Optional[int] = None) -> Workload:
    """Build workload ``cfg``; ``nk`` limits the number of k-points (first nk of the seeded list)."""
    spec = _CONFIGS[cfg]
    rng = np.random.Generator(np.random.PCG64(20261202 + cfg))
    cr = make_crystal(cfg, rng)
    nkt = spec["nk"]
    if spec["nrange"] is not None:
        lo, hi = spec["nrange"]
        targets = rng.integers(lo, hi + 1, size=nkt)
    else:
        targets = [None] * nkt
    kps = [make_kpoint(cr, rng, (int(t) if t is not None else None), spec["gkmax"]) for t in targets]
    if nk is not None:
        kps = kps[:nk]
    ngkmax = max(k.ngp for k in kps)
    kps = [pad_kpoint(k, ngkmax) for k in kps]
    return Workload(cfg=cfg, crystal=cr, kpoints=kps, ngkmax=ngkmax)


def make_small(natm_per_species=(1, 1), lmax=3, lorbl=(0, 1), n_target=40, nk=2, seed=7,
               cell=(7.0, 7.5, 8.0), rmt=(1.4, 1.2)) -> Workload:
    """A tiny workload for brute-force checks (same generator rules, smaller sizes)."""
    key = 1000 + seed
    _CONFIGS[key] = dict(cell=cell, rmt=rmt[:len(natm_per_species)], natoms=natm_per_species, lmax=lmax,
                         lorbl=lorbl, nk=nk, nrange=(n_target, n_target), gkmax=None)
    try:
        rng = np.random.Generator(np.random.PCG64(seed))
        cr = make_crystal(key, rng)
        kps = [make_kpoint(cr, rng, n_target, None) for _ in range(nk)]
    finally:
        del _CONFIGS[key]
    ngkmax = max(k.ngp for k in kps)
    kps = [pad_kpoint(k, ngkmax) for k in kps]
    return Workload(cfg=key, crystal=cr, kpoints=kps, ngkmax=ngkmax)


def with_positions(cr: Crystal, atposc: np.ndarray) -> Crystal:
    """The crystal with atoms moved to atposc (3, natm): Theta(G) is recomputed
    (gencfun.f90:52-91); V(G), the muffin-tin matrices and the form factors are kept, as the
    IBS force term holds them fixed (forcek.f90:267-275)."""
    pos = np.asfortranarray(np.asarray(atposc, dtype=float))
    cf, _ = step_function(cr.omega, cr.rmt, cr.natoms, cr.spidx, pos, cr.vgc, cr.gc)
    return dataclasses.replace(cr, atposc=pos, cfunig=cf)


def kpoint_at(kp: KPoint, cr: Crystal) -> KPoint:
    """The k-point's structure factors for the positions of cr (gensfacgp.f90:38-49)."""
    out = dataclasses.replace(kp)
    sf = np.zeros_like(kp.sfacgp)
    n = kp.ngp
    sf[:n] = structure_factors(kp.vgpc[:, :n], cr.atposc)
    out.sfacgp = sf
    return out


@dataclasses.dataclass
class RadialData:
    """Muffin-tin radial data of one SCF iteration (the inputs of hmlint and olprad)."""
    lmaxvr: int
    nrmt: np.ndarray          # (nspecies,) int32
    spr: np.ndarray           # (nrmtmax, nspecies) radial mesh, Fortran order
    apwfr: np.ndarray         # (nrmtmax, 2, apwordmax, lmaxapw+1, natm): u, du/dr
    lofr: np.ndarray          # (nrmtmax, 2, nlomax, natm): u, du/dr
    veffmt: np.ndarray        # (lmmaxvr, nrmtmax, natm): coefficients of the real harmonics


def make_radial(cr: Crystal, nr=(300,), lmaxvr=6, seed=0, rmin=1e-5) -> RadialData:
    """Seeded radial data: per species a logarithmic mesh of nr[is] points from rmin to R
    (the form of genrmesh.f90), radial functions u = r^l (1 + a r + b r^2) e^(-c r) with
    a, b in U(-0.5, 0.5), c in U(0.5, 1.5) per function and atom (derivative analytic), and a
    potential whose spherical part is (-Z/r + d e^(-r)) sqrt(4 pi) (Z in U(5, 30) per species)
    and whose lm > 0 coefficients are N(0, 0.3) r^2 e^(-r)."""
    rng = np.random.default_rng(20261202 + 77 + seed)
    ns = cr.nspecies
    nrs = np.array([nr[min(i, len(nr) - 1)] for i in range(ns)], dtype=np.int32)
    nmax = int(nrs.max())
    spr = np.zeros((nmax, ns), order="F")
    for i in range(ns):
        n = int(nrs[i])
        t = math.log(cr.rmt[i] / rmin) / (n - 1)
        spr[:n, i] = rmin * np.exp(np.arange(n) * t)
        spr[n - 1, i] = cr.rmt[i]

    def fn(r, l):
        a, b = rng.uniform(-0.5, 0.5, 2)
        c = rng.uniform(0.5, 1.5)
        p = 1.0 + a * r + b * r * r
        dp = a + 2.0 * b * r
        e = np.exp(-c * r)
        u = r ** l * p * e
        du = (l * r ** (l - 1) if l > 0 else 0.0) * p * e + r ** l * (dp - c * p) * e
        return u, du

    apwfr = np.zeros((nmax, 2, cr.apwordmax, cr.lmaxapw + 1, cr.natm), order="F")
    lofr = np.zeros((nmax, 2, max(cr.nlomax, 1), cr.natm), order="F")
    lmv = (lmaxvr + 1) ** 2
    veffmt = np.zeros((lmv, nmax, cr.natm), order="F")
    zs = rng.uniform(5.0, 30.0, ns)
    for a in range(cr.natm):
        is_ = int(cr.spidx[a])
        n = int(nrs[is_])
        r = spr[:n, is_]
        for l in range(cr.lmaxapw + 1):
            for io in range(cr.apwordmax):
                apwfr[:n, 0, io, l, a], apwfr[:n, 1, io, l, a] = fn(r, l)
        for ilo in range(int(cr.nlorb[is_])):
            lofr[:n, 0, ilo, a], lofr[:n, 1, ilo, a] = fn(r, int(cr.lorbl[ilo, is_]))
        d = rng.uniform(-2.0, 2.0)
        veffmt[0, :n, a] = (-zs[is_] / r + d * np.exp(-r)) * math.sqrt(4.0 * math.pi)
        veffmt[1:, :n, a] = rng.normal(0.0, 0.3, (lmv - 1, 1)) * (r * r * np.exp(-r))[None, :]
    return RadialData(lmaxvr=lmaxvr, nrmt=nrs, spr=spr, apwfr=apwfr, lofr=lofr, veffmt=veffmt)
