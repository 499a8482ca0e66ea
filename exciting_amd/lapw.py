"""Python binding of the gfx950 LAPW H+O setup (ctypes over ``include/exc_lapw.h``).

Mirrors the reference's per-k interface:

* ``Context(crystal)``                 — uploads the SCF state (module data of the reference)
* ``Context.match(kp)``                — match.f90:13, host arrays in, apwalm out
* ``Context.hamiltonandoverlapsetup``  — hamiltonandoverlapsetup.f90:8, (H, O) out
* ``Context.match_dev / setup_dev``    — the same on device tensors (torch, cuda)

The library is loaded from ``exciting_amd/_lib/libexcgpu.so``; there is no fallback: if
the library is missing, ``load_library`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIBPATH = Path(__file__).resolve().parent / "_lib" / "libexcgpu.so"

EXC_OK = 0
EXC_ERR_ARG = -1
EXC_ERR_HIP = -2
EXC_ERR_UNSUPPORTED = -3
EXC_ERR_STATE = -4


class ExcError(RuntimeError):
    def __init__(self, where, code):
        super().__init__(f"{where} returned {code}")
        self.code = code


class ExcDims(C.Structure):
    _fields_ = [("natmtot", C.c_int), ("nspecies", C.c_int), ("lmaxapw", C.c_int), ("lmaxmat", C.c_int),
                ("apwordmax", C.c_int), ("nlomax", C.c_int), ("lolmmax", C.c_int), ("maxnlo", C.c_int),
                ("ngrtot", C.c_int), ("ngvec", C.c_int), ("intgv", C.c_int * 6), ("omega", C.c_double)]


class KPointDev(C.Structure):
    """exc_kpoint_dev: one k-point of a grouped launch (device pointers)."""
    _fields_ = [("ngp", C.c_int), ("ngkmax", C.c_int), ("ldh", C.c_int), ("reserved", C.c_int),
                ("gpc", C.c_void_p), ("tpgpc", C.c_void_p), ("sfacgp", C.c_void_p), ("igpig", C.c_void_p),
                ("vgpc", C.c_void_p), ("apwalm", C.c_void_p), ("h", C.c_void_p), ("o", C.c_void_p)]


EXC_BATCH_MATCH = 1
EXC_BATCH_GRAPH = 2
EXC_BATCH_SPLIT_BF16 = 4
EXC_BATCH_OLP_LEGENDRE = 8
EXC_BATCH_SPLIT_CRT = 16

_lib = None

_P = C.c_void_p
_I = C.c_int

# name -> argtypes (restype int)
SIGNATURES = {
    "exc_ctx_create": [C.POINTER(ExcDims), C.POINTER(_P)],
    "exc_ctx_destroy": [_P],
    "exc_ctx_set_stream": [_P, _P],
    "exc_ctx_set_atoms": [_P, _P, _P, _P],
    "exc_ctx_set_lo": [_P, _P, _P, _P],
    "exc_ctx_set_radial": [_P, _P],
    "exc_ctx_set_mt": [_P, _P, _P, _P, _P, _P],
    "exc_ctx_set_gvec": [_P, _P, _P, _P, _P, _P],
    "exc_match": [_P, _I, _I, _P, _P, _P, _P],
    "exc_match_dev": [_P, _I, _I, _P, _P, _P, _P],
    "exc_hamiltonandoverlapsetup": [_P, _I, _I, _P, _P, _P, _I, _P, _P],
    "exc_hamiltonandoverlapsetup_dev": [_P, _I, _I, _P, _P, _P, _I, _P, _P],
    "exc_hamiltonandoverlapsetup_noalloc_dev": [_P, _I, _I, _P, _P, _P, _I, _P, _P],
    "exc_hamiltonsetup": [_P, _I, _I, _P, _P, _P, _I, _P],
    "exc_hamiltonsetup_dev": [_P, _I, _I, _P, _P, _P, _I, _P],
    "exc_overlapsetup": [_P, _I, _I, _P, _P, _P, _I, _P],
    "exc_overlapsetup_dev": [_P, _I, _I, _P, _P, _P, _I, _P],
    "exc_ctx_nlotot": [_P],
    "exc_ctx_status": [_P],
    "exc_ctx_set_timing": [_P, _I],
    "exc_ctx_read_timing": [_P, _P, _P],
    "exc_ctx_set_variant": [_P, _I, _I],
    "exc_ctx_set_moduli": [_P, _I],
    "exc_crt_moduli": [_I, _P, _P],
    "exc_ctx_num_variants": [_I],
    "exc_setup_batched_dev": [_P, _I, _P, _I],
    "exc_ho_vjp_dev": [_P, _I, _I, _P, _P, _P, _P, _I, _I, _P, _P, _P, _P, _P],
    "exc_ho_vjp": [_P, _I, _I, _P, _P, _P, _P, _I, _I, _P, _P, _P, _P, _P],
    "exc_eigen_dev": [_P, _I, _I, _P, _P, _I, _P, _P, _I],
    "exc_eigen": [_P, _I, _I, _P, _P, _I, _P, _P, _I],
    "exc_eigen_batched_dev": [_P, _I, _P, _P, _P, _P, _I, _P, _P, _P],
    "exc_dmat_wfsize": [_P],
    "exc_ctx_hmlint": [_P, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "exc_ctx_set_radial_fr": [_P, _I, _P, _P, _P],
    "exc_ctx_gencfun": [_P, _P, _P, _P],
    "exc_genrhomt_dev": [_P, _P, _P],
    "exc_genrhomt": [_P, _P, _P],
    "exc_rhovalk_dev": [_P, _I, _P, _P, _I, _I, _P, _I, _I, _P, _P, _P, C.c_double, C.c_double, _P],
    "exc_rhovalk": [_P, _I, _P, _P, _I, _I, _P, _I, _I, _P, _P, _P, C.c_double, C.c_double, _P],
    "exc_genrhoir_dev": [_P, _I, _P, _I, _I, _P, _P, C.c_double, C.c_double, _P],
    "exc_genrhoir_batched_dev": [_P, _I, _P, _I, _P, _P, _P, _P, C.c_double, _P],
    "exc_genrhoir": [_P, _I, _P, _I, _I, _P, _P, C.c_double, C.c_double, _P],
    "exc_ctx_genveffig": [_P, _P, _P],
    "exc_gendmatmt_dev": [_P, _I, _I, _P, _I, _I, _P, _P, _P, C.c_double, _P],
    "exc_gendmatmt": [_P, _I, _I, _P, _I, _I, _P, _P, _P, C.c_double, _P],
    "exc_gendmatmt_batched_dev": [_P, _I, _P, _I, _P, _P, _P, _P, _P, _P],
    "exc_ctx_set_meffig": [_P, _P],
    "exc_ctx_set_iora": [_P, _P, _P, _P, _P],
    "exc_ctx_crt_form": [_P, _P, _P, _P],
    "exc_ctx_set_cfdamp": [_P, C.c_double, C.c_double],
    "exc_gengpvec": [_P, _P, _P, _P, C.c_double, _I, _P, _P, _P, _P, _P, _P],
    "exc_gensfacgp": [_P, _I, _P, _I, _P],
    "exc_gengpvec_dev": [_P, _I, _P, _P, _P, C.c_double, _I, _P, _P, _P, _P, _P, _P, _P],
}

# stage order of exc_ctx_read_timing
STAGES = ("match", "ystage", "mt", "lo", "bmt", "graph", "split", "bsplit", "vjp", "olp", "eigen", "dmat", "hmlint",
          "scf", "crtprep", "crtgemm", "crtmerge")
# stages that take a variant but have no timer of their own (EXC_STAGE_CRTFORM)
VARIANT_STAGES = STAGES + ("crtform",)


def load_library(path: os.PathLike | None = None):
    """Load libexcgpu.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIBPATH
    if not p.exists():
        raise OSError(f"{p} not built: run `python -m exciting_amd.build`")
    try:
        # torch ships its own HIP runtime (loaded as "libamdhip64.so"); importing it first lets
        # the library's libamdhip64.so.7 dependency resolve to that same runtime instead of a
        # second copy, so device memory and streams are shared with torch
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(p))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    if path is None:
        _lib = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()            # torch tensor


def _f(a, dtype):
    """Fortran-ordered contiguous copy (or the array itself when it already is one)."""
    return np.asfortranarray(a, dtype=dtype)


def dims_of(cr) -> ExcDims:
    d = ExcDims()
    d.natmtot = cr.natm
    d.nspecies = cr.nspecies
    d.lmaxapw = cr.lmaxapw
    d.lmaxmat = cr.lmaxmat
    d.apwordmax = cr.apwordmax
    d.nlomax = cr.nlomax
    d.lolmmax = cr.lolmmax
    d.maxnlo = cr.maxnlo
    d.ngrtot = cr.ngrtot
    d.ngvec = cr.ngvec
    iv = np.asarray(cr.intgv, dtype=np.int32)
    for i in range(3):
        d.intgv[i] = int(iv[i, 0])
        d.intgv[3 + i] = int(iv[i, 1])
    d.omega = cr.omega
    return d


class Context:
    """One device context holding the SCF-constant state of a crystal."""

    def __init__(self, cr, stream=None):
        self.lib = load_library()
        self.cr = cr
        h = C.c_void_p()
        self._keep = []
        self._call("exc_ctx_create", C.byref(dims_of(cr)), C.byref(h))
        self.h = h
        if stream is not None:
            self.set_stream(stream)
        self.upload(cr)

    def _call(self, name, *args, ok_positive=False):
        rc = getattr(self.lib, name)(*args)
        if rc < 0 or (rc > 0 and not ok_positive):
            raise ExcError(name, rc)
        return rc

    def set_stream(self, stream):
        """stream: a torch.cuda.Stream, an int handle, or None."""
        handle = getattr(stream, "cuda_stream", stream)
        self._call("exc_ctx_set_stream", self.h, C.c_void_p(handle) if handle else None)

    def upload(self, cr):
        k = self._keep = []

        def hold(a):
            k.append(a)
            return _ptr(a)

        self._call("exc_ctx_set_atoms", self.h, hold(_f(cr.natoms, np.int32)), hold(_f(cr.rmt, np.float64)),
                   hold(_f(cr.atposc, np.float64)))
        self._call("exc_ctx_set_lo", self.h, hold(_f(cr.nlorb, np.int32)), hold(_f(cr.lorbl, np.int32)),
                   hold(_f(cr.idxlo, np.int32)))
        self.radial_info = self._call("exc_ctx_set_radial", self.h, hold(_f(cr.dmat, np.float64)), ok_positive=True)
        self._call("exc_ctx_set_mt", self.h, hold(_f(cr.haa, np.complex128)), hold(_f(cr.loa, np.complex128)),
                   hold(_f(cr.lolo, np.complex128)), hold(_f(cr.oalo, np.float64)), hold(_f(cr.ololo, np.float64)))
        self._call("exc_ctx_set_gvec", self.h, hold(_f(cr.ivg, np.int32)), hold(_f(cr.ivgig, np.int32)),
                   hold(_f(cr.cfunig, np.complex128)), hold(_f(cr.veffig, np.complex128)),
                   hold(_f(cr.ffacg, np.float64)))
        self._keep = []

    @property
    def nlotot(self):
        return self.lib.exc_ctx_nlotot(self.h)

    def status(self):
        """exc_ctx_status: raises ExcError if a setup since the last call met a basis difference
        G_i - G_j outside the first ngvec G-vectors (the reference's 2 gkmax > gmaxvr stop)."""
        self._call("exc_ctx_status", self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.exc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ host-array entry points
    def match(self, kp, ngkmax=None, out=None):
        """apwalm(ngkmax, apwordmax, lmmaxapw, natm) for k-point kp (match.f90:13); `out`: a
        preallocated Fortran-ordered complex128 array of that shape."""
        cr = self.cr
        ngkmax = kp.igpig.shape[0] if ngkmax is None else ngkmax
        shape = (ngkmax, cr.apwordmax, cr.lmmaxapw, cr.natm)
        if out is None:
            apw = np.zeros(shape, dtype=np.complex128, order="F")
        else:
            if out.shape != shape or out.dtype != np.complex128 or not out.flags.f_contiguous:
                raise ValueError("match: out must be a Fortran-ordered complex128 array of shape %s" % (shape,))
            apw = out
        gpc = _f(kp.gpc[:ngkmax], np.float64)
        tp = _f(kp.tpgpc[:, :ngkmax], np.float64)
        sf = _f(kp.sfacgp[:ngkmax], np.complex128)
        self.match_info = self._call("exc_match", self.h, kp.ngp, ngkmax, _ptr(gpc), _ptr(tp), _ptr(sf), _ptr(apw),
                                     ok_positive=True)
        return apw

    def hamiltonandoverlapsetup(self, kp, apwalm, ldh=None, out=None, resident=False):
        """(H, O) of k-point kp (hamiltonandoverlapsetup.f90:8), Fortran-ordered (ldh, nmat);
        `out`: a preallocated pair of such arrays.  resident: apwalm is the array the last
        match filled, and the library's device copy of it is used (NULL to the entry, as the
        Fortran driver passes it): no upload."""
        nmat = kp.ngp + self.nlotot
        ldh = nmat if ldh is None else ldh
        ngkmax = apwalm.shape[0]
        if out is None:
            H = np.zeros((ldh, nmat), dtype=np.complex128, order="F")
            O = np.zeros((ldh, nmat), dtype=np.complex128, order="F")
        else:
            H, O = out
            for m in (H, O):
                if m.shape != (ldh, nmat) or m.dtype != np.complex128 or not m.flags.f_contiguous:
                    raise ValueError("hamiltonandoverlapsetup: out must be Fortran-ordered complex128 (ldh, nmat)")
        apw = None if resident else _f(apwalm, np.complex128)
        ig = _f(kp.igpig[:ngkmax], np.int32)
        vg = _f(kp.vgpc[:, :ngkmax], np.float64)
        self._call("exc_hamiltonandoverlapsetup", self.h, kp.ngp, ngkmax, _ptr(apw), _ptr(ig), _ptr(vg), ldh,
                   _ptr(H), _ptr(O))
        return H, O

    def eigen(self, H, O, nev):
        """Lowest nev eigenpairs of H c = e O c (solvewithlapack, ZHEGVX range 'I') from host
        arrays; H and O are overwritten as ZHEGVX leaves them.  Returns (eval, evec(nmat, nev))."""
        nmat = H.shape[1]
        ldh = H.shape[0]
        for m in (H, O):
            if m.shape != (ldh, nmat) or m.dtype != np.complex128 or not m.flags.f_contiguous:
                raise ValueError("eigen: H, O must be Fortran-ordered complex128 (ldh, nmat)")
        ev = np.zeros(nev)
        vec = np.zeros((nmat, nev), dtype=np.complex128, order="F")
        self._call("exc_eigen", self.h, nmat, ldh, _ptr(H), _ptr(O), nev, _ptr(ev), _ptr(vec), nmat,
                   ok_positive=False)
        return ev, vec

    def eigen_dev(self, nmat, ldh, h, o, nev, eval_, evec, ldv):
        """Device version: torch tensors h, o (stored as (nmat, ldh) row-major = Fortran (ldh, nmat)),
        eval_(nev) float64, evec (nev, ldv) complex128."""
        return self._call("exc_eigen_dev", self.h, nmat, ldh, _ptr(h), _ptr(o), nev, _ptr(eval_), _ptr(evec), ldv)

    def eigen_batched_dev(self, nmats, hs, os_, nev, evals, evecs):
        """exc_eigen_batched_dev over lists of torch tensors (per k as for eigen_dev, ld = nmat)."""
        nk = len(nmats)
        ints = (C.c_int * nk)(*nmats)
        ptrs = [(C.c_void_p * nk)(*[_ptr(t) for t in lst]) for lst in (hs, os_, evals, evecs)]
        return self._call("exc_eigen_batched_dev", self.h, nk, C.cast(ints, _P), C.cast(ints, _P),
                          C.cast(ptrs[0], _P), C.cast(ptrs[1], _P), nev, C.cast(ptrs[2], _P), C.cast(ptrs[3], _P),
                          C.cast(ints, _P))

    def genrhomt(self, dm, lmaxvr, nrmtmax):
        """exc_genrhomt from a host density matrix (wfsize, wfsize, natm) Fortran-ordered:
        rhomt((lmaxvr+1)^2, nrmtmax, natm) with the radial data of the last hmlint call."""
        dm = np.asfortranarray(dm, dtype=np.complex128)
        out = np.zeros(((lmaxvr + 1) ** 2, nrmtmax, self.cr.natm), order="F")
        self._call("exc_genrhomt", self.h, _ptr(dm), _ptr(out))
        return out

    def rhovalk(self, apwalm, ngp, evec, occ, wkpt, zbshtvr, rfshtvr, lmaxvr, nrmtmax, lradstep=1, epsocc=1e-8,
                rhomt=None, evecsv=None):
        """exc_rhovalk from host arrays (the radial data of the last hmlint call): apwalm
        (ngkmax, apwordmax, lmmaxapw, natm) and evec (ldv, nst) Fortran complex128, occ (nst,),
        zbshtvr / rfshtvr (lmmaxvr, lmmaxvr) Fortran complex128 / float64; rhomt (lmmaxvr,
        nrmtmax, natm) Fortran float64 accumulated (new when None); evecsv (nst, nst) Fortran
        complex128 or None for the identity."""
        lmv = (lmaxvr + 1) ** 2
        if rhomt is None:
            rhomt = np.zeros((lmv, nrmtmax, self.cr.natm), order="F")
        for a in (apwalm, evec, zbshtvr) + ((evecsv,) if evecsv is not None else ()):
            if a.dtype != np.complex128 or not a.flags.f_contiguous:
                raise ValueError("rhovalk: complex arrays must be Fortran-ordered complex128")
        if rhomt.shape != (lmv, nrmtmax, self.cr.natm) or not rhomt.flags.f_contiguous:
            raise ValueError("rhovalk: rhomt must be Fortran (lmmaxvr, nrmtmax, natm)")
        rf = np.asfortranarray(rfshtvr, dtype=np.float64)
        occ = np.ascontiguousarray(occ, dtype=np.float64)
        self._call("exc_rhovalk", self.h, int(lradstep), _ptr(zbshtvr), _ptr(rf), int(ngp), apwalm.shape[0],
                   _ptr(apwalm), evec.shape[1], evec.shape[0], _ptr(evec),
                   _ptr(evecsv) if evecsv is not None else None, _ptr(occ), float(wkpt), float(epsocc), _ptr(rhomt))
        return rhomt

    def rhovalk_dev(self, lradstep, zbshtvr, rfshtvr, ngp, ngkmax, apwalm, nst, ldv, evec, occ, wkpt, rhomt,
                    epsocc=1e-8, evecsv=None):
        """Device version (torch tensors; rhomt (natm, nrmtmax, lmmaxvr) row-major = Fortran
        (lmmaxvr, nrmtmax, natm); zbshtvr / rfshtvr transposed likewise)."""
        return self._call("exc_rhovalk_dev", self.h, int(lradstep), _ptr(zbshtvr), _ptr(rfshtvr), int(ngp),
                          int(ngkmax), _ptr(apwalm), int(nst), int(ldv), _ptr(evec),
                          _ptr(evecsv) if evecsv is not None else None, _ptr(occ), float(wkpt), float(epsocc),
                          _ptr(rhomt))

    def genrhoir(self, kp, evec, occ, wkpt, epsocc=1e-8, rhoir=None):
        """exc_genrhoir from host arrays: evec (ldv, nst) Fortran complex128; rhoir (ngrtot,)
        accumulated (new when None)."""
        evec = np.asfortranarray(evec, dtype=np.complex128)
        occ = np.ascontiguousarray(occ, dtype=np.float64)
        ig = np.ascontiguousarray(kp.igpig[:kp.ngp], dtype=np.int32)
        out = np.zeros(self.cr.ngrtot) if rhoir is None else rhoir
        self._call("exc_genrhoir", self.h, kp.ngp, _ptr(ig), evec.shape[1], evec.shape[0], _ptr(evec), _ptr(occ),
                   float(wkpt), float(epsocc), _ptr(out))
        return out

    def genrhoir_batched_dev(self, batch, nst, evecs, occs, wkpts, rhoir, epsocc=1e-8):
        """exc_genrhoir_batched_dev over a DeviceBatch: evecs[k] (nst, nmat_k) tensors."""
        nk = batch.nk
        ldv = (C.c_int * nk)(*batch.nmat)
        ev = (C.c_void_p * nk)(*[_ptr(t) for t in evecs])
        oc = (C.c_void_p * nk)(*[_ptr(t) for t in occs])
        wk = (C.c_double * nk)(*[float(x) for x in wkpts])
        return self._call("exc_genrhoir_batched_dev", self.h, nk, C.cast(batch.desc, _P), nst, C.cast(ldv, _P),
                          C.cast(ev, _P), C.cast(oc, _P), C.cast(wk, _P), float(epsocc), _ptr(rhoir))

    def set_cfdamp(self, cfdamp, gmaxvr):
        """exc_ctx_set_cfdamp: gencfun's damping of Theta(G) (0: none)."""
        self._call("exc_ctx_set_cfdamp", self.h, float(cfdamp), float(gmaxvr))

    def gencfun(self):
        """exc_ctx_gencfun with the crystal's bvec: (cfunig(ngrtot), cfunir(ngrtot)) as host arrays."""
        cr = self.cr
        bvec = np.asfortranarray(cr.bvec, dtype=np.float64)
        cf = np.zeros(cr.ngrtot, dtype=np.complex128)
        cir = np.zeros(cr.ngrtot)
        self._call("exc_ctx_gencfun", self.h, _ptr(bvec), _ptr(cf), _ptr(cir))
        return cf, cir

    def set_meffig(self, meffig):
        """exc_ctx_set_meffig: the interstitial kinetic factor of H with ValenceRelativity /=
        'none' (ngvec complex), or None for the 'none' form (cfunig)."""
        if meffig is None:
            self._call("exc_ctx_set_meffig", self.h, None)
            return
        m = np.ascontiguousarray(meffig, dtype=np.complex128)
        self._call("exc_ctx_set_meffig", self.h, _ptr(m))

    def set_iora(self, h1aa=None, h1loa=None, h1lolo=None, m2effig=None):
        """exc_ctx_set_iora: the 'iora*' overlap terms (Fortran layouts: h1aa(apwordmax,
        apwordmax, 0:lmaxapw, natmtot), h1loa(apwordmax, nlomax, natmtot), h1lolo(nlomax,
        nlomax, natmtot), m2effig(ngvec) complex); None turns a term off."""
        keep = []

        def arr(x, dt):
            if x is None:
                return None
            a = np.asfortranarray(x, dtype=dt)
            keep.append(a)
            return _ptr(a)
        self._call("exc_ctx_set_iora", self.h, arr(h1aa, np.float64), arr(h1loa, np.float64),
                   arr(h1lolo, np.float64), arr(m2effig, np.complex128))

    def gengpvec(self, vpl, vpc, gkmax, ngkmax):
        """exc_gengpvec with the crystal's bvec: (ngp, igpig, vgpl, vgpc, gpc, tpgpc) as host
        arrays (Fortran layouts, rows past ngp zero)."""
        bvec = np.asfortranarray(self.cr.bvec, dtype=np.float64)
        vl = np.ascontiguousarray(vpl, dtype=np.float64)
        vc = np.ascontiguousarray(vpc, dtype=np.float64)
        ngp = C.c_int(0)
        igpig = np.zeros(ngkmax, dtype=np.int32)
        vgpl = np.zeros((3, ngkmax), order="F")
        vgpc = np.zeros((3, ngkmax), order="F")
        gpc = np.zeros(ngkmax)
        tpgpc = np.zeros((2, ngkmax), order="F")
        self._call("exc_gengpvec", self.h, _ptr(bvec), _ptr(vl), _ptr(vc), float(gkmax), int(ngkmax), C.byref(ngp),
                   _ptr(igpig), _ptr(vgpl), _ptr(vgpc), _ptr(gpc), _ptr(tpgpc))
        return ngp.value, igpig, vgpl, vgpc, gpc, tpgpc

    def gensfacgp(self, ngp, vgpc, ld):
        """exc_gensfacgp: S_a(G+k) of vgpc(3, >= ngp) as a host array (ld, natmtot)."""
        v = np.asfortranarray(vgpc[:, :ngp], dtype=np.float64)
        out = np.zeros((ld, self.cr.natm), dtype=np.complex128, order="F")
        self._call("exc_gensfacgp", self.h, int(ngp), _ptr(v), int(ld), _ptr(out))
        return out

    def gengpvec_dev(self, vpl, vpc, gkmax, ngkmax, sfacgp=True):
        """exc_gengpvec_dev for nk k-points (vpl, vpc: (nk, 3)): ngp (list) and torch device
        tensors igpig (nk, ngkmax), vgpl, vgpc (nk, ngkmax, 3), gpc (nk, ngkmax),
        tpgpc (nk, ngkmax, 2), sfacgp (nk, natm, ngkmax) complex or None."""
        import torch
        vl = np.ascontiguousarray(vpl, dtype=np.float64).reshape(-1, 3)
        vc = np.ascontiguousarray(vpc, dtype=np.float64).reshape(-1, 3)
        nk = vl.shape[0]
        bvec = np.asfortranarray(self.cr.bvec, dtype=np.float64)
        dev = torch.device("cuda")
        igpig = torch.empty((nk, ngkmax), dtype=torch.int32, device=dev)
        vgpl = torch.empty((nk, ngkmax, 3), dtype=torch.float64, device=dev)
        vgpc = torch.empty((nk, ngkmax, 3), dtype=torch.float64, device=dev)
        gpc = torch.empty((nk, ngkmax), dtype=torch.float64, device=dev)
        tpgpc = torch.empty((nk, ngkmax, 2), dtype=torch.float64, device=dev)
        sf = torch.empty((nk, self.cr.natm, ngkmax), dtype=torch.complex128, device=dev) if sfacgp else None
        ngp = (C.c_int * nk)()
        self._call("exc_gengpvec_dev", self.h, nk, _ptr(bvec), _ptr(vl), _ptr(vc), float(gkmax), int(ngkmax),
                   C.cast(ngp, _P), _ptr(igpig), _ptr(vgpl), _ptr(vgpc), _ptr(gpc), _ptr(tpgpc),
                   _ptr(sf) if sf is not None else None)
        return list(ngp), igpig, vgpl, vgpc, gpc, tpgpc, sf

    def genveffig(self, veffir):
        """exc_ctx_genveffig: V(G) (ngvec,) from V(r) on the grid (ngrtot,)."""
        v = np.ascontiguousarray(veffir, dtype=np.float64)
        out = np.zeros(self.cr.ngvec, dtype=np.complex128)
        self._call("exc_ctx_genveffig", self.h, _ptr(v), _ptr(out))
        return out

    def set_radial_fr(self, rd):
        """exc_ctx_set_radial_fr: match's boundary values from the radial functions of rd
        (returns the singular-matching info, 0 when none)."""
        spr = np.asfortranarray(rd.spr, dtype=np.float64)
        apw = np.asfortranarray(rd.apwfr, dtype=np.float64)
        nrmt = np.ascontiguousarray(rd.nrmt, dtype=np.int32)
        return self._call("exc_ctx_set_radial_fr", self.h, rd.spr.shape[0], _ptr(nrmt), _ptr(spr), _ptr(apw),
                          ok_positive=True)

    def hmlint(self, rd, out=True):
        """exc_ctx_hmlint from synth.RadialData: computes h_a, loa, lolo, oalo, ololo on the GPU
        and installs them in the context; with out, also returns them as host arrays in the
        shapes of Crystal.haa / loa / lolo / oalo / ololo."""
        cr = self.cr
        nrmtmax = rd.spr.shape[0]
        spr = np.asfortranarray(rd.spr, dtype=np.float64)
        apw = np.asfortranarray(rd.apwfr, dtype=np.float64)
        lofr = np.asfortranarray(rd.lofr, dtype=np.float64)
        v = np.asfortranarray(rd.veffmt, dtype=np.float64)
        nrmt = np.ascontiguousarray(rd.nrmt, dtype=np.int32)
        res = None
        if out:
            mn, nl = max(cr.maxnlo, 1), max(cr.nlomax, 1)
            res = (np.zeros((cr.maxaa, cr.maxaa, cr.natm), dtype=np.complex128, order="F"),
                   np.zeros((mn, cr.maxaa, cr.natm), dtype=np.complex128, order="F"),
                   np.zeros((mn, mn, cr.natm), dtype=np.complex128, order="F"),
                   np.zeros((cr.apwordmax, nl, cr.natm), order="F"),
                   np.zeros((nl, nl, cr.natm), order="F"))
        self._call("exc_ctx_hmlint", self.h, rd.lmaxvr, nrmtmax, _ptr(nrmt), _ptr(spr), _ptr(apw), _ptr(lofr),
                   _ptr(v), *[_ptr(a) if res else None for a in (res or (None,) * 5)])
        return res

    def dmat_wfsize(self):
        """Row count of the muffin-tin density matrix (maxaa + largest lo count of an atom)."""
        rc = self.lib.exc_dmat_wfsize(self.h)
        if rc < 0:
            raise ExcError("exc_dmat_wfsize", rc)
        return rc

    def gendmatmt(self, apwalm, ngp, evec, occ, wkpt, evecsv=None, dm=None):
        """gendmatmt from host arrays: apwalm (ngkmax, apwordmax, lmmaxapw, natm) and evec
        (ldv, nst) Fortran-ordered complex128, occ (nst,), evecsv (nst, nst) or None; dm
        (wfsize, wfsize, natm) Fortran-ordered complex128 is accumulated (new when None)."""
        ws = self.dmat_wfsize()
        if dm is None:
            dm = np.zeros((ws, ws, self.cr.natm), dtype=np.complex128, order="F")
        for a in (apwalm, evec, dm) + ((evecsv,) if evecsv is not None else ()):
            if a.dtype != np.complex128 or not a.flags.f_contiguous:
                raise ValueError("gendmatmt: complex arrays must be Fortran-ordered complex128")
        if dm.shape != (ws, ws, self.cr.natm):
            raise ValueError("gendmatmt: dm must be (wfsize, wfsize, natm)")
        occ = np.ascontiguousarray(occ, dtype=np.float64)
        nst = evec.shape[1]
        self._call("exc_gendmatmt", self.h, ngp, apwalm.shape[0], _ptr(apwalm), nst, evec.shape[0], _ptr(evec),
                   _ptr(evecsv) if evecsv is not None else None, _ptr(occ), float(wkpt), _ptr(dm))
        return dm

    def gendmatmt_dev(self, ngp, ngkmax, apwalm, nst, ldv, evec, occ, wkpt, dm, evecsv=None):
        """Device version (torch tensors; dm (natm, wfsize, wfsize) row-major = Fortran (wfsize, wfsize, natm))."""
        return self._call("exc_gendmatmt_dev", self.h, ngp, ngkmax, _ptr(apwalm), nst, ldv, _ptr(evec),
                          _ptr(evecsv) if evecsv is not None else None, _ptr(occ), float(wkpt), _ptr(dm))

    def gendmatmt_batched_dev(self, batch, nst, evecs, occs, wkpts, dm, evecsvs=None):
        """exc_gendmatmt_batched_dev over a DeviceBatch: evecs[k] (nst, nmat_k) tensors (ldv = nmat_k)."""
        nk = batch.nk
        ldv = (C.c_int * nk)(*batch.nmat)
        ev = (C.c_void_p * nk)(*[_ptr(t) for t in evecs])
        oc = (C.c_void_p * nk)(*[_ptr(t) for t in occs])
        wk = (C.c_double * nk)(*[float(x) for x in wkpts])
        sv = (C.c_void_p * nk)(*[_ptr(t) if t is not None else None for t in evecsvs]) if evecsvs else None
        return self._call("exc_gendmatmt_batched_dev", self.h, nk, C.cast(batch.desc, _P), nst, C.cast(ldv, _P),
                          C.cast(ev, _P), C.cast(sv, _P) if sv is not None else None, C.cast(oc, _P),
                          C.cast(wk, _P), _ptr(dm))

    # ------------------------------------------------------------------ device entry points
    def match_dev(self, ngp, ngkmax, gpc, tpgpc, sfacgp, apwalm):
        return self._call("exc_match_dev", self.h, ngp, ngkmax, _ptr(gpc), _ptr(tpgpc), _ptr(sfacgp), _ptr(apwalm),
                          ok_positive=True)

    def set_moduli(self, n):
        """Number of moduli of the modular split (EXC_BATCH_SPLIT_CRT); 0: the built-in count."""
        self._call("exc_ctx_set_moduli", self.h, int(n))

    def crt_form(self):
        """(form, moduli, sigma) of the last modular-split setup: form 0 A and Yt, 1 Gram."""
        f, n, s = C.c_int(), C.c_int(), C.c_double()
        self._call("exc_ctx_crt_form", self.h, C.byref(f), C.byref(n), C.byref(s))
        return f.value, n.value, s.value

    def set_timing(self, on=True):
        self._call("exc_ctx_set_timing", self.h, 1 if on else 0)

    def set_variant(self, stage, variant):
        """Pin the tile variant of a stage ('ystage' or 'mt'); -1 = built-in choice."""
        self._call("exc_ctx_set_variant", self.h, VARIANT_STAGES.index(stage), int(variant))

    def num_variants(self, stage):
        return self.lib.exc_ctx_num_variants(VARIANT_STAGES.index(stage))

    def read_timing(self):
        """{stage: (total_ms, launches)} recorded since the last read (synchronises the stream)."""
        n = len(STAGES)             # EXC_NSTAGE
        ms = (C.c_double * n)()
        cnt = (C.c_int * n)()
        self._call("exc_ctx_read_timing", self.h, C.cast(ms, C.c_void_p), C.cast(cnt, C.c_void_p))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(STAGES)}

    def setup_batched_dev(self, batch, flags=EXC_BATCH_MATCH):
        """All k-points of `batch` (a DeviceBatch) in one grouped launch per stage."""
        return self._call("exc_setup_batched_dev", self.h, batch.nk, C.cast(batch.desc, C.c_void_p), flags)

    def vjp_dev(self, ngp, ngkmax, apwalm, igpig, vgpc, sfacgp, evec, ldc, nst, wh, wo, dedtau, dedv):
        """Accumulate dL/dtau (3, natm) and dL/dV(G) for D = C diag(wh) C^H, W = C diag(wo) C^H."""
        return self._call("exc_ho_vjp_dev", self.h, ngp, ngkmax, _ptr(apwalm), _ptr(igpig), _ptr(vgpc),
                          _ptr(sfacgp), nst, ldc, _ptr(evec), _ptr(wh), _ptr(wo), _ptr(dedtau), _ptr(dedv))

    def setup_dev(self, ngp, ngkmax, apwalm, igpig, vgpc, ldh, H, O):
        return self._call("exc_hamiltonandoverlapsetup_dev", self.h, ngp, ngkmax, _ptr(apwalm), _ptr(igpig),
                          _ptr(vgpc), ldh, _ptr(H), _ptr(O))


class DeviceKPoint:
    """Per-k inputs resident in device memory (torch tensors in the reference's layouts)."""

    def __init__(self, kp, device="cuda"):
        import torch
        n = kp.igpig.shape[0]
        self.ngp = kp.ngp
        self.ngkmax = n
        self.gpc = torch.from_numpy(np.ascontiguousarray(kp.gpc)).to(device)
        self.tpgpc = torch.from_numpy(np.ascontiguousarray(kp.tpgpc.T)).to(device)          # (ngkmax, 2) == (2, ngkmax) F
        self.sfacgp = torch.from_numpy(np.ascontiguousarray(kp.sfacgp.T)).to(device)        # (natm, ngkmax) == F
        self.igpig = torch.from_numpy(np.ascontiguousarray(kp.igpig.astype(np.int32))).to(device)
        self.vgpc = torch.from_numpy(np.ascontiguousarray(kp.vgpc.T)).to(device)            # (ngkmax, 3) == (3, ngkmax) F


def fortran_view(t, shape):
    """A torch tensor allocated C-contiguous with reversed shape, seen as a Fortran array."""
    return t.permute(*reversed(range(len(shape))))


def alloc_apwalm_dev(cr, ngkmax, device="cuda"):
    import torch
    return torch.empty((cr.natm, cr.lmmaxapw, cr.apwordmax, ngkmax), dtype=torch.complex128, device=device)


def alloc_matrix_dev(ldh, nmat, device="cuda"):
    """A (ldh, nmat) Fortran-ordered complex matrix: stored as a (nmat, ldh) C tensor."""
    import torch
    return torch.empty((nmat, ldh), dtype=torch.complex128, device=device)


class DeviceBatch:
    """Device inputs and outputs of many k-points plus the exc_kpoint_dev array describing them."""

    def __init__(self, cr, kpoints, device="cuda"):
        self.dks = [DeviceKPoint(kp, device) for kp in kpoints]
        self.nk = len(kpoints)
        self.nmat = [kp.ngp + cr.nlotot for kp in kpoints]
        self.apw = [alloc_apwalm_dev(cr, d.ngkmax, device) for d in self.dks]
        self.H = [alloc_matrix_dev(n, n, device) for n in self.nmat]
        self.O = [alloc_matrix_dev(n, n, device) for n in self.nmat]
        self.desc = (KPointDev * self.nk)()
        for i, d in enumerate(self.dks):
            e = self.desc[i]
            e.ngp, e.ngkmax, e.ldh = d.ngp, d.ngkmax, self.nmat[i]
            e.gpc, e.tpgpc, e.sfacgp = _ptr(d.gpc), _ptr(d.tpgpc), _ptr(d.sfacgp)
            e.igpig, e.vgpc = _ptr(d.igpig), _ptr(d.vgpc)
            e.apwalm, e.h, e.o = _ptr(self.apw[i]), _ptr(self.H[i]), _ptr(self.O[i])
