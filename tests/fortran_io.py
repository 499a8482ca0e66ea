"""Stream-file exchange with the Fortran driver test program (exc_driver_main.f90)."""
import numpy as np


def _w(f, a, dtype):
    np.asarray(a, dtype=dtype).ravel(order="F").tofile(f)


def write_input(path, cr, kp):
    hdr = [cr.natm, cr.nspecies, cr.lmaxapw, cr.lmaxmat, cr.apwordmax, cr.nlomax, cr.lolmmax, cr.maxnlo,
           cr.ngrtot, cr.ngvec, *cr.intgv[:, 0], *cr.intgv[:, 1], kp.ngp, kp.igpig.shape[0]]
    with open(path, "wb") as f:
        _w(f, hdr, np.int32)
        _w(f, [cr.omega], np.float64)
        _w(f, cr.natoms, np.int32); _w(f, cr.rmt, np.float64); _w(f, cr.atposc, np.float64)
        _w(f, cr.nlorb, np.int32); _w(f, cr.lorbl, np.int32); _w(f, cr.idxlo, np.int32)
        _w(f, cr.dmat, np.float64); _w(f, cr.haa, np.complex128); _w(f, cr.loa, np.complex128)
        _w(f, cr.lolo, np.complex128); _w(f, cr.oalo, np.float64); _w(f, cr.ololo, np.float64)
        _w(f, cr.ivg, np.int32); _w(f, cr.ivgig, np.int32); _w(f, cr.cfunig, np.complex128)
        _w(f, cr.veffig, np.complex128); _w(f, cr.ffacg, np.float64)
        _w(f, kp.gpc, np.float64); _w(f, kp.tpgpc, np.float64); _w(f, kp.sfacgp, np.complex128)
        _w(f, kp.igpig, np.int32); _w(f, kp.vgpc, np.float64)


def read_output(path, cr, ngkmax):
    with open(path, "rb") as f:
        nmat = int(np.fromfile(f, np.int32, 1)[0])
        na = ngkmax * cr.apwordmax * cr.lmmaxapw * cr.natm
        apw = np.fromfile(f, np.complex128, na).reshape((ngkmax, cr.apwordmax, cr.lmmaxapw, cr.natm), order="F")
        H = np.fromfile(f, np.complex128, nmat * nmat).reshape((nmat, nmat), order="F")
        O = np.fromfile(f, np.complex128, nmat * nmat).reshape((nmat, nmat), order="F")
    return apw, H, O


def write_eigen(path, C, e, occ, wkpt):
    with open(path, "wb") as f:
        np.asarray([C.shape[1]], dtype=np.int32).tofile(f)
        np.asarray([wkpt], dtype=np.float64).tofile(f)
        np.asfortranarray(C, dtype=np.complex128).ravel(order="F").tofile(f)
        np.asarray(e, dtype=np.float64).tofile(f)
        np.asarray(occ, dtype=np.float64).tofile(f)


def read_forces(path, cr, ngkmax):
    with open(path, "rb") as f:
        nmat = int(np.fromfile(f, np.int32, 1)[0])
        skip = ngkmax * cr.apwordmax * cr.lmmaxapw * cr.natm + 2 * nmat * nmat
        np.fromfile(f, np.complex128, skip)
        return np.fromfile(f, np.float64, 3 * cr.natm).reshape((3, cr.natm), order="F")


def read_eigen_output(path, cr, ngkmax, nev):
    """evalfv(nev), evecfv(nmat, nev) appended by the driver's solve step (no forces)."""
    with open(path, "rb") as f:
        nmat = int(np.fromfile(f, np.int32, 1)[0])
        skip = ngkmax * cr.apwordmax * cr.lmmaxapw * cr.natm + 2 * nmat * nmat
        np.fromfile(f, np.complex128, skip)
        e = np.fromfile(f, np.float64, nev)
        C = np.fromfile(f, np.complex128, nmat * nev).reshape((nmat, nev), order="F")
    return e, C


def read_dmat(path, cr, ngkmax, ws):
    """dm(wfsize, wfsize, natm) appended by the driver after forceibs (gendmatmt_gpu)."""
    with open(path, "rb") as f:
        nmat = int(np.fromfile(f, np.int32, 1)[0])
        skip = ngkmax * cr.apwordmax * cr.lmmaxapw * cr.natm + 2 * nmat * nmat
        np.fromfile(f, np.complex128, skip)
        np.fromfile(f, np.float64, 3 * cr.natm)
        return np.fromfile(f, np.complex128, ws * ws * cr.natm).reshape((ws, ws, cr.natm), order="F")


def write_radial(path, cr, rd):
    """Radial data for the driver's hmlint_gpu step: lmaxvr, nrmtmax, nrmt, spr, apwfr, lofr, veffmt,
    then the transforms zbshtvr, rfshtvr of the oracle's genshtmat (for rhovalk_gpu)."""
    from oracle import rhovalk_ref as V
    zb, _, rf = V.genshtmat(rd.lmaxvr)
    with open(path, "wb") as f:
        _w(f, [rd.lmaxvr, rd.spr.shape[0]], np.int32)
        _w(f, rd.nrmt, np.int32)
        _w(f, rd.spr, np.float64)
        _w(f, rd.apwfr, np.float64)
        _w(f, rd.lofr, np.float64)
        _w(f, rd.veffmt, np.float64)
        _w(f, zb, np.complex128)
        _w(f, rf, np.float64)


def write_scf(path, cr, veffir):
    """bvec(3,3) and V(r) on the FFT grid for the driver's gencfun_gpu / genveffig_gpu step."""
    with open(path, "wb") as f:
        _w(f, cr.bvec, np.float64)
        _w(f, veffir, np.float64)


def read_rhoir(path, cr, ngkmax, ws):
    """rhoir(ngrtot) appended after dm (genrhoir_gpu)."""
    with open(path, "rb") as f:
        nmat = int(np.fromfile(f, np.int32, 1)[0])
        skip = ngkmax * cr.apwordmax * cr.lmmaxapw * cr.natm + 2 * nmat * nmat + ws * ws * cr.natm
        np.fromfile(f, np.complex128, skip)
        np.fromfile(f, np.float64, 3 * cr.natm)
        return np.fromfile(f, np.float64, cr.ngrtot)


def read_rhomt(path, cr, ngkmax, ws, lmv, nrmtmax, direct=False):
    """rhomt(lmv, nrmtmax, natm) appended after dm and rhoir when the driver had radial data
    (genrhomt_gpu); with direct, the one after it (rhovalk_gpu)."""
    with open(path, "rb") as f:
        nmat = int(np.fromfile(f, np.int32, 1)[0])
        skip = ngkmax * cr.apwordmax * cr.lmmaxapw * cr.natm + 2 * nmat * nmat + ws * ws * cr.natm
        np.fromfile(f, np.complex128, skip)
        np.fromfile(f, np.float64, 3 * cr.natm + cr.ngrtot + (lmv * nrmtmax * cr.natm if direct else 0))
        return np.fromfile(f, np.float64, lmv * nrmtmax * cr.natm).reshape((lmv, nrmtmax, cr.natm), order="F")
