"""The direct-route muffin-tin density on the device (exc_rhovalk / exc_rhovalk_dev, the
counterpart of rhovalk.f90:99-207) against the oracle's rhovalk (pinned in
test_oracle_rhovalk.py against the reference's sphcover and the density-matrix route).

Bar: rhomt within 1e-12 of the oracle, relative to its largest value (the same sums in
another order)."""
import numpy as np
import pytest

from conftest import small, workload

pytestmark = pytest.mark.gpu


def _case(which, lmaxvr, nst, seed):
    from exciting_amd import synth
    from oracle import lapw_ref as R
    w = small("d") if which == "small-d" else workload(1, 1)
    cr, kp = w.crystal, w.kpoints[0]
    rd = synth.make_radial(cr, nr=(120, 100) if which == "small-d" else (500, 460), lmaxvr=lmaxvr, seed=seed)
    apw = R.match(cr, kp)
    rng = np.random.default_rng(seed)
    nmat = kp.ngp + cr.nlotot
    C = np.asfortranarray(rng.standard_normal((nmat, nst)) + 1j * rng.standard_normal((nmat, nst)))
    occ = rng.uniform(0.0, 2.0, nst)
    occ[1] = 0.0
    if nst > 3:
        occ[3] = 1e-12
    return cr, kp, rd, apw, C, occ


@pytest.mark.parametrize("which,lmaxvr,lradstep", [("small-d", 4, 1), ("small-d", 4, 3), ("workload1", 8, 1),
                                                   ("workload1", 6, 2)])
def test_rhovalk_against_oracle(which, lmaxvr, lradstep):
    from exciting_amd import lapw
    from oracle import rhovalk_ref as V
    cr, kp, rd, apw, C, occ = _case(which, lmaxvr, 9 if which == "small-d" else 40, 21)
    ctx = lapw.Context(cr)
    ctx.hmlint(rd, out=False)
    zb, _, rf = V.genshtmat(lmaxvr)
    nrmtmax = rd.spr.shape[0]
    rng = np.random.default_rng(4)
    rho0 = np.asfortranarray(rng.uniform(-1, 1, ((lmaxvr + 1) ** 2, nrmtmax, cr.natm)))
    got = ctx.rhovalk(np.asfortranarray(apw), kp.ngp, C, occ, 0.35, zb, rf, lmaxvr, nrmtmax, lradstep=lradstep,
                      rhomt=rho0.copy(order="F"))
    ref = V.rhovalk(cr, rd, apw, kp.ngp, C, occ, 0.35, lradstep=lradstep, sht=(zb, None, rf), rhomt=rho0.copy())
    e = np.abs(got - ref).max() / np.abs(ref - rho0).max()
    print(f"{which} lmaxvr {lmaxvr} lradstep {lradstep}: rhovalk max rel diff {e:.2e}", flush=True)
    assert e <= 1e-12


def test_rhovalk_band_limited_equals_density_matrix_route():
    """Wavefunctions of l <= 2 (radial functions zero above): the device rhovalk equals the
    device density-matrix route (exc_gendmatmt + exc_genrhomt) to rounding."""
    from exciting_amd import lapw
    from oracle import rhovalk_ref as V
    cr, kp, rd, apw, C, occ = _case("small-d", 4, 9, 5)
    rd.apwfr[:, :, :, 3:, :] = 0.0
    occ = np.where(occ > 1e-8, occ, 0.5)
    ctx = lapw.Context(cr)
    ctx.hmlint(rd, out=False)
    zb, _, rf = V.genshtmat(4)
    nrmtmax = rd.spr.shape[0]
    a = np.asfortranarray(apw)
    rho = ctx.rhovalk(a, kp.ngp, C, occ, 0.4, zb, rf, 4, nrmtmax)
    dm = ctx.gendmatmt(a, kp.ngp, C, occ, 0.4)
    rho_dm = ctx.genrhomt(dm, 4, nrmtmax)
    e = np.abs(rho - rho_dm).max() / np.abs(rho_dm).max()
    print(f"rhovalk vs density-matrix route on the device: {e:.2e}", flush=True)
    assert e <= 1e-12


def test_rhovalk_dev_equals_host_entry():
    import torch
    from exciting_amd import lapw
    from oracle import rhovalk_ref as V
    cr, kp, rd, apw, C, occ = _case("workload1", 8, 24, 9)
    ctx = lapw.Context(cr)
    ctx.hmlint(rd, out=False)
    zb, _, rf = V.genshtmat(8)
    nrmtmax = rd.spr.shape[0]
    a = np.asfortranarray(apw)
    host = ctx.rhovalk(a, kp.ngp, C, occ, 0.25, zb, rf, 8, nrmtmax)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x.T)).cuda()   # Fortran order on the device
    rho = torch.zeros((cr.natm, nrmtmax, 81), dtype=torch.float64, device="cuda")
    ctx.rhovalk_dev(1, t(zb), t(rf), kp.ngp, a.shape[0], t(a), C.shape[1], C.shape[0], t(C),
                    torch.from_numpy(occ).cuda(), 0.25, rho)
    torch.cuda.synchronize()
    dev = rho.cpu().numpy().T
    assert np.array_equal(dev, host)


def test_rhovalk_with_evecsv():
    """The states evecfv evecsv (tevecsv true, one spinor) against the oracle's sum of the
    first-variational wavefunctions weighted by evecsv (no entry below epsocc in a random
    unitary evecsv, so the reference's skip leaves nothing out)."""
    from exciting_amd import lapw
    from oracle import rhovalk_ref as V
    cr, kp, rd, apw, C, occ = _case("small-d", 4, 9, 31)
    rng = np.random.default_rng(6)
    X = rng.standard_normal((9, 9)) + 1j * rng.standard_normal((9, 9))
    U = np.asfortranarray(np.linalg.qr(X)[0])
    assert (np.abs(U.real) + np.abs(U.imag)).min() > 1e-8
    ctx = lapw.Context(cr)
    ctx.hmlint(rd, out=False)
    zb, _, rf = V.genshtmat(4)
    nrmtmax = rd.spr.shape[0]
    got = ctx.rhovalk(np.asfortranarray(apw), kp.ngp, C, occ, 0.35, zb, rf, 4, nrmtmax, evecsv=U)
    ref = V.rhovalk(cr, rd, apw, kp.ngp, C, occ, 0.35, sht=(zb, None, rf), evecsv=U)
    e = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"rhovalk with evecsv: max rel diff {e:.2e}", flush=True)
    assert e <= 1e-12
