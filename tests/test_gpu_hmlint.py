"""Muffin-tin integrals on the device (exc_ctx_hmlint: hmlint = mt_kin + mt_pot, olprad) against
the oracle (oracle/hmlint_ref.py, pinned in test_oracle_hmlint.py), and the setup consuming
the matrices the call installs.

Bar: max |difference| <= 1e-12 of the largest element (FP64; the device sums the spline
integral as a weight vector, the oracle interval by interval)."""
import dataclasses

import numpy as np
import pytest

from conftest import small, workload

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("which", ["small-d", "workload1"])
def test_hmlint_against_oracle(which):
    from exciting_amd import lapw, synth
    from oracle import hmlint_ref as M
    w = small("d") if which == "small-d" else workload(1, 1)
    cr = w.crystal
    rd = synth.make_radial(cr, nr=(150, 120) if which == "small-d" else (260, 230),
                           lmaxvr=4 if which == "small-d" else 6)
    ctx = lapw.Context(cr)
    haa, loa, lolo, oalo, ololo = ctx.hmlint(rd)
    rh, rl, rll = M.hmlint(cr, rd)
    ro, roo = M.olprad(cr, rd)
    mn, nl = cr.maxnlo, cr.nlomax
    errs = [_rel(haa, rh), _rel(loa[:mn], rl[:mn]), _rel(lolo[:mn, :mn], rll[:mn, :mn]),
            _rel(oalo[:, :nl], ro[:, :nl]), _rel(ololo[:nl, :nl], roo[:nl, :nl])]
    print(f"{which}: hmlint/olprad max rel diff haa {errs[0]:.2e} loa {errs[1]:.2e} lolo {errs[2]:.2e} "
          f"oalo {errs[3]:.2e} ololo {errs[4]:.2e}", flush=True)
    assert max(errs) <= 1e-12


def test_setup_uses_hmlint_matrices():
    """After exc_ctx_hmlint the per-k setup uses the matrices it computed: H and O against the
    oracle setup with the oracle's hmlint/olprad matrices in place of the seeded ones."""
    from exciting_amd import lapw, synth
    from oracle import hmlint_ref as M
    from oracle import lapw_ref as R
    w = small("d")
    cr = w.crystal
    rd = synth.make_radial(cr, nr=(150, 120), lmaxvr=4)
    ctx = lapw.Context(cr)
    ctx.hmlint(rd, out=False)
    rh, rl, rll = M.hmlint(cr, rd)
    ro, roo = M.olprad(cr, rd)
    mn, nl = cr.maxnlo, cr.nlomax
    cr2 = dataclasses.replace(cr, haa=rh, loa=np.asfortranarray(rl[:mn]), lolo=np.asfortranarray(rll[:mn, :mn]),
                              oalo=np.asfortranarray(ro[:, :nl]), ololo=np.asfortranarray(roo[:nl, :nl]))
    kp = w.kpoints[0]
    apw = ctx.match(kp)
    H, O = ctx.hamiltonandoverlapsetup(kp, apw)
    H_ref, O_ref, _ = R.setup(cr2, kp)
    eh = np.linalg.norm(H - H_ref) / np.linalg.norm(H_ref)
    eo = np.linalg.norm(O - O_ref) / np.linalg.norm(O_ref)
    print(f"setup after hmlint: H rel {eh:.2e}, O rel {eo:.2e}", flush=True)
    assert eh <= 1e-12 and eo <= 1e-12


def test_match_and_setup_from_radial_data():
    """The whole muffin-tin state from one set of radial data: match's boundary values from the
    radial functions (exc_ctx_set_radial_fr, polynom at R) and the matrices from hmlint /
    olprad; apwalm, H and O against the oracle with the same derived inputs."""
    from exciting_amd import lapw, synth
    from oracle import hmlint_ref as M
    from oracle import lapw_ref as R
    w = small("d")
    cr = w.crystal
    rd = synth.make_radial(cr, nr=(150, 120), lmaxvr=4)
    ctx = lapw.Context(cr)
    assert ctx.set_radial_fr(rd) == 0
    ctx.hmlint(rd, out=False)
    rh, rl, rll = M.hmlint(cr, rd)
    ro, roo = M.olprad(cr, rd)
    mn, nl = cr.maxnlo, cr.nlomax
    cr2 = dataclasses.replace(cr, dmat=M.dmat_from_radial(cr, rd), haa=rh, loa=np.asfortranarray(rl[:mn]),
                              lolo=np.asfortranarray(rll[:mn, :mn]), oalo=np.asfortranarray(ro[:, :nl]),
                              ololo=np.asfortranarray(roo[:nl, :nl]))
    kp = w.kpoints[0]
    apw = ctx.match(kp)
    H, O = ctx.hamiltonandoverlapsetup(kp, apw)
    H_ref, O_ref, A_ref = R.setup(cr2, kp)
    n = kp.ngp
    ea = np.abs(apw[:n] - A_ref[:n]).max() / np.abs(A_ref[:n]).max()
    eh = np.linalg.norm(H - H_ref) / np.linalg.norm(H_ref)
    eo = np.linalg.norm(O - O_ref) / np.linalg.norm(O_ref)
    print(f"from radial data: apwalm {ea:.2e}, H {eh:.2e}, O {eo:.2e}", flush=True)
    assert ea <= 1e-12 and eh <= 1e-12 and eo <= 1e-12


@pytest.mark.parametrize("which", ["small-d", "workload1"])
def test_genrhomt_against_oracle(which):
    """The muffin-tin density from a density matrix (exc_genrhomt) on the radial data of the
    preceding hmlint call, against the oracle's genrhomt (pinned by the potential-energy
    identity in test_oracle_hmlint.py)."""
    from exciting_amd import lapw, synth
    from oracle import hmlint_ref as M
    from oracle import lapw_ref as R
    w = small("d") if which == "small-d" else workload(1, 1)
    cr, kp = w.crystal, w.kpoints[0]
    rd = synth.make_radial(cr, nr=(150, 120) if which == "small-d" else (260, 230),
                           lmaxvr=4 if which == "small-d" else 6)
    ctx = lapw.Context(cr)
    ctx.hmlint(rd, out=False)
    rng = np.random.default_rng(41)
    nmat, nst = kp.ngp + cr.nlotot, 20
    C = rng.standard_normal((nmat, nst)) + 1j * rng.standard_normal((nmat, nst))
    dm = R.gendmatmt(cr, R.match(cr, kp), kp.ngp, C, rng.uniform(0, 2, nst), 0.5)
    rho = ctx.genrhomt(dm, rd.lmaxvr, rd.spr.shape[0])
    ref = M.genrhomt(cr, rd, dm)
    e = np.abs(rho - ref).max() / np.abs(ref).max()
    print(f"{which}: genrhomt max rel diff {e:.2e}", flush=True)
    assert e <= 1e-12


def test_density_chain_on_device():
    """One SCF pass's device chain at workload 1 (4 k-points): grouped setup, grouped
    eigensolve, density matrix of all k-points, muffin-tin density; the density against the
    oracle's genrhomt of the density matrix the device produced."""
    import torch
    from exciting_amd import lapw, synth
    from oracle import hmlint_ref as M
    w = workload(1, 4)
    cr = w.crystal
    rd = synth.make_radial(cr, nr=(260, 230), lmaxvr=6)
    s = torch.cuda.Stream()
    ctx = lapw.Context(cr, stream=s)
    nev = 100
    with torch.cuda.stream(s):
        ctx.hmlint(rd, out=False)
        b = lapw.DeviceBatch(cr, w.kpoints)
        ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_OLP_LEGENDRE)
        evs = [torch.zeros(nev, dtype=torch.float64, device="cuda") for _ in range(b.nk)]
        vecs = [torch.zeros((nev, n), dtype=torch.complex128, device="cuda") for n in b.nmat]
        ctx.eigen_batched_dev(b.nmat, b.H, b.O, nev, evs, vecs)
        occs = [torch.full((nev,), 2.0, dtype=torch.float64, device="cuda") for _ in range(b.nk)]
        ws = ctx.dmat_wfsize()
        dm = torch.zeros((cr.natm, ws, ws), dtype=torch.complex128, device="cuda")
        ctx.gendmatmt_batched_dev(b, nev, vecs, occs, [1.0 / b.nk] * b.nk, dm)
        lmv, nrm = (rd.lmaxvr + 1) ** 2, rd.spr.shape[0]
        rho = torch.zeros((cr.natm, nrm, lmv), dtype=torch.float64, device="cuda")
        ctx._call("exc_genrhomt_dev", ctx.h, dm.data_ptr(), rho.data_ptr())
        torch.cuda.synchronize()
    dmh = np.asfortranarray(np.transpose(dm.cpu().numpy(), (2, 1, 0)))
    got = np.transpose(rho.cpu().numpy(), (2, 1, 0))
    ref = M.genrhomt(cr, rd, dmh)
    e = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"device density chain: max rel diff {e:.2e}", flush=True)
    assert e <= 1e-12
