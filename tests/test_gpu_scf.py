"""Theta(G), Theta(r) and V(G) on the device (exc_ctx_gencfun, exc_ctx_genveffig: hipFFT on the
G box) against the oracle (gencfun / genveffig, pinned in test_oracle_scf.py), and the setup
reading the installed tables.

Bars: Theta and V within 1e-12 of the oracle (absolute, values of order 1); H, O relative
Frobenius 1e-12."""
import dataclasses

import numpy as np
import pytest

from conftest import small, workload

pytestmark = pytest.mark.gpu


def _veffir(cr, seed):
    """A real potential on the grid: smooth random plane-wave content plus a constant."""
    from oracle import lapw_ref as R
    rng = np.random.default_rng(seed)
    z = np.zeros(cr.ngrtot, dtype=np.complex128)
    low = np.nonzero(cr.gc < 0.6 * cr.gmaxvr)[0]
    z[R.igfft(cr)[low]] = rng.standard_normal(len(low)) + 1j * rng.standard_normal(len(low))
    v = R.dft3(z, R.fft_grid(cr), 1).real if cr.ngrtot < 5000 else R._fft3(z, R.fft_grid(cr), 1).real
    return v / np.abs(v).max() - 0.5


@pytest.mark.parametrize("which", ["small-d", "workload1"])
def test_gencfun_genveffig_against_oracle(which):
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = small("d") if which == "small-d" else workload(1, 1)
    cr = w.crystal
    ctx = lapw.Context(cr)
    cf, cir = ctx.gencfun()
    cf_ref, cir_ref = R.gencfun(cr)
    v = _veffir(cr, 3)
    vg = ctx.genveffig(v)
    vg_ref = R.genveffig(cr, v, cir_ref)
    e = [np.abs(cf - cf_ref).max(), np.abs(cir - cir_ref).max(), np.abs(vg - vg_ref).max()]
    print(f"{which}: grid {R.fft_grid(cr)}: Theta(G) {e[0]:.2e}, Theta(r) {e[1]:.2e}, V(G) {e[2]:.2e}", flush=True)
    assert max(e) <= 1e-12


def test_gencfun_damped_against_oracle():
    """exc_ctx_set_cfdamp: the damped Theta(G) within 1e-12 of the oracle's, Theta(r) the
    undamped one, the setup reads the damped table; cfdamp 0 restores the undamped values."""
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 1)
    cr = w.crystal
    ctx = lapw.Context(cr)
    ctx.set_cfdamp(0.6, cr.gmaxvr)
    cf, cir = ctx.gencfun()
    cf_ref, cir_ref = R.gencfun(cr, cfdamp=0.6, gmaxvr=cr.gmaxvr)
    e = [np.abs(cf - cf_ref).max(), np.abs(cir - cir_ref).max()]
    print(f"damped: Theta(G) {e[0]:.2e}, Theta(r) {e[1]:.2e}", flush=True)
    assert max(e) <= 1e-12
    cr2 = dataclasses.replace(cr, cfunig=cf_ref)
    kp = w.kpoints[0]
    H_ref, O_ref, A = R.setup(cr2, kp)
    H, O = ctx.hamiltonandoverlapsetup(kp, A)
    eh = np.linalg.norm(H - H_ref) / np.linalg.norm(H_ref)
    eo = np.linalg.norm(O - O_ref) / np.linalg.norm(O_ref)
    print(f"setup on the damped Theta: H {eh:.2e}, O {eo:.2e}", flush=True)
    assert eh <= 1e-12 and eo <= 1e-12
    ctx.set_cfdamp(0.0, cr.gmaxvr)
    cf0, _ = ctx.gencfun()
    assert np.abs(cf0 - R.gencfun(cr)[0]).max() <= 1e-12
    with pytest.raises(lapw.ExcError):
        ctx.set_cfdamp(-1.0, cr.gmaxvr)
    ctx.close()


def test_setup_reads_the_new_tables():
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = small("d")
    cr = w.crystal
    ctx = lapw.Context(cr)
    ctx.gencfun()
    v = _veffir(cr, 5)
    ctx.genveffig(v)
    cf_ref, cir_ref = R.gencfun(cr)
    cr2 = dataclasses.replace(cr, cfunig=cf_ref, veffig=R.genveffig(cr, v, cir_ref))
    kp = w.kpoints[0]
    apw = ctx.match(kp)
    H, O = ctx.hamiltonandoverlapsetup(kp, apw)
    H_ref, O_ref, _ = R.setup(cr2, kp)
    eh = np.linalg.norm(H - H_ref) / np.linalg.norm(H_ref)
    eo = np.linalg.norm(O - O_ref) / np.linalg.norm(O_ref)
    print(f"setup after gencfun/genveffig: H {eh:.2e}, O {eo:.2e}", flush=True)
    assert eh <= 1e-12 and eo <= 1e-12


@pytest.mark.parametrize("which", ["small-d", "workload1"])
def test_genrhoir_against_oracle(which):
    """The interstitial density of one k-point (exc_genrhoir, batched FFTs of 64 states) against
    the oracle's genrhoir (pinned by Parseval and direct sums in test_oracle_scf.py); 70 states
    at workload 1 cross the chunk boundary; a zero and a below-epsocc occupation are skipped."""
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = small("d") if which == "small-d" else workload(1, 1)
    cr, kp = w.crystal, w.kpoints[0]
    ctx = lapw.Context(cr)
    ctx.gencfun()
    rng = np.random.default_rng(13)
    nmat, nst = kp.ngp + cr.nlotot, (9 if which == "small-d" else 70)
    C = rng.standard_normal((nmat, nst)) + 1j * rng.standard_normal((nmat, nst))
    occ = rng.uniform(0.0, 2.0, nst)
    occ[1], occ[2] = 0.0, 1e-12
    rho0 = rng.uniform(0, 1, cr.ngrtot)
    got = ctx.genrhoir(kp, C, occ, 0.35, rhoir=rho0.copy())
    ref = R.genrhoir(cr, kp, C, occ, 0.35, rhoir=rho0.copy())
    e = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"{which}: genrhoir max rel diff {e:.2e}", flush=True)
    assert e <= 1e-12


def test_genrhoir_batched_after_eigensolve():
    import torch
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 4)
    cr = w.crystal
    s = torch.cuda.Stream()
    ctx = lapw.Context(cr, stream=s)
    nev = 100
    rng = np.random.default_rng(17)
    occs_h = [rng.uniform(0.0, 2.0, nev) for _ in w.kpoints]
    wk = list(rng.uniform(0.1, 0.4, len(w.kpoints)))
    with torch.cuda.stream(s):
        ctx.gencfun()
        b = lapw.DeviceBatch(cr, w.kpoints)
        ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_OLP_LEGENDRE)
        evs = [torch.zeros(nev, dtype=torch.float64, device="cuda") for _ in range(b.nk)]
        vecs = [torch.zeros((nev, n), dtype=torch.complex128, device="cuda") for n in b.nmat]
        ctx.eigen_batched_dev(b.nmat, b.H, b.O, nev, evs, vecs)
        occs = [torch.from_numpy(o).cuda() for o in occs_h]
        rho = torch.zeros(cr.ngrtot, dtype=torch.float64, device="cuda")
        ctx.genrhoir_batched_dev(b, nev, vecs, occs, wk, rho)
        torch.cuda.synchronize()
    ref = None
    for k, kp in enumerate(w.kpoints):
        ref = R.genrhoir(cr, kp, np.asfortranarray(vecs[k].cpu().numpy().T), occs_h[k], wk[k], rhoir=ref)
    got = rho.cpu().numpy()
    e = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"genrhoir batched over {b.nk} k-points: max rel diff {e:.2e}", flush=True)
    assert e <= 1e-12
