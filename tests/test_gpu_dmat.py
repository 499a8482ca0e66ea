"""Muffin-tin density matrix on the device (exc_gendmatmt / exc_gendmatmt_dev /
exc_gendmatmt_batched_dev) against the oracle's gendmatmt (gendmatmt.f90:59-172), whose row
order and expectation-value identities are pinned on the CPU (test_oracle_dmat.py).

Bar: relative Frobenius difference <= 1e-12 of the accumulated matrices (FP64, different
summation order only)."""
import numpy as np
import pytest

from conftest import small, workload

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def _inputs(cr, kp, nst, seed):
    rng = np.random.default_rng(seed)
    nmat = kp.ngp + cr.nlotot
    C = np.asfortranarray(rng.standard_normal((nmat, nst)) + 1j * rng.standard_normal((nmat, nst)))
    occ = rng.uniform(0.0, 2.0, nst)
    return C, occ


@pytest.mark.parametrize("which", ["small-d", "workload1"])
def test_gendmatmt_host_against_oracle(which):
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = small("d") if which == "small-d" else workload(1, 1)
    cr, kp = w.crystal, w.kpoints[0]
    ctx = lapw.Context(cr)
    apw = np.asfortranarray(R.match(cr, kp))
    nst = 40 if which == "small-d" else 200
    C, occ = _inputs(cr, kp, nst, 3)
    ws = ctx.dmat_wfsize()
    assert ws == R.dmat_wfsize(cr)
    rng = np.random.default_rng(4)
    dm0 = np.asfortranarray(rng.standard_normal((ws, ws, cr.natm)) + 0j)   # accumulated into
    ref = R.gendmatmt(cr, apw, kp.ngp, C, occ, 0.25, dm=dm0.copy(order="F"))
    got = ctx.gendmatmt(apw, kp.ngp, C, occ, 0.25, dm=dm0.copy(order="F"))
    r = _rel(got, ref)
    print(f"{which}: gendmatmt host entry, rel diff {r:.2e}", flush=True)
    assert r <= 1e-12


def test_gendmatmt_second_variational():
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = small("d")
    cr, kp = w.crystal, w.kpoints[0]
    ctx = lapw.Context(cr)
    apw = np.asfortranarray(R.match(cr, kp))
    nst = 24
    C, occ = _inputs(cr, kp, nst, 8)
    rng = np.random.default_rng(9)
    sv = np.asfortranarray(rng.standard_normal((nst, nst)) + 1j * rng.standard_normal((nst, nst)))
    ref = R.gendmatmt(cr, apw, kp.ngp, C, occ, 0.5, evecsv=sv)
    got = ctx.gendmatmt(apw, kp.ngp, C, occ, 0.5, evecsv=sv)
    r = _rel(got, ref)
    print(f"gendmatmt with evecsv: rel diff {r:.2e}", flush=True)
    assert r <= 1e-12


def test_gendmatmt_batched_after_setup_and_eigensolve():
    """The device pipeline of one SCF pass over a batch: grouped setup, grouped eigensolve,
    then the density matrix of all k-points at once; against the oracle on the same
    device-built apwalm and eigenvectors, and against the per-k device entry."""
    import torch
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 4)
    cr = w.crystal
    s = torch.cuda.Stream()
    ctx = lapw.Context(cr, stream=s)
    nev = 120
    rng = np.random.default_rng(21)
    occs_h = [rng.uniform(0.0, 2.0, nev) for _ in w.kpoints]
    wk = list(rng.uniform(0.05, 0.3, len(w.kpoints)))
    ws = ctx.dmat_wfsize()
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(cr, w.kpoints)
        ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_OLP_LEGENDRE)
        evs = [torch.zeros(nev, dtype=torch.float64, device="cuda") for _ in range(b.nk)]
        vecs = [torch.zeros((nev, n), dtype=torch.complex128, device="cuda") for n in b.nmat]
        ctx.eigen_batched_dev(b.nmat, b.H, b.O, nev, evs, vecs)
        occs = [torch.from_numpy(o).cuda() for o in occs_h]
        dm = torch.zeros((cr.natm, ws, ws), dtype=torch.complex128, device="cuda")
        ctx.gendmatmt_batched_dev(b, nev, vecs, occs, wk, dm)
        dm1 = torch.zeros_like(dm)
        for k, d in enumerate(b.dks):
            ctx.gendmatmt_dev(d.ngp, d.ngkmax, b.apw[k], nev, b.nmat[k], vecs[k], occs[k], wk[k], dm1)
        torch.cuda.synchronize()
    ref = None
    for k, kp in enumerate(w.kpoints):
        apw = np.asfortranarray(np.transpose(b.apw[k].cpu().numpy(), (3, 2, 1, 0)))
        C = np.asfortranarray(vecs[k].cpu().numpy().T)
        ref = R.gendmatmt(cr, apw, kp.ngp, C, occs_h[k], wk[k], dm=ref)
    got = np.asfortranarray(np.transpose(dm.cpu().numpy(), (2, 1, 0)))
    got1 = np.asfortranarray(np.transpose(dm1.cpu().numpy(), (2, 1, 0)))
    r, r1 = _rel(got, ref), _rel(got1, ref)
    print(f"gendmatmt batched over {b.nk} k-points: rel diff {r:.2e} (per-k entry {r1:.2e})", flush=True)
    assert r <= 1e-12 and r1 <= 1e-12
