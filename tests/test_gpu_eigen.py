"""Generalized eigensolve on the device (exc_eigen / exc_eigen_dev, rocSOLVER zhegvdx), the
counterpart of solvewithlapack's ZHEGVX(1, 'V', 'I', 'U', 1..nev) (modfvsystem.f90:691-693).

Against the oracle's LAPACK solve of the same H, O (oracle/lapw_ref.eigen: zhegvx subset):
eigenvalues within 1e-10 Ha; eigenvectors checked by what fixes them up to phase —
relative residual ||H c - e O c|| / (||H|| + |e| ||O||) <= 1e-12 and O-orthonormality
||C^H O C - I||_max <= 1e-10.  Then the whole path on the device (grouped setup followed by
exc_eigen_dev on the same HBM buffers), and LAPACK's info convention for an O that is not
positive definite."""
import numpy as np
import pytest

from conftest import small, workload

pytestmark = pytest.mark.gpu


def _check(H_ref, O_ref, e, C, e_ref, tag):
    de = np.abs(e - e_ref).max()
    R = H_ref @ C - (O_ref @ C) * e[None, :]
    scale = np.linalg.norm(H_ref, 2) + np.abs(e).max() * np.linalg.norm(O_ref, 2)
    res = np.linalg.norm(R, axis=0).max() / scale
    orth = np.abs(C.conj().T @ O_ref @ C - np.eye(C.shape[1])).max()
    print(f"{tag}: max |d eig| {de:.2e} Ha, residual {res:.2e}, orthonormality {orth:.2e}", flush=True)
    assert de <= 1e-10
    assert res <= 1e-12
    assert orth <= 1e-10


@pytest.mark.parametrize("which", ["small-d", "config1"])
def test_eigen_host_against_lapack(which):
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = small("d") if which == "small-d" else workload(1, 1)
    ctx = lapw.Context(w.crystal)
    kp = w.kpoints[0]
    H_ref, O_ref, _ = R.setup(w.crystal, kp)
    nev = min(200, H_ref.shape[0])
    e_ref, _ = R.eigen(H_ref, O_ref, nev)
    e, C = ctx.eigen(np.asfortranarray(H_ref.copy()), np.asfortranarray(O_ref.copy()), nev)
    _check(H_ref, O_ref, e, C, e_ref, f"{which} N={kp.ngp} host")


def test_eigen_on_device_after_grouped_setup():
    import torch
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 2)
    s = torch.cuda.Stream()
    ctx = lapw.Context(w.crystal, stream=s)
    nev = 200
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(w.crystal, w.kpoints)
        ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_OLP_LEGENDRE)
        out = []
        for i in range(b.nk):
            n = b.nmat[i]
            ev = torch.zeros(nev, dtype=torch.float64, device="cuda")
            vec = torch.zeros((nev, n), dtype=torch.complex128, device="cuda")
            ctx.eigen_dev(n, n, b.H[i], b.O[i], nev, ev, vec, n)
            out.append((ev, vec))
        torch.cuda.synchronize()
    for i, kp in enumerate(w.kpoints):
        H_ref, O_ref, _ = R.setup(w.crystal, kp)
        e_ref, _ = R.eigen(H_ref, O_ref, nev)
        e = out[i][0].cpu().numpy()
        C = out[i][1].cpu().numpy().T
        _check(H_ref, O_ref, e, C, e_ref, f"config1 k{i} N={kp.ngp} device pipeline")


def test_eigen_reports_indefinite_overlap():
    from exciting_amd import lapw
    w = small("a")
    ctx = lapw.Context(w.crystal)
    n = 12
    H = np.asfortranarray(np.eye(n, dtype=np.complex128))
    O = np.asfortranarray(-np.eye(n, dtype=np.complex128))
    with pytest.raises(lapw.ExcError) as ei:
        ctx.eigen(H, O, 3)
    assert ei.value.code == n + 1          # leading minor of order 1 not positive definite


@pytest.mark.parametrize("variant", range(5))
def test_eigen_solver_variants(variant):
    """Every solver path (0 zhegvdx, 1 zhegvx, 2 zhegvd, 3 zhegv, 4 level-3 reduction +
    zheevd) to the same bars."""
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 1)
    ctx = lapw.Context(w.crystal)
    ctx.set_variant("eigen", variant)
    kp = w.kpoints[0]
    H_ref, O_ref, _ = R.setup(w.crystal, kp)
    nev = 200
    e_ref, _ = R.eigen(H_ref, O_ref, nev)
    e, C = ctx.eigen(np.asfortranarray(H_ref.copy()), np.asfortranarray(O_ref.copy()), nev)
    _check(H_ref, O_ref, e, C, e_ref, f"solver variant {variant}")


def test_eigen_batched_equals_single():
    """The multi-k entry (concurrent worker streams) gives the single-k results bit for bit."""
    import torch
    from exciting_amd import lapw
    w = workload(1, 6)
    s = torch.cuda.Stream()
    ctx = lapw.Context(w.crystal, stream=s)
    nev = 200
    flags = lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_OLP_LEGENDRE
    res = {}
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(w.crystal, w.kpoints)
        for mode in ("single", "batched"):
            ctx.setup_batched_dev(b, flags)
            evs = [torch.zeros(nev, dtype=torch.float64, device="cuda") for _ in range(b.nk)]
            vecs = [torch.zeros((nev, n), dtype=torch.complex128, device="cuda") for n in b.nmat]
            if mode == "single":
                for i in range(b.nk):
                    ctx.eigen_dev(b.nmat[i], b.nmat[i], b.H[i], b.O[i], nev, evs[i], vecs[i], b.nmat[i])
            else:
                ctx.eigen_batched_dev(b.nmat, b.H, b.O, nev, evs, vecs)
            torch.cuda.synchronize()
            res[mode] = ([e.cpu() for e in evs], [v.cpu() for v in vecs])
            print(f"eigen {mode}: {b.nk} k-points", flush=True)
    for i in range(len(w.kpoints)):
        assert torch.equal(res["single"][0][i], res["batched"][0][i])
        assert torch.equal(res["single"][1][i], res["batched"][1][i])
