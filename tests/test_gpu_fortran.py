"""The Fortran driver (amdflang, iso_c_binding) with the reference signatures, on the GPU.

Runs exciting_amd/_lib/exc_driver_main: match(ngp, gpc, tpgpc, sfacgp, apwalm) followed by
hamiltonsetup(system, ngp, apwalm, igpig, vgpc) and overlapsetup(...) as seceqnfv.f90:76-77
calls them (apwalm taken from the library's device copy of match's result, or copied from the
host again), or by the fused hamiltonandoverlapsetup, for one k-point, and checks apwalm, H
and O against the oracle (relative Frobenius error <= 1e-12); the three call sequences run the
same kernels and must give the same matrices bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import small, workload
from fortran_io import (read_dmat, read_eigen_output, read_forces, read_output, read_rhoir, read_rhomt,
                        write_eigen, write_input, write_radial, write_scf)

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "exciting_amd", "_lib",
                   "exc_driver_main")


@pytest.mark.parametrize("which", ["small-b", "workload0"])
def test_fortran_driver(tmp_path, which):
    from oracle import lapw_ref as R
    if not os.path.exists(EXE):
        pytest.fail("Fortran driver not built: run python -m exciting_amd.build")
    w = small("b") if which == "small-b" else workload(0)
    cr, kp = w.crystal, w.kpoints[0]
    fin = tmp_path / "in.bin"
    write_input(fin, cr, kp)
    H_ref, O_ref, A_ref = R.setup(cr, kp)
    n = kp.ngp
    res = {}
    for mode in ("split", "copy", "fused"):
        fout = tmp_path / f"out_{mode}.bin"
        print(f"running Fortran driver on {which}, setup calls {mode}", flush=True)
        r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120,
                           env=dict(os.environ, EXC_DRIVER_SETUP=mode))
        print(r.stdout, r.stderr)
        assert r.returncode == 0
        apw, H, O = read_output(fout, cr, kp.igpig.shape[0])
        assert np.abs(apw[:n] - A_ref[:n]).max() <= 1e-12 * np.abs(A_ref[:n]).max()
        assert np.linalg.norm(H - H_ref) <= 1e-12 * np.linalg.norm(H_ref)
        assert np.linalg.norm(O - O_ref) <= 1e-12 * np.linalg.norm(O_ref)
        res[mode] = (H, O)
    for mode in ("copy", "fused"):
        assert np.array_equal(res[mode][0], res["split"][0]) and np.array_equal(res[mode][1], res["split"][1]), mode


def test_fortran_forcek(tmp_path):
    """forcek_gpu (Fortran, through exc_ho_vjp) against the restated reference forcek, and
    gendmatmt_gpu (through exc_gendmatmt) against the restated gendmatmt."""
    from oracle import lapw_ref as R
    w = small("b")
    cr, kp = w.crystal, w.kpoints[0]
    H, O, A = R.setup(cr, kp)
    nst = 10
    e, C = R.eigen(H, O, nst)
    occ = np.random.default_rng(2).uniform(0.0, 2.0, nst)
    wk = 0.25
    fin, fout, feig = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "eig.bin"
    write_input(fin, cr, kp)
    write_eigen(feig, C, e, occ, wk)
    r = subprocess.run([EXE, str(fin), str(fout), str(feig)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    F = read_forces(fout, cr, kp.igpig.shape[0])
    F_ref = R.forcek_ibs(cr, kp, C, e, occ, apwalm=A, wkpt=wk)
    print("forceibs max |diff|", np.abs(F - F_ref).max(), "max |F|", np.abs(F_ref).max())
    assert np.abs(F - F_ref).max() <= 1e-9 * max(1.0, np.abs(F_ref).max())
    # gendmatmt_gpu of the same states, on the driver's own apwalm
    apw, _, _ = read_output(fout, cr, kp.igpig.shape[0])
    dm = read_dmat(fout, cr, kp.igpig.shape[0], R.dmat_wfsize(cr))
    dm_ref = R.gendmatmt(cr, apw, kp.ngp, C, occ, wk)
    rel = np.linalg.norm((dm - dm_ref).ravel()) / np.linalg.norm(dm_ref.ravel())
    print("gendmatmt_gpu rel diff", rel)
    assert rel <= 1e-12


@pytest.mark.parametrize("which", ["small-b", "workload0"])
def test_fortran_solvewithgpu(tmp_path, which):
    """solvewithgpu (Fortran, through exc_eigen) after the setup: eigenvalues against the
    oracle's LAPACK solve of the oracle's H, O within 1e-10 Ha, residual and O-orthonormality."""
    from oracle import lapw_ref as R
    w = small("b") if which == "small-b" else workload(0)
    cr, kp = w.crystal, w.kpoints[0]
    nev = 20
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, cr, kp)
    r = subprocess.run([EXE, str(fin), str(fout), "-", str(nev)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    e, C = read_eigen_output(fout, cr, kp.igpig.shape[0], nev)
    H, O, _ = R.setup(cr, kp)
    e_ref, _ = R.eigen(H, O, nev)
    res = np.linalg.norm(H @ C - (O @ C) * e[None, :], axis=0).max() / (np.linalg.norm(H, 2) + np.abs(e).max() * np.linalg.norm(O, 2))
    orth = np.abs(C.conj().T @ O @ C - np.eye(nev)).max()
    print(f"{which}: max |d eig| {np.abs(e - e_ref).max():.2e} Ha, residual {res:.2e}, orthonormality {orth:.2e}")
    assert np.abs(e - e_ref).max() <= 1e-10
    assert res <= 1e-12 and orth <= 1e-10


def test_fortran_hmlint_then_setup(tmp_path):
    """set_radial_gpu and hmlint_gpu (Fortran, through exc_ctx_set_radial_fr / exc_ctx_hmlint)
    before the setup: H and O against the oracle setup with the oracle's boundary values and
    hmlint/olprad matrices in place of the seeded ones."""
    import dataclasses
    from exciting_amd import synth
    from oracle import hmlint_ref as M
    from oracle import lapw_ref as R
    w = small("b")
    cr, kp = w.crystal, w.kpoints[0]
    rd = synth.make_radial(cr, nr=(140, 110), lmaxvr=4)
    fin, fout, frad = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "radial.bin"
    write_input(fin, cr, kp)
    write_radial(frad, cr, rd)
    r = subprocess.run([EXE, str(fin), str(fout), "-", "0", str(frad)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    _, H, O = read_output(fout, cr, kp.igpig.shape[0])
    rh, rl, rll = M.hmlint(cr, rd)
    ro, roo = M.olprad(cr, rd)
    mn, nl = cr.maxnlo, cr.nlomax
    cr2 = dataclasses.replace(cr, dmat=M.dmat_from_radial(cr, rd), haa=rh, loa=np.asfortranarray(rl[:mn]),
                              lolo=np.asfortranarray(rll[:mn, :mn]), oalo=np.asfortranarray(ro[:, :nl]),
                              ololo=np.asfortranarray(roo[:nl, :nl]))
    H_ref, O_ref, _ = R.setup(cr2, kp)
    eh = np.linalg.norm(H - H_ref) / np.linalg.norm(H_ref)
    eo = np.linalg.norm(O - O_ref) / np.linalg.norm(O_ref)
    print(f"Fortran set_radial_gpu + hmlint_gpu + setup: H rel {eh:.2e}, O rel {eo:.2e}")
    assert eh <= 1e-12 and eo <= 1e-12


def test_fortran_gencfun_genveffig_then_setup(tmp_path):
    """gencfun_gpu and genveffig_gpu (Fortran) before the setup: H and O against the oracle
    setup with the oracle's Theta(G) and V(G)."""
    import dataclasses
    from oracle import lapw_ref as R
    from test_gpu_scf import _veffir
    w = small("b")
    cr, kp = w.crystal, w.kpoints[0]
    v = _veffir(cr, 7)
    fin, fout, fscf = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "scf.bin"
    write_input(fin, cr, kp)
    write_scf(fscf, cr, v)
    r = subprocess.run([EXE, str(fin), str(fout), "-", "0", "-", str(fscf)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    _, H, O = read_output(fout, cr, kp.igpig.shape[0])
    cf, cir = R.gencfun(cr)
    cr2 = dataclasses.replace(cr, cfunig=cf, veffig=R.genveffig(cr, v, cir))
    H_ref, O_ref, _ = R.setup(cr2, kp)
    eh = np.linalg.norm(H - H_ref) / np.linalg.norm(H_ref)
    eo = np.linalg.norm(O - O_ref) / np.linalg.norm(O_ref)
    print(f"Fortran gencfun_gpu + genveffig_gpu + setup: H rel {eh:.2e}, O rel {eo:.2e}")
    assert eh <= 1e-12 and eo <= 1e-12


def test_fortran_gendmatmt_genrhomt(tmp_path):
    """gendmatmt_gpu then genrhomt_gpu (Fortran) on radial data: the density against the
    oracle's genrhomt of the density matrix the driver wrote; genrhoir_gpu's interstitial
    density against the oracle's genrhoir."""
    from exciting_amd import synth
    from oracle import hmlint_ref as M
    from oracle import lapw_ref as R
    w = small("b")
    cr, kp = w.crystal, w.kpoints[0]
    rd = synth.make_radial(cr, nr=(140, 110), lmaxvr=4)
    nmat = kp.ngp + cr.nlotot
    rng = np.random.default_rng(12)
    nst = 8
    C = rng.standard_normal((nmat, nst)) + 1j * rng.standard_normal((nmat, nst))
    e, occ = np.sort(rng.standard_normal(nst)), rng.uniform(0, 2, nst)
    fin, fout, feig, frad = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "eig.bin", tmp_path / "rad.bin"
    write_input(fin, cr, kp)
    write_eigen(feig, C, e, occ, 0.3)
    write_radial(frad, cr, rd)
    r = subprocess.run([EXE, str(fin), str(fout), str(feig), "0", str(frad)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    ngkmax, ws = kp.igpig.shape[0], R.dmat_wfsize(cr)
    dm = read_dmat(fout, cr, ngkmax, ws)
    rho = read_rhomt(fout, cr, ngkmax, ws, (rd.lmaxvr + 1) ** 2, rd.spr.shape[0])
    ref = M.genrhomt(cr, rd, dm)
    rel = np.abs(rho - ref).max() / np.abs(ref).max()
    print("Fortran genrhomt_gpu rel diff", rel)
    assert rel <= 1e-12
    rir = read_rhoir(fout, cr, ngkmax, ws)
    rir_ref = R.genrhoir(cr, kp, C, occ, 0.3)
    rel2 = np.abs(rir - rir_ref).max() / np.abs(rir_ref).max()
    print("Fortran genrhoir_gpu rel diff", rel2)
    assert rel2 <= 1e-12


def test_fortran_rhovalk(tmp_path):
    """rhovalk_gpu (Fortran) with lmaxvr 4 <= lmaxapw: the direct-route muffin-tin density of the
    driver's states against the oracle's rhovalk on the apwalm the driver wrote."""
    from exciting_amd import synth
    from oracle import lapw_ref as R
    from oracle import rhovalk_ref as V
    w = small("d")
    cr, kp = w.crystal, w.kpoints[0]
    rd = synth.make_radial(cr, nr=(120, 100), lmaxvr=4)
    nmat = kp.ngp + cr.nlotot
    rng = np.random.default_rng(14)
    nst = 6
    C = rng.standard_normal((nmat, nst)) + 1j * rng.standard_normal((nmat, nst))
    e, occ = np.sort(rng.standard_normal(nst)), rng.uniform(0.1, 2, nst)
    fin, fout, feig, frad = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "eig.bin", tmp_path / "rad.bin"
    write_input(fin, cr, kp)
    write_eigen(feig, C, e, occ, 0.3)
    write_radial(frad, cr, rd)
    r = subprocess.run([EXE, str(fin), str(fout), str(feig), "0", str(frad)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    ngkmax, ws = kp.igpig.shape[0], R.dmat_wfsize(cr)
    apw, _, _ = read_output(fout, cr, ngkmax)
    rho = read_rhomt(fout, cr, ngkmax, ws, (rd.lmaxvr + 1) ** 2, rd.spr.shape[0], direct=True)
    ref = V.rhovalk(cr, rd, apw, kp.ngp, C, occ, 0.3)
    rel = np.abs(rho - ref).max() / np.abs(ref).max()
    print("Fortran rhovalk_gpu rel diff", rel)
    assert rel <= 1e-12
