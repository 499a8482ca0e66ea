"""Writes tests/golden/workload3_k0_probes.npz: the oracle's H and O of the first k-point of
workload 3 (32 atoms, 4 species, lmax 10, N_G+k = 5008), too large to store (2 x 420 MB),
kept as what the GPU parity test needs of them:

* probes H X and O X for a seeded complex Gaussian X (nmat x 4): for any matrix E,
  E|E X|_F^2 = 4 |E|_F^2, so |(H_gpu - H) X|_F / |H X|_F estimates the relative Frobenius
  error of H_gpu (and likewise O) without the matrices;
* |H|_F, |O|_F, and the lowest 200 eigenvalues of H c = e O c (LAPACK zhegvx, range 'I',
  upper triangles, as solvewithlapack calls it);
* the same eigenvalues of H and O summed in the reference's own order (one zgemm per atom,
  hamiltonsetup.f90:90-103, the oracle's per_atom form): how far two correct FP64 summation
  orders of the same matrices move these eigenvalues on this spectrum (its lowest eigenvalue
  is about -1.3e4 Ha), the floor of any FP64 comparison here.

The workload is rebuilt from its seed (exciting_amd.synth.make_config(3)); the oracle's setup
of one k-point takes about 2 minutes on 8 cores.  Run from the repository root:

    python tests/golden/make_workload_probes.py
"""
import os
import sys
import time

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from exciting_amd import synth  # noqa: E402
from oracle import lapw_ref as R  # noqa: E402

NPROBE = 4
NEV = 200


def probes(nmat, seed=20261203):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nmat, NPROBE)) + 1j * rng.standard_normal((nmat, NPROBE))


def main():
    w = synth.make_config(3, nk=1)
    cr, kp = w.crystal, w.kpoints[0]
    t0 = time.time()
    H, O, _ = R.setup(cr, kp)
    print(f"oracle setup N={kp.ngp} nmat={H.shape[0]}: {time.time() - t0:.1f} s", flush=True)
    X = probes(H.shape[0])
    kw = dict(lower=False, subset_by_index=[0, NEV - 1], driver="gvx", eigvals_only=True)
    e = scipy.linalg.eigh(H, O, **kw)
    H1, O1, _ = R.setup(cr, kp, per_atom=True)
    e1 = scipy.linalg.eigh(H1, O1, **kw)
    print(f"per-atom order vs l-block order: max |de| {np.abs(e1 - e).max():.2e} Ha", flush=True)
    out = os.path.join(HERE, "workload3_k0_probes.npz")
    np.savez_compressed(out, ngp=kp.ngp, hx=H @ X, ox=O @ X, hnorm=np.linalg.norm(H), onorm=np.linalg.norm(O),
                        eig=e, eig_per_atom=e1)
    print("wrote", out)


if __name__ == "__main__":
    main()
