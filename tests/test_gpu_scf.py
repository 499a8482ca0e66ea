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


@pytest.mark.parametrize("which", ["small-d", "config1"])
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
