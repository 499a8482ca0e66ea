"""ValenceRelativity /= 'none' in the interstitial (exc_ctx_set_meffig, hmlistln.f90:61-72): H's
kinetic term takes meffig(G_i - G_j) instead of Theta, O keeps Theta.

* per-k FP64 entry, grouped FP64 launch (graph replay, overlap by the addition theorem) and
  the modular split, all against the oracle with the same meffig: FP64 H, O <= 1e-12
  (relative Frobenius), the split's H within 1e-9 (its own scheme's error, as at 'none');
* meffig = cfunig reproduces the 'none' result bit for bit; clearing it restores 'none';
* the adjoint refuses (it differentiates the Theta form).
The meffig used is seeded and Hermitian (meffig(-G) = conj(meffig(G))): cfunig times
1 + 0.05 cos(1.3 |G|).
"""
import numpy as np
import pytest

from conftest import small, workload

pytestmark = pytest.mark.gpu


def _meffig(cr):
    g = cr.gc[:cr.ngvec]
    return cr.cfunig[:cr.ngvec] * (1.0 + 0.05 * np.cos(1.3 * g))


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_meffig_per_k_and_grouped_against_oracle():
    import torch
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 2)
    cr = w.crystal
    m = _meffig(cr)
    s = torch.cuda.Stream()
    ctx = lapw.Context(cr, stream=s)
    ctx.set_meffig(m)
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(cr, w.kpoints)
        for rep in range(2):                       # eager, then a graph capture and replay
            for t in b.H + b.O:
                t.fill_(float("nan"))
            flags = lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_OLP_LEGENDRE | (lapw.EXC_BATCH_GRAPH if rep else 0)
            ctx.setup_batched_dev(b, flags)
            torch.cuda.synchronize()
            for i, kp in enumerate(w.kpoints):
                H_ref, O_ref, _ = R.setup(cr, kp, meffig=m)
                H = b.H[i].cpu().numpy().T
                O = b.O[i].cpu().numpy().T
                eh, eo = _rel(H, H_ref), _rel(O, O_ref)
                print(f"grouped rep {rep} k{i}: |dH|/|H| {eh:.2e} |dO|/|O| {eo:.2e}", flush=True)
                assert eh <= 1e-12 and eo <= 1e-12
    # the per-k host entry
    kp = w.kpoints[0]
    H_ref, O_ref, A = R.setup(cr, kp, meffig=m)
    H, O = ctx.hamiltonandoverlapsetup(kp, A)
    print(f"per-k: |dH|/|H| {_rel(H, H_ref):.2e}", flush=True)
    assert _rel(H, H_ref) <= 1e-12 and _rel(O, O_ref) <= 1e-12
    # the difference from 'none' is real: the same run without meffig differs
    H0_ref, _, _ = R.setup(cr, kp)
    assert _rel(H0_ref, H_ref) > 1e-6
    ctx.close()


def test_meffig_equal_to_theta_and_cleared():
    import torch
    from exciting_amd import lapw
    w = workload(1, 2)
    cr = w.crystal
    s = torch.cuda.Stream()
    ctx = lapw.Context(cr, stream=s)
    flags = lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_OLP_LEGENDRE | lapw.EXC_BATCH_GRAPH
    out = []
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(cr, w.kpoints)
        for m in (None, cr.cfunig[:cr.ngvec].copy(), _meffig(cr), None):
            ctx.set_meffig(m)
            ctx.setup_batched_dev(b, flags)
            torch.cuda.synchronize()
            out.append([h.cpu().clone() for h in b.H])
    for i in range(len(w.kpoints)):
        assert torch.equal(out[1][i], out[0][i])       # meffig = Theta: the 'none' bits
        assert torch.equal(out[3][i], out[0][i])       # cleared: back to 'none' (graph recaptured)
        assert not torch.equal(out[2][i], out[0][i])
    ctx.close()


def test_meffig_modular_split():
    import torch
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 2)
    cr = w.crystal
    m = _meffig(cr)
    s = torch.cuda.Stream()
    ctx = lapw.Context(cr, stream=s)
    ctx.set_meffig(m)
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(cr, w.kpoints)
        ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_OLP_LEGENDRE | lapw.EXC_BATCH_SPLIT_CRT)
        torch.cuda.synchronize()
    for i, kp in enumerate(w.kpoints):
        H_ref, O_ref, _ = R.setup(cr, kp, meffig=m)
        eh = _rel(b.H[i].cpu().numpy().T, H_ref)
        eo = _rel(b.O[i].cpu().numpy().T, O_ref)
        print(f"modular split k{i}: |dH|/|H| {eh:.2e} |dO|/|O| {eo:.2e}", flush=True)
        assert eh <= 1e-9 and eo <= 1e-12
    ctx.close()


def test_meffig_refused_by_adjoint():
    from exciting_amd import lapw
    w = small("a")
    cr = w.crystal
    ctx = lapw.Context(cr)
    ctx.set_meffig(_meffig(cr))
    # refused before any argument is read
    with pytest.raises(lapw.ExcError) as e:
        ctx._call("exc_ho_vjp_dev", ctx.h, 1, 1, None, None, None, None, 1, 1, None, None, None, None, None)
    assert e.value.code == lapw.EXC_ERR_UNSUPPORTED
    ctx.set_meffig(None)
    with pytest.raises(lapw.ExcError) as e:
        ctx._call("exc_ho_vjp_dev", ctx.h, 1, 1, None, None, None, None, 1, 1, None, None, None, None, None)
    assert e.value.code == lapw.EXC_ERR_ARG           # the 'none' form: the usual argument check
    ctx.close()
