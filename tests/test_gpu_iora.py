"""ValenceRelativity 'iora*' overlap terms (exc_ctx_set_iora; overlapsetup.f90:81-94, 126, 144,
olpistln.f90:53-62) against the oracle, whose terms are pinned in test_oracle_setup.py:

* per-k entries (host fused, host O only, device fused): H and O within 1e-12 (relative
  Frobenius) of the oracle's with the same seeded h1aa, h1loa, h1lolo, m2effig; the device
  entry equals the host entry bit for bit;
* grouped launch with the addition-theorem overlap (eager and graph replay) and the modular
  split: O within 1e-12, H within 1e-12 (FP64) or 1e-9 (the split's own error);
* refusals: the grouped MFMA overlap and the split products' overlap with h1aa, the
  no-allocation entry with h1aa, the adjoint, a non-symmetric h1aa and I + h1aa not positive
  definite; clearing the terms restores the 'none' result bit for bit (graph recaptured).
"""
import numpy as np
import pytest

from conftest import iora_terms, workload

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_iora_per_k_against_oracle():
    import torch
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 1)
    cr, kp = w.crystal, w.kpoints[0]
    t = iora_terms(cr, 21)
    ctx = lapw.Context(cr)
    ctx.set_iora(**t)
    H_ref, O_ref, A = R.setup(cr, kp, iora=t)
    H, O = ctx.hamiltonandoverlapsetup(kp, A)
    eh, eo = _rel(H, H_ref), _rel(O, O_ref)
    print(f"per-k host: |dH|/|H| {eh:.2e} |dO|/|O| {eo:.2e}", flush=True)
    assert eh <= 1e-12 and eo <= 1e-12
    _, O0, _ = R.setup(cr, kp, A)
    assert _rel(O0, O_ref) > 1e-6                       # the terms are there
    # O alone through the host entry
    nmat = kp.ngp + cr.nlotot
    ngkmax = A.shape[0]
    Oo = np.zeros((nmat, nmat), dtype=np.complex128, order="F")
    apw = np.asfortranarray(A, dtype=np.complex128)
    ig = np.ascontiguousarray(kp.igpig[:ngkmax], dtype=np.int32)
    vg = np.asfortranarray(kp.vgpc[:, :ngkmax], dtype=np.float64)
    ctx._call("exc_overlapsetup", ctx.h, kp.ngp, ngkmax, lapw._ptr(apw), lapw._ptr(ig), lapw._ptr(vg), nmat,
              lapw._ptr(Oo))
    assert np.array_equal(Oo, O)
    # the device entry on device arrays
    b = lapw.DeviceBatch(cr, [kp])
    d = b.dks[0]
    ctx.match_dev(d.ngp, d.ngkmax, d.gpc, d.tpgpc, d.sfacgp, b.apw[0])
    ctx.setup_dev(d.ngp, d.ngkmax, b.apw[0], d.igpig, d.vgpc, nmat, b.H[0], b.O[0])
    torch.cuda.synchronize()
    Hd, Od = b.H[0].cpu().numpy().T, b.O[0].cpu().numpy().T
    print(f"per-k device (its own match): |dH|/|H| {_rel(Hd, H_ref):.2e} |dO|/|O| {_rel(Od, O_ref):.2e}",
          flush=True)
    assert _rel(Hd, H_ref) <= 1e-12 and _rel(Od, O_ref) <= 1e-12
    ctx.close()


def test_iora_grouped_against_oracle():
    import torch
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 2)
    cr = w.crystal
    t = iora_terms(cr, 22)
    s = torch.cuda.Stream()
    ctx = lapw.Context(cr, stream=s)
    ctx.set_iora(**t)
    refs = [R.setup(cr, kp, iora=t) for kp in w.kpoints]
    runs = [("FP64 eager", lapw.EXC_BATCH_OLP_LEGENDRE, 1e-12),
            ("FP64 graph", lapw.EXC_BATCH_OLP_LEGENDRE | lapw.EXC_BATCH_GRAPH, 1e-12),
            ("FP64 graph replay", lapw.EXC_BATCH_OLP_LEGENDRE | lapw.EXC_BATCH_GRAPH, 1e-12),
            ("modular split", lapw.EXC_BATCH_OLP_LEGENDRE | lapw.EXC_BATCH_SPLIT_CRT, 1e-9)]
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(cr, w.kpoints)
        for name, fl, hbar in runs:
            for m in b.H + b.O:
                m.fill_(float("nan"))
            ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH | fl)
            torch.cuda.synchronize()
            if fl & lapw.EXC_BATCH_SPLIT_CRT:
                assert ctx.crt_form()[0] == 0          # h1aa: the (A, Yt) form
            for i, (H_ref, O_ref, _) in enumerate(refs):
                eh = _rel(b.H[i].cpu().numpy().T, H_ref)
                eo = _rel(b.O[i].cpu().numpy().T, O_ref)
                print(f"{name} k{i}: |dH|/|H| {eh:.2e} |dO|/|O| {eo:.2e}", flush=True)
                assert eh <= hbar and eo <= 1e-12
    ctx.close()


def test_iora_refusals_and_clear():
    import torch
    from exciting_amd import lapw
    w = workload(1, 2)
    cr = w.crystal
    t = iora_terms(cr, 23)
    s = torch.cuda.Stream()
    ctx = lapw.Context(cr, stream=s)
    flags = lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_OLP_LEGENDRE | lapw.EXC_BATCH_GRAPH

    def code(f):
        with pytest.raises(lapw.ExcError) as e:
            f()
        return e.value.code

    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(cr, w.kpoints)
        ctx.setup_batched_dev(b, flags)
        torch.cuda.synchronize()
        ref = [(h.cpu().clone(), o.cpu().clone()) for h, o in zip(b.H, b.O)]
        # input checks leave the state as it was
        bad = t["h1aa"].copy()
        bad[0, 1, 2, 0] += 1e-3
        assert code(lambda: ctx.set_iora(h1aa=bad)) == lapw.EXC_ERR_ARG
        neg = np.zeros_like(t["h1aa"])
        neg[0, 0] = neg[1, 1] = -2.0
        assert code(lambda: ctx.set_iora(h1aa=neg)) == lapw.EXC_ERR_UNSUPPORTED
        ctx.setup_batched_dev(b, flags)
        torch.cuda.synchronize()
        for i in range(b.nk):
            assert torch.equal(b.H[i].cpu(), ref[i][0]) and torch.equal(b.O[i].cpu(), ref[i][1])
        ctx.set_iora(**t)
        # the forms that cannot take the terms
        assert code(lambda: ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH)) == lapw.EXC_ERR_UNSUPPORTED
        d = b.dks[0]
        rc = ctx.lib.exc_hamiltonandoverlapsetup_noalloc_dev(ctx.h, d.ngp, d.ngkmax, b.apw[0].data_ptr(),
                                                             d.igpig.data_ptr(), d.vgpc.data_ptr(), b.nmat[0],
                                                             b.H[0].data_ptr(), b.O[0].data_ptr())
        assert rc == lapw.EXC_ERR_UNSUPPORTED
        assert code(lambda: ctx._call("exc_ho_vjp_dev", ctx.h, 1, 1, None, None, None, None, 1, 1, None, None,
                                      None, None, None)) == lapw.EXC_ERR_UNSUPPORTED
        ctx.set_iora(m2effig=t["m2effig"])
        assert code(lambda: ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_SPLIT_CRT)) == \
            lapw.EXC_ERR_UNSUPPORTED
        ctx.setup_batched_dev(b, flags)                 # the interstitial term alone: the graph path
        torch.cuda.synchronize()
        assert not torch.equal(b.O[0].cpu(), ref[0][1]) and torch.equal(b.H[0].cpu(), ref[0][0])
        # cleared: the 'none' bits again (the graph is captured again)
        ctx.set_iora()
        ctx.setup_batched_dev(b, flags)
        torch.cuda.synchronize()
        for i in range(b.nk):
            assert torch.equal(b.H[i].cpu(), ref[i][0]) and torch.equal(b.O[i].cpu(), ref[i][1])
    ctx.status()
    ctx.close()


def test_iora_interstitial_term_with_gram_split():
    """m2effig alone leaves the modular split in its Gram form, whose merge takes O's APW block
    as O minus O's interstitial element — with the m2effig term: H equals the 'none' run's
    to 1e-13 (the recovered block differs from the FP64 one by its rounding), O gains the
    term."""
    import torch
    from exciting_amd import lapw
    w = workload(1, 1)
    cr = w.crystal
    t = iora_terms(cr, 24)
    s = torch.cuda.Stream()
    ctx = lapw.Context(cr, stream=s)
    flags = lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_OLP_LEGENDRE | lapw.EXC_BATCH_SPLIT_CRT
    out = []
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(cr, w.kpoints)
        for m in (None, t["m2effig"]):
            ctx.set_iora(m2effig=m)
            ctx.setup_batched_dev(b, flags)
            torch.cuda.synchronize()
            assert ctx.crt_form()[0] == 1
            out.append((b.H[0].cpu().numpy().T.copy(), b.O[0].cpu().numpy().T.copy()))
    dh = _rel(out[1][0], out[0][0])
    do = _rel(out[1][1], out[0][1])
    print(f"Gram split with m2effig: H vs 'none' {dh:.2e}, O vs 'none' {do:.2e}", flush=True)
    assert dh <= 1e-13 and do > 1e-8
    ctx.close()
