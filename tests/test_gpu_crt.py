"""Modular split of the muffin-tin contraction (EXC_BATCH_SPLIT_CRT) on the GPU.

The integer products are exact, so the GPU result is fixed by the scheme itself:
* O without the addition theorem, O_MT = conj(A) A^T from apwalm as the GPU's match wrote it:
  the emulation (oracle/crt_ref.py) on the same apwalm gives the same FP64 values; the GPU adds
  Theta(G - G') once, so every APW-APW element of the upper triangle equals emulation +
  olpistln to within one rounding of that sum (|d| <= 2^-52 (|O_MT| + |Theta|)), and the strict
  lower triangle is conj(upper) bit for bit;
* H: Yt on the GPU and on the host differ by rounding, which moves a few of Yt's integers by
  one unit, so H is held to twice the scheme's own error bound: each rounded row is within
  sqrt(2K)/2 units of the exact one, a unit is below 2^(1 - beta) |X_i|, so an element of the
  product is within 2^(1 - beta) sqrt(2K) |A_i| |Yt_j| of the exact one (beta = 44 with 12
  moduli, 40 with the default 11);
* the lowest 200 eigenvalues within 1e-8 Ha of the FP64 oracle's, on workload 1 (2 k-points)
  and on workload 3's first k-point (the committed oracle probes): the contract's precision bar;
* every GEMM variant gives the same H and O bit for bit (integer sums do not depend on order),
  and graph replays equal the eager launch;
* the Gram form (E = A conj(U) (Lambda + sigma)^(1/2), H_MT = conj(E) E^T - sigma O_APW), the
  default with the addition-theorem overlap: against the emulation on the host's E to the
  same kind of bound, and both forms' eigenvalues within 1e-8 Ha.
"""
import os

import numpy as np
import pytest

from conftest import small, workload

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS = 2.0 ** -52


def _run(w, nk, flags, variant=None, nmod=None):
    import torch
    from exciting_amd import lapw
    s = torch.cuda.Stream()
    ctx = lapw.Context(w.crystal, stream=s)
    if variant is not None:
        ctx.set_variant("crtgemm", variant)
    if nmod is not None:
        ctx.set_moduli(nmod)
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(w.crystal, w.kpoints[:nk])
        for t in b.H + b.O:
            t.fill_(float("nan"))
        ctx.setup_batched_dev(b, flags)
        torch.cuda.synchronize()
    ctx.status()
    return ctx, b


def _yt(cr, A, n):
    return np.concatenate([A[:n, :, :, a].reshape(n, -1, order="F") @ cr.haa[:, :, a].T for a in range(cr.natm)],
                          axis=1)


@pytest.mark.parametrize("which", ["small-d", "workload1"])
def test_crt_against_emulation(which):
    from exciting_amd import lapw
    from oracle import crt_ref as Q
    from oracle import lapw_ref as R
    w = small("d") if which == "small-d" else workload(1, 2)
    cr = w.crystal
    nk = len(w.kpoints) if which == "small-d" else 2
    nmod = 12
    beta = Q.beta(nmod)
    ctx, b = _run(w, nk, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_SPLIT_CRT, nmod=nmod)
    for i in range(nk):
        kp = w.kpoints[i]
        n = kp.ngp
        A4 = b.apw[i].cpu().numpy().transpose(3, 2, 1, 0)          # (ngkmax, io, lm, atom), the GPU's apwalm
        H = b.H[i].cpu().numpy().T
        O = b.O[i].cpu().numpy().T
        A = A4[:n].reshape(n, -1, order="F")
        Y = _yt(cr, A4, n)
        Oe, _, _ = Q.product(A, A, nmod)
        He, _, _ = Q.product(A, Y, nmod)
        Oi = np.zeros((n, n), dtype=complex)
        R.olpistln(cr, kp, Oi)
        Hi = np.zeros((n + cr.nlotot,) * 2, dtype=complex)
        R.hmlistln(cr, kp, Hi)
        Hi = Hi[:n, :n]
        iu = np.triu_indices(n)
        dO = np.abs(O[:n, :n][iu] - (Oe[iu] + Oi[iu]))
        tolO = EPS * (np.abs(Oe[iu]) + np.abs(Oi[iu]))
        na, ny = np.linalg.norm(A, axis=1), np.linalg.norm(Y, axis=1)
        dH = np.abs(H[:n, :n][iu] - (He[iu] + Hi[iu]))
        tolH = (2.0 ** (2 - beta) * np.sqrt(2 * A.shape[1]) * np.outer(na, ny)[iu]
                + 4 * EPS * (np.abs(He[iu]) + np.abs(Hi[iu])))
        il = np.tril_indices(n, -1)
        print(f"{which} k{i} N={n}: O vs emulation max |d|/tol {np.max(dO / np.maximum(tolO, 1e-300)):.2f}, "
              f"H max |d|/bound {np.max(dH / tolH):.2e}", flush=True)
        assert (dO <= tolO).all()
        assert (dH <= tolH).all()
        assert np.array_equal(O[:n, :n][il], np.conj(O[:n, :n].T[il]))
        assert np.array_equal(H[:n, :n][il], np.conj(H[:n, :n].T[il]))
        # the lo rows and columns are the FP64 path's
        H_ref, O_ref, _ = R.setup(cr, kp)
        for M, Mr in ((H, H_ref), (O, O_ref)):
            assert np.linalg.norm(M[n:, :] - Mr[n:, :]) <= 1e-12 * np.linalg.norm(Mr)
    ctx.close()


def _gram_e(cr, A, n, sigma):
    """E = A conj(U) (Lambda + sigma)^(1/2) per atom from h_a = U Lambda U^H (numpy's eigh; the
    library's rocSOLVER vectors differ by phases and within degenerate spaces, conj(E) E^T
    does not)."""
    out = []
    for a in range(cr.natm):
        lam, U = np.linalg.eigh(cr.haa[:, :, a])
        out.append(A[:n, :, :, a].reshape(n, -1, order="F") @ (np.conj(U) * np.sqrt(np.maximum(lam + sigma, 0.0))))
    return np.concatenate(out, axis=1)


def test_crt_gram_form_against_emulation():
    """The Gram form (the default with the addition-theorem overlap): the library reports form 1,
    10 moduli and sigma = -min lambda(h_a) (1 + 1e-6) + 1e-10 max |lambda|; H's APW block equals the emulation's
    conj(E) E^T (oracle/crt_ref.py on the host's E) - sigma O_APW + the interstitial term to
    the scheme's own bound (each rounded row of E within sqrt(2K)/2 units of the exact one, a
    unit below 2^(1 - beta) |E_i|), plus the roundings of the FP64 terms."""
    from exciting_amd import lapw
    from oracle import crt_ref as Q
    from oracle import lapw_ref as R
    w = workload(1, 2)
    cr = w.crystal
    ctx, b = _run(w, 2, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_SPLIT_CRT | lapw.EXC_BATCH_OLP_LEGENDRE)
    form, nmod, sigma = ctx.crt_form()
    lmin = min(np.linalg.eigvalsh(cr.haa[:, :, a])[0] for a in range(cr.natm))
    print(f"form {form}, {nmod} moduli, sigma {sigma:.6f} (min lambda {lmin:.6f})", flush=True)
    assert form == 1 and nmod == 10
    lmax = max(np.abs(np.linalg.eigvalsh(cr.haa[:, :, a])).max() for a in range(cr.natm))
    assert abs(sigma - (max(0.0, -lmin) * (1 + 1e-6) + 1e-10 * lmax)) <= 1e-12 * max(1.0, abs(lmin))
    beta = Q.beta(nmod)
    for i in range(2):
        kp = w.kpoints[i]
        n = kp.ngp
        A4 = b.apw[i].cpu().numpy().transpose(3, 2, 1, 0)
        H = b.H[i].cpu().numpy().T
        O = b.O[i].cpu().numpy().T
        E = _gram_e(cr, A4, n, sigma)
        Ge, _, _ = Q.product(E, E, nmod)
        Oi = np.zeros((n, n), dtype=complex)
        R.olpistln(cr, kp, Oi)
        Hi = np.zeros((n + cr.nlotot,) * 2, dtype=complex)
        R.hmlistln(cr, kp, Hi)
        Hi = Hi[:n, :n]
        iu = np.triu_indices(n)
        Oapw = O[:n, :n][iu] - Oi[iu]
        ref = Ge[iu] - sigma * Oapw + Hi[iu]
        ne = np.linalg.norm(E, axis=1)
        tol = (2.0 ** (2 - beta) * np.sqrt(2 * E.shape[1]) * np.outer(ne, ne)[iu]
               + 8 * EPS * (np.abs(Ge[iu]) + sigma * np.abs(O[:n, :n][iu]) + np.abs(Hi[iu])))
        d = np.abs(H[:n, :n][iu] - ref)
        print(f"Gram form k{i}: H max |d|/bound {np.max(d / tol):.2e}", flush=True)
        assert (d <= tol).all()
    ctx.close()


@pytest.mark.parametrize("form,nmod", [(0, 0), (1, 0), (1, 11), (0, 10)])
def test_crt_forms_eigenvalues(form, nmod):
    """Both forms of the modular split, at the built-in and at a pinned moduli count: the
    reported form and count, and the lowest 200 eigenvalues of workload 1's first k-point
    within 1e-8 Ha of the FP64 oracle's."""
    import torch
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 1)
    s = torch.cuda.Stream()
    ctx = lapw.Context(w.crystal, stream=s)
    ctx.set_variant("crtform", form)
    ctx.set_moduli(nmod)
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(w.crystal, w.kpoints[:1])
        ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_SPLIT_CRT | lapw.EXC_BATCH_OLP_LEGENDRE)
        torch.cuda.synchronize()
    f, n, sg = ctx.crt_form()
    assert f == form and n == (nmod or (10 if form else 11)) and (sg > 0.0) == (form == 1)
    H_ref, O_ref, _ = R.setup(w.crystal, w.kpoints[0])
    de = np.abs(_eig(b.H[0].cpu().numpy().T, b.O[0].cpu().numpy().T) - _eig(H_ref, O_ref)).max()
    print(f"form {f}, {n} moduli: lowest 200 eigenvalues vs FP64 oracle {de:.2e} Ha", flush=True)
    assert de <= 1e-8
    ctx.close()


def test_crt_gram_after_eigensolve_and_new_h():
    """The Gram form's eigen-decomposition runs on a rocBLAS handle of its own: after an
    eigensolve on the same context (whose workers pin their workspaces) and a new set of h_a
    (the context's tables uploaded again: the decomposition is redone), the split still takes
    the Gram form and its eigenvalues stay within 1e-8 Ha of the FP64 oracle's."""
    import torch
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 1)
    cr = w.crystal
    s = torch.cuda.Stream()
    ctx = lapw.Context(cr, stream=s)
    H_ref, O_ref, A = R.setup(cr, w.kpoints[0])
    H, O = ctx.hamiltonandoverlapsetup(w.kpoints[0], A)
    ev, _ = ctx.eigen(H, O, 50)
    assert np.abs(ev - _eig(H_ref, O_ref, 50)).max() <= 1e-8
    flags = lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_SPLIT_CRT | lapw.EXC_BATCH_OLP_LEGENDRE | lapw.EXC_BATCH_GRAPH
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(cr, w.kpoints[:1])
        for rep in range(2):
            if rep:                                    # h_a set again: the table is rebuilt
                hold = [np.asfortranarray(x, dtype=t) for x, t in ((cr.haa, np.complex128), (cr.loa, np.complex128),
                        (cr.lolo, np.complex128), (cr.oalo, np.float64), (cr.ololo, np.float64))]
                ctx._call("exc_ctx_set_mt", ctx.h, *[lapw._ptr(x) for x in hold])
            ctx.setup_batched_dev(b, flags)
            torch.cuda.synchronize()
            assert ctx.crt_form()[0] == 1
            de = np.abs(_eig(b.H[0].cpu().numpy().T, b.O[0].cpu().numpy().T) - _eig(H_ref, O_ref)).max()
            print(f"Gram form after an eigensolve, pass {rep}: {de:.2e} Ha", flush=True)
            assert de <= 1e-8
    ctx.close()


def test_crt_gram_form_left_for_a_large_shift():
    """With h_a - 30 I (eigenvalues from about -31 to -7) the shift sigma would exceed half of
    the largest |eigenvalue|: the built-in choice takes the (A, Yt) form at 11 moduli, and H
    and O stay within the split's bar of the FP64 oracle's on the same h."""
    import dataclasses
    import torch
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 1)
    cr = w.crystal
    haa = cr.haa.copy()
    for a in range(cr.natm):
        haa[:, :, a] -= 30.0 * np.eye(haa.shape[0])
    cr2 = dataclasses.replace(cr, haa=haa)
    s = torch.cuda.Stream()
    ctx = lapw.Context(cr2, stream=s)
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(cr2, w.kpoints[:1])
        ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_SPLIT_CRT | lapw.EXC_BATCH_OLP_LEGENDRE)
        torch.cuda.synchronize()
    form, nmod, _ = ctx.crt_form()
    H_ref, O_ref, _ = R.setup(cr2, w.kpoints[0])
    H, O = b.H[0].cpu().numpy().T, b.O[0].cpu().numpy().T
    eh = np.linalg.norm(H - H_ref) / np.linalg.norm(H_ref)
    eo = np.linalg.norm(O - O_ref) / np.linalg.norm(O_ref)
    print(f"h - 30 I: form {form}, {nmod} moduli, |dH|/|H| {eh:.2e} |dO|/|O| {eo:.2e}", flush=True)
    assert form == 0 and nmod == 11
    assert eh <= 1e-9 and eo <= 1e-12
    ctx.close()


def _eig(H, O, nev=200):
    import scipy.linalg
    m = min(nev, H.shape[0])
    return scipy.linalg.eigh(H, O, lower=False, subset_by_index=[0, m - 1], driver="gvx", eigvals_only=True)


def test_crt_eigenvalues_workload1():
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 2)
    ctx, b = _run(w, 2, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_SPLIT_CRT | lapw.EXC_BATCH_OLP_LEGENDRE)
    for i in range(2):
        H_ref, O_ref, _ = R.setup(w.crystal, w.kpoints[i])
        e0 = _eig(H_ref, O_ref)
        e1 = _eig(b.H[i].cpu().numpy().T, b.O[i].cpu().numpy().T)
        de = np.abs(e1 - e0).max()
        print(f"workload 1 k{i}: lowest 200 eigenvalues vs FP64 oracle {de:.2e} Ha", flush=True)
        assert de <= 1e-8
    ctx.close()


def test_crt_eigenvalues_workload3():
    import torch
    from exciting_amd import lapw
    from golden.make_workload_probes import NEV, probes
    ref = np.load(os.path.join(GOLDEN, "workload3_k0_probes.npz"))
    w = workload(3)
    ctx, b = _run(w, 8, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_SPLIT_CRT | lapw.EXC_BATCH_OLP_LEGENDRE |
                  lapw.EXC_BATCH_GRAPH)
    H = b.H[0].cpu().numpy().T
    O = b.O[0].cpu().numpy().T
    del b
    torch.cuda.empty_cache()
    X = probes(H.shape[0])
    eh = np.linalg.norm(H @ X - ref["hx"]) / np.linalg.norm(ref["hx"])
    eo = np.linalg.norm(O @ X - ref["ox"]) / np.linalg.norm(ref["ox"])
    e = _eig(H, O, NEV)
    de = np.abs(e - ref["eig"]).max()
    form, nmod, _ = ctx.crt_form()
    print(f"workload 3 k0, modular split (form {form}, {nmod} moduli): H {eh:.2e} O {eo:.2e} (probes), lowest "
          f"{NEV} eigenvalues {de:.2e} Ha", flush=True)
    assert eo <= 1e-12 and eh <= 1e-9
    assert de <= 1e-8
    ctx.close()


def test_crt_variants_bit_identical_and_graph():
    import torch
    from exciting_amd import lapw
    w = workload(1, 3)
    flags = lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_SPLIT_CRT
    out = {}
    s = torch.cuda.Stream()
    ctx = lapw.Context(w.crystal, stream=s)
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(w.crystal, w.kpoints)
        for v in range(ctx.num_variants("crtgemm")):
            print(f"workload 1 nk {b.nk} modular split GEMM variant {v}", flush=True)
            ctx.set_variant("crtgemm", v)
            for t in b.H + b.O:
                t.fill_(float("nan"))
            ctx.setup_batched_dev(b, flags)
            torch.cuda.synchronize()
            out[v] = ([t.clone() for t in b.H], [t.clone() for t in b.O])
        for v in out:
            for i in range(b.nk):
                assert torch.equal(out[v][0][i], out[0][0][i]), f"variant {v} H k{i}"
                assert torch.equal(out[v][1][i], out[0][1][i]), f"variant {v} O k{i}"
        ctx.set_variant("crtgemm", -1)
        ref = out[0]
        for rep in range(3):
            for t in b.H + b.O:
                t.fill_(float("nan"))
            x = torch.randn(256, 256, device="cuda", dtype=torch.float64)
            (x @ x).sum().item()
            ctx.setup_batched_dev(b, flags | lapw.EXC_BATCH_GRAPH)
            torch.cuda.synchronize()
            print(f"modular split graph replay {rep}", flush=True)
            for i in range(b.nk):
                assert torch.equal(b.H[i], ref[0][i]) and torch.equal(b.O[i], ref[1][i]), f"replay {rep} k{i}"
    ctx.status()
    ctx.close()


def test_crt_caller_apwalm_equals_match():
    """Without EXC_BATCH_MATCH the caller's apwalm is used and A's row exponents come from a
    pass over it (with the match they are summed from the norms the match kernel writes in
    closed form).  On the same apwalm the two runs agree bit for bit (the exponents are the
    same integers)."""
    import torch
    from exciting_amd import lapw
    w = workload(1, 2)
    s = torch.cuda.Stream()
    ctx = lapw.Context(w.crystal, stream=s)
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(w.crystal, w.kpoints)
        ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_SPLIT_CRT)
        torch.cuda.synchronize()
        ref = [(h.clone(), o.clone()) for h, o in zip(b.H, b.O)]
        for t in b.H + b.O:
            t.fill_(float("nan"))
        ctx.setup_batched_dev(b, lapw.EXC_BATCH_SPLIT_CRT)
        torch.cuda.synchronize()
        for i in range(b.nk):
            print(f"workload 1 k{i}: caller apwalm against match", flush=True)
            assert torch.equal(b.H[i], ref[i][0]) and torch.equal(b.O[i], ref[i][1]), f"k{i}"
    ctx.close()


def test_crt_moduli_count():
    """Fewer moduli give fewer bits and a larger error, within each count's own bound
    2^(1 - beta) sqrt(2K) |A_i| |A_j| against the FP64 oracle."""
    from exciting_amd import lapw
    from oracle import crt_ref as Q
    from oracle import lapw_ref as R
    w = small("d")
    cr, kp = w.crystal, w.kpoints[0]
    n = kp.ngp
    H_ref, O_ref, A = R.setup(cr, kp)
    errs = []
    for nmod in (9, 11, 12):
        ctx, b = _run(w, 1, lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_SPLIT_CRT, nmod=nmod)
        O = b.O[0].cpu().numpy().T
        Am = A[:n].reshape(n, -1, order="F")
        na = np.linalg.norm(Am, axis=1)
        d = np.abs(O[:n, :n] - O_ref[:n, :n])
        bound = (2.0 ** (1 - Q.beta(nmod)) * np.sqrt(2 * Am.shape[1]) * np.outer(na, na)
                 + 1e-14 * np.abs(O_ref[:n, :n]))
        errs.append(np.linalg.norm(O - O_ref) / np.linalg.norm(O_ref))
        print(f"small-d, {nmod} moduli (beta {Q.beta(nmod)}): |dO|/|O| {errs[-1]:.2e}", flush=True)
        assert (d <= bound).all()
        ctx.close()
    assert errs[0] > errs[2]
