"""Grouped setup with EXC_BATCH_OLP_LEGENDRE: the muffin-tin part of O by the spherical-
harmonic addition theorem from the radial factors of match (k_olp.hip).

* apwalm and H bit-identical to the grouped run without the flag (it changes O only);
* O against the oracle, which follows the reference's conj(A) A^T contraction
  (overlapsetup / olpaa): relative Frobenius error <= 1e-12;
* eigenvalues of (H, O) against the oracle's (LAPACK zhegv on both) <= 1e-10 Ha;
* replays of the captured graph equal the eager result.
The identity itself is pinned on the CPU (tests/test_oracle_setup.py::
test_overlap_addition_theorem)."""
import numpy as np
import pytest

from conftest import small, workload

pytestmark = pytest.mark.gpu


def _run(w, flags, olp_variant=None):
    import torch
    from exciting_amd import lapw
    s = torch.cuda.Stream()
    ctx = lapw.Context(w.crystal, stream=s)
    if olp_variant is not None:
        ctx.set_variant("olp", olp_variant)
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(w.crystal, w.kpoints)
        for t in b.H + b.O:
            t.fill_(float("nan"))
        ctx.setup_batched_dev(b, flags)
        torch.cuda.synchronize()
    return ctx, s, b, [(b.apw[i].cpu(), b.H[i].cpu(), b.O[i].cpu()) for i in range(b.nk)]


@pytest.mark.parametrize("variant", [0, 2, 3])
@pytest.mark.parametrize("which", ["small-d", "workload0", "workload1"])
def test_olp_legendre_against_oracle(which, variant):
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = {"small-d": lambda: small("d"), "workload0": lambda: workload(0), "workload1": lambda: workload(1, 2)}[which]()
    base = lapw.EXC_BATCH_MATCH
    ref = _run(w, base)[3]
    res = _run(w, base | lapw.EXC_BATCH_OLP_LEGENDRE, olp_variant=variant)[3]
    for i, kp in enumerate(w.kpoints):
        a0, h0, _ = ref[i]
        a1, h1, o1 = res[i]
        n = kp.ngp
        assert np.array_equal(a1.numpy()[..., :n], a0.numpy()[..., :n]), f"k{i}: apwalm differs"
        assert np.array_equal(h1.numpy(), h0.numpy()), f"k{i}: H differs"
        H_ref, O_ref, _ = R.setup(w.crystal, kp)
        O = o1.numpy().T
        H = h1.numpy().T
        eo = np.linalg.norm(O - O_ref) / np.linalg.norm(O_ref)
        nev = min(200, n)
        e_ref, _ = R.eigen(H_ref, O_ref, nev)
        e, _ = R.eigen(np.asfortranarray(H), np.asfortranarray(O), nev)
        de = np.abs(e - e_ref).max()
        print(f"{which} overlap variant {variant} k{i} N={n}: |dO|/|O| {eo:.2e}, max |d eig| (lowest {nev}) "
              f"{de:.2e} Ha", flush=True)
        assert eo <= 1e-12
        assert de <= 1e-10
        il = np.tril_indices(O.shape[0], -1)
        assert np.array_equal(O[il], np.conj(O.T[il]))


@pytest.mark.parametrize("shape", [0, 1])
def test_olp_legendre_graph_replays(shape):
    """Graph replays equal the eager launch bit for bit, with the overlap and lo kernels on a
    branch of their own (graph variant 0) or in line (1)."""
    import torch
    from exciting_amd import lapw
    w = workload(1, 4)
    flags = lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_OLP_LEGENDRE
    ctx, s, b, eager = _run(w, flags)
    ctx.set_variant("graph", shape)
    with torch.cuda.stream(s):
        for rep in range(3):
            for t in b.H + b.O:
                t.fill_(float("nan"))
            x = torch.randn(512, 512, device="cuda", dtype=torch.float64)
            (x @ x).sum().item()
            ctx.setup_batched_dev(b, flags | lapw.EXC_BATCH_GRAPH)
            torch.cuda.synchronize()
            print(f"graph variant {shape} replay {rep}", flush=True)
            for i in range(b.nk):
                assert torch.equal(b.H[i].cpu(), eager[i][1]) and torch.equal(b.O[i].cpu(), eager[i][2])


def test_olp_legendre_needs_match():
    from exciting_amd import lapw
    w = small("a")
    with pytest.raises(RuntimeError):
        _run(w, lapw.EXC_BATCH_OLP_LEGENDRE)


def test_olp_variants_bit_identical():
    """The overlap kernel's thread blockings (4 x 2 and 4 x 4 elements per thread) run the same
    Clenshaw recurrence per element in the same order: O bit for bit; the row-parallel
    variant 2 (scaled recurrence) agrees to within rounding, and its pipelined form (variant
    3) bit for bit."""
    import torch
    from exciting_amd import lapw
    w = workload(1, 3)
    s = torch.cuda.Stream()
    ctx = lapw.Context(w.crystal, stream=s)
    flags = lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_OLP_LEGENDRE
    out = {}
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(w.crystal, w.kpoints)
        for v in range(ctx.num_variants("olp")):
            ctx.set_variant("olp", v)
            for t in b.O:
                t.fill_(float("nan"))
            ctx.setup_batched_dev(b, flags)
            torch.cuda.synchronize()
            out[v] = [t.clone() for t in b.O]
    for i in range(b.nk):
        assert torch.equal(out[1][i], out[0][i]), f"overlap variant 1 k{i}"
        # variant 3 runs variant 2's operations in the same order (pipelined loads only)
        assert torch.equal(out[3][i], out[2][i]), f"overlap variant 3 k{i}"
        # variant 2 runs the recurrence scaled (another rounding): within a few units of the
        # last place of the largest element
        d = (out[2][i] - out[0][i]).abs().max().item()
        print(f"overlap variant 2 k{i}: max |dO| {d:.2e} (max |O| {out[0][i].abs().max().item():.2e})", flush=True)
        assert d <= 1e-14 * out[0][i].abs().max().item()
