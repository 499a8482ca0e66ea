"""Grouped (batched) setup of many k-points: one launch per stage, the muffin-tin stage a
work list of tiles over all k-points, optionally captured in a hipGraph.

Its arithmetic is the per-k path's (same device code, same summation order), so H, O and
apwalm must equal the per-k results bit for bit, for every four-product tile variant; the
variants with Gauss's three real products per complex product (grouped MT 0, 2, 3 and Y 0,
1: other roundings) give H and O within 1e-14 of the per-k results relative to their norm.
The per-k path is itself checked against the oracle in test_gpu_setup.py, and one k here is
re-checked."""
import numpy as np
import pytest

from conftest import workload

pytestmark = pytest.mark.gpu


def _perk(ctx, batch, four_product=True):
    """The per-k results; with four_product, on the four-product per-k variants (MT 1, Y 3)."""
    import torch
    out = []
    if four_product:
        ctx.set_variant("mt", 1)
        ctx.set_variant("ystage", 3)
    for i, d in enumerate(batch.dks):
        ctx.match_dev(d.ngp, d.ngkmax, d.gpc, d.tpgpc, d.sfacgp, batch.apw[i])
        ctx.setup_dev(d.ngp, d.ngkmax, batch.apw[i], d.igpig, d.vgpc, batch.nmat[i], batch.H[i], batch.O[i])
    torch.cuda.synchronize()
    for i in range(batch.nk):
        out.append((batch.apw[i].clone(), batch.H[i].clone(), batch.O[i].clone()))
    if four_product:
        ctx.set_variant("mt", -1)
        ctx.set_variant("ystage", -1)
    return out


def _clear(batch):
    for i in range(batch.nk):
        batch.apw[i].zero_()
        batch.H[i].zero_()
        batch.O[i].zero_()


GAUSS_BMT = (0, 2, 3)   # grouped MT variants with Gauss's three real products per complex product
GAUSS_Y = (0, 1, 4)  # Y variants likewise


def _same(batch, ref, what, rtol=0.0):
    import torch
    for i in range(batch.nk):
        a, h, o = ref[i]
        n = batch.dks[i].ngp
        assert bool((batch.apw[i][..., :n] == a[..., :n]).all()), f"{what}: apwalm differs at k={i}"
        if rtol == 0.0:
            assert bool((batch.H[i] == h).all()), f"{what}: H differs at k={i}"
            assert bool((batch.O[i] == o).all()), f"{what}: O differs at k={i}"
        else:
            eh = (torch.linalg.norm(batch.H[i] - h) / torch.linalg.norm(h)).item()
            eo = (torch.linalg.norm(batch.O[i] - o) / torch.linalg.norm(o)).item()
            assert eh <= rtol and eo <= rtol, f"{what}: H {eh:.2e}, O {eo:.2e} at k={i}"


@pytest.mark.parametrize("cfg,nk", [(1, 16), (2, 6)])
def test_batched_equals_per_k(cfg, nk):
    import torch
    from exciting_amd import lapw
    w = workload(cfg, nk)
    s = torch.cuda.Stream()
    ctx = lapw.Context(w.crystal, stream=s)
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(w.crystal, w.kpoints)
        ref = _perk(ctx, b)
        ctx.set_variant("ystage", 2)   # a four-product Y: MT variants bit for bit with the per-k path
        for v in range(ctx.num_variants("bmt")):
            print(f"cfg {cfg} nk {nk} grouped MT variant {v}", flush=True)
            ctx.set_variant("bmt", v)
            _clear(b)
            ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH)
            torch.cuda.synchronize()
            _same(b, ref, f"variant {v}", 1e-14 if v in GAUSS_BMT else 0.0)
        ctx.set_variant("bmt", 1)   # the four-product MT kernel: Y variants bit for bit with the per-k path
        for v in range(ctx.num_variants("ystage")):
            print(f"cfg {cfg} nk {nk} grouped Y variant {v}", flush=True)
            ctx.set_variant("ystage", v)
            _clear(b)
            ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH)
            torch.cuda.synchronize()
            _same(b, ref, f"Y variant {v}", 1e-14 if v in GAUSS_Y else 0.0)
        ctx.set_variant("ystage", -1)


def test_graph_replays_match_eager():
    """Replays of the captured grouped launch (default variants) equal the eager grouped
    launch bit for bit, with the program's own eager work between them; the eager grouped
    result is within 1e-14 of the per-k path (the default MT kernel uses Gauss's three
    products)."""
    import torch
    from exciting_amd import lapw
    w = workload(1, 16)
    s = torch.cuda.Stream()
    ctx = lapw.Context(w.crystal, stream=s)
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(w.crystal, w.kpoints)
        perk = _perk(ctx, b)
        _clear(b)
        ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH)
        torch.cuda.synchronize()
        _same(b, perk, "eager grouped vs per-k", 1e-14)
        ref = [(b.apw[i].clone(), b.H[i].clone(), b.O[i].clone()) for i in range(b.nk)]
        flags = lapw.EXC_BATCH_MATCH | lapw.EXC_BATCH_GRAPH
        for rep in range(4):
            _clear(b)
            torch.cuda.synchronize()
            # eager work of the program between replays
            x = torch.randn(512, 512, device="cuda", dtype=torch.float64)
            (x @ x).sum().item()
            ctx.setup_batched_dev(b, flags)          # first call captures, later calls replay
            torch.cuda.synchronize()
            print(f"graph replay {rep}", flush=True)
            _same(b, ref, f"graph replay {rep}")


def test_batched_against_oracle_one_k():
    import torch
    from exciting_amd import lapw
    from oracle import lapw_ref as R
    w = workload(1, 3)
    s = torch.cuda.Stream()
    ctx = lapw.Context(w.crystal, stream=s)
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(w.crystal, w.kpoints)
        ctx.setup_batched_dev(b, lapw.EXC_BATCH_MATCH)
        torch.cuda.synchronize()
    kp = w.kpoints[2]
    H_ref, O_ref, _ = R.setup(w.crystal, kp)
    H = b.H[2].cpu().numpy().T
    O = b.O[2].cpu().numpy().T
    assert np.linalg.norm(H - H_ref) <= 1e-12 * np.linalg.norm(H_ref)
    assert np.linalg.norm(O - O_ref) <= 1e-12 * np.linalg.norm(O_ref)
