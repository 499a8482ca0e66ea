"""exc_hamiltonandoverlapsetup_noalloc_dev, the setup entry that allocates nothing (its Y-stage
workspace is the library's fixed device array), against exc_hamiltonandoverlapsetup_dev: the
same kernels and variants, so H and O must agree bit for bit; two calls in a row, the second
with a smaller basis (the array keeps the first call's contents, which the second must not
read), and a basis past the array's limit is refused with EXC_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest

from conftest import workload

pytestmark = pytest.mark.gpu


def test_noalloc_setup_matches_allocating_entry():
    import torch
    from exciting_amd import lapw
    w = workload(1, 2)
    cr = w.crystal
    kps = sorted(w.kpoints, key=lambda k: -k.ngp)          # larger basis first, then the smaller one
    assert kps[0].ngp > kps[1].ngp
    s = torch.cuda.Stream()
    ctx = lapw.Context(cr, stream=s)
    with torch.cuda.stream(s):
        b = lapw.DeviceBatch(cr, kps)
        for i, d in enumerate(b.dks):
            ctx.match_dev(d.ngp, d.ngkmax, d.gpc, d.tpgpc, d.sfacgp, b.apw[i])
        for i, d in enumerate(b.dks):
            n = b.nmat[i]
            Hn = torch.full((n, n), float("nan"), dtype=torch.complex128, device="cuda")
            On = torch.full_like(Hn, float("nan"))
            print(f"no-alloc setup call {i}: ngp {d.ngp}", flush=True)
            rc = ctx.lib.exc_hamiltonandoverlapsetup_noalloc_dev(ctx.h, d.ngp, d.ngkmax, b.apw[i].data_ptr(),
                                                                 d.igpig.data_ptr(), d.vgpc.data_ptr(), n,
                                                                 Hn.data_ptr(), On.data_ptr())
            assert rc == 0, rc
            ctx.setup_dev(d.ngp, d.ngkmax, b.apw[i], d.igpig, d.vgpc, n, b.H[i], b.O[i])
            torch.cuda.synchronize()
            assert torch.equal(Hn, b.H[i]) and torch.equal(On, b.O[i]), f"call {i} differs"
        # a basis whose workspace exceeds the fixed array is refused before anything runs
        big = (1 << 30) // (16 * cr.natm * cr.maxaa) + 1
        rc = ctx.lib.exc_hamiltonandoverlapsetup_noalloc_dev(ctx.h, big, big, b.apw[0].data_ptr(),
                                                             b.dks[0].igpig.data_ptr(), b.dks[0].vgpc.data_ptr(),
                                                             big + cr.nlotot, b.H[0].data_ptr(), b.O[0].data_ptr())
        assert rc == lapw.EXC_ERR_ARG
    ctx.status()
    ctx.close()
