"""Host-side logic that runs without a GPU: workload generation rules, the autotuner's
table persistence, the work-list decode used by the per-k muffin-tin launch."""
import json
import math

import numpy as np

from conftest import workload
from exciting_amd import autotune, synth


def test_workload_rules():
    w = workload(2, 4)
    cr = w.crystal
    assert cr.natm == 16 and cr.nspecies == 3 and cr.nlotot == 64
    # minimum periodic distance >= R_a + R_b
    frac = np.linalg.solve(cr.avec, cr.atposc)
    for a in range(cr.natm):
        for b in range(a):
            d = frac[:, a] - frac[:, b]
            d -= np.round(d)
            best = min(np.linalg.norm(cr.avec @ (d + np.array([i, j, k])))
                       for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1))
            assert best >= cr.rmt[cr.spidx[a]] + cr.rmt[cr.spidx[b]] - 1e-12
    # V(-G) = conj V(G), V(0) real; G table covers all basis differences (2 gkmax < gmaxvr)
    key = {tuple(cr.ivg[:, g]): g for g in range(cr.ngvec)}
    for g in range(0, cr.ngvec, 97):
        assert abs(cr.veffig[key[tuple(-cr.ivg[:, g])]] - np.conj(cr.veffig[g])) < 1e-15
    assert cr.veffig[0].imag == 0.0
    for kp in w.kpoints:
        assert 2 * kp.gkmax < cr.gmaxvr
        assert 2400 <= kp.ngp <= 2700
    # Theta(0) = 1 - sum 4 pi R^3 / (3 Omega)
    t0 = 1.0 - sum(4 * math.pi * cr.rmt[s] ** 3 / (3 * cr.omega) for s in cr.spidx)
    assert abs(cr.cfunig[0] - t0) < 1e-14


def test_seeds_are_deterministic():
    a = synth.make_config(1, nk=2)
    b = synth.make_config(1, nk=2)
    assert np.array_equal(a.crystal.haa, b.crystal.haa)
    assert np.array_equal(a.kpoints[1].vgpc, b.kpoints[1].vgpc)


def test_autotune_table_roundtrip(tmp_path):
    p = tmp_path / "tune.json"
    autotune.save(p, (autotune.bucket(1230), 242), {"bmt": 4, "ystage": 1})
    autotune.save(p, (autotune.bucket(2550), 242), {"bmt": 5})
    assert autotune.load(p, (1280, 242)) == {"bmt": 4, "ystage": 1}
    assert autotune.load(p, (2560, 242)) == {"bmt": 5}
    assert autotune.load(p, (5120, 242)) is None
    json.loads(p.read_text())
    assert autotune.bucket(1) == 256 and autotune.bucket(256) == 256 and autotune.bucket(257) == 512


def test_upper_triangle_tile_decode():
    # the per-k MT launch enumerates (ti, tj), ti <= tj, column by column
    def decode(t):
        j = int((math.sqrt(8.0 * t + 1.0) - 1.0) * 0.5)
        while (j + 1) * (j + 2) // 2 <= t:
            j += 1
        while j * (j + 1) // 2 > t:
            j -= 1
        return t - j * (j + 1) // 2, j
    seen = [decode(t) for t in range(40 * 41 // 2)]
    assert seen == [(i, j) for j in range(40) for i in range(j + 1)]


def test_autotune_apply_pins_table_choice(tmp_path):
    class Ctx:
        def __init__(self):
            self.pinned = {}

        def set_variant(self, st, v):
            self.pinned[st] = v

    cr = workload(1, 2).crystal
    ngps = [1230, 1174]
    p = tmp_path / "tune.json"
    assert autotune.apply(Ctx(), cr, ngps, p) == {}
    autotune.save(p, autotune.key(cr, ngps), {"bmt": 19, "ystage": 12})
    c = Ctx()
    assert autotune.apply(c, cr, ngps, p) == {"bmt": 19, "ystage": 12}
    assert c.pinned == {"bmt": 19, "ystage": 12}
    assert autotune.key(cr, ngps) == "N1280-naa242-natm8"
    # the committed table is well-formed: every entry names known stages and valid variants
    if autotune.TABLE.exists():
        from exciting_amd import lapw
        for k, v in json.loads(autotune.TABLE.read_text()).items():
            assert k.startswith("N") and "-naa" in k and "-natm" in k
            for st, var in v.items():
                assert st in lapw.STAGES and isinstance(var, int) and var >= 0
