"""The G+k oracle (oracle/gvec_ref.py) pinned by what the mathematics fixes:

* the selection by brute force: every chosen G has |G+k| < gkmax, every other G of the first
  ngvec has |G+k| >= gkmax, and the chosen ones keep the G order (gengpvec.f90:50-75);
* sphcrd by reconstruction, v = r (sin t cos p, sin t sin p, cos t), and its special cases
  (the z axis, the -x axis, the zero vector; sphcrd.f90:34-49);
* gensfacgp by |S| = 1, S = 1 for an atom at the origin, S = i for (G+k).r = pi/2, and
  S(v) S(-v) = 1;
* the generator the workloads are built from (exciting_amd/synth.py) agrees with it on the
  seeded k-points: the same igpig, and vgpc, gpc, tpgpc, sfacgp to rounding.
"""
import math

import numpy as np

from conftest import small, workload


def test_selection_brute_force():
    from oracle import gvec_ref as G
    w = workload(1, 3)
    cr = w.crystal
    for kp in w.kpoints:
        ngp, igpig, vgpl, vgpc, gpc, tp = G.gengpvec(cr.ivg, cr.bvec, cr.ngvec, kp.vkl, kp.vkc, kp.gkmax)
        assert ngp == kp.ngp
        vgc = cr.bvec @ cr.ivg[:, :cr.ngvec].astype(float)
        g = np.linalg.norm(vgc + kp.vkc[:, None], axis=0)
        chosen = np.zeros(cr.ngvec, bool)
        chosen[igpig - 1] = True
        assert (g[chosen] < kp.gkmax).all() and (g[~chosen] >= kp.gkmax).all()
        assert (np.diff(igpig) > 0).all()                       # the G order
        assert np.array_equal(vgpl, cr.ivg[:, igpig - 1] + kp.vkl[:, None])


def test_sphcrd_reconstruction_and_special_cases():
    from oracle import gvec_ref as G
    rng = np.random.default_rng(7)
    v = rng.normal(size=(3, 500)) * rng.uniform(0.1, 10.0, 500)
    r, th, ph = G.sphcrd(v)
    back = r * np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    assert np.abs(back - v).max() <= 1e-14 * np.abs(v).max() * 10
    assert (th >= 0).all() and (th <= math.pi).all() and (ph >= 0).all() and (ph < 2 * math.pi).all()
    special = np.array([[0.0, 0.0, 2.0], [-3.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, -1.0]]).T
    r, th, ph = G.sphcrd(special)
    assert np.array_equal(r, [2.0, 3.0, 0.0, 1.0])
    assert th[0] == 0.0 and ph[0] == 0.0                        # +z: no azimuth
    assert th[1] == math.acos(0.0) and ph[1] == math.pi          # -x
    assert th[2] == 0.0 and ph[2] == 0.0                         # zero vector
    assert th[3] == math.pi and ph[3] == 0.0                     # -z


def test_structure_factor_identities():
    from oracle import gvec_ref as G
    rng = np.random.default_rng(3)
    v = rng.normal(size=(3, 64))
    at = np.concatenate([np.zeros((3, 1)), rng.normal(size=(3, 3))], axis=1)
    s = G.gensfacgp(v, at)
    assert np.allclose(np.abs(s), 1.0, atol=1e-15, rtol=0)
    assert np.array_equal(s[:, 0], np.ones(64, dtype=complex))   # atom at the origin
    sm = G.gensfacgp(-v, at)
    assert np.abs(s * sm - 1.0).max() <= 4e-16
    q = G.gensfacgp(np.array([[math.pi / 2], [0.0], [0.0]]), np.array([[1.0], [0.0], [0.0]]))
    assert abs(q[0, 0] - 1j) <= 1e-16


def test_generator_agrees():
    from oracle import gvec_ref as G
    for w in (small("a"), workload(1, 4)):
        cr = w.crystal
        for kp in w.kpoints:
            ngp, igpig, vgpl, vgpc, gpc, tp = G.gengpvec(cr.ivg, cr.bvec, cr.ngvec, kp.vkl, kp.vkc, kp.gkmax)
            n = kp.ngp
            assert ngp == n and np.array_equal(igpig, kp.igpig[:n])
            assert np.abs(vgpc - kp.vgpc[:, :n]).max() <= 1e-13
            assert np.abs(gpc - kp.gpc[:n]).max() <= 1e-13
            assert np.abs(tp - kp.tpgpc[:, :n]).max() <= 1e-12
            sf = G.gensfacgp(vgpc, cr.atposc)
            assert np.abs(sf - kp.sfacgp[:n]).max() <= 1e-12
