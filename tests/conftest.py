import os
import sys
from functools import lru_cache

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built library")
    config.addinivalue_line("markers", "slow: longer CPU-side checks")


@lru_cache(maxsize=None)
def workload(cfg, nk=None):
    from exciting_amd import synth
    return synth.make_config(cfg, nk=nk)


@lru_cache(maxsize=None)
def small(key):
    """Tiny workloads with different tile tails and lo layouts (brute-force sizes)."""
    from exciting_amd import synth
    specs = {
        "a": dict(natm_per_species=(1, 1), lmax=3, lorbl=(0, 1), n_target=40, nk=2, seed=7),
        "b": dict(natm_per_species=(2, 1), lmax=2, lorbl=(0, 0, 1, 1), n_target=77, nk=2, seed=11),
        "c": dict(natm_per_species=(1,), lmax=4, lorbl=(1,), n_target=150, nk=1, seed=13, rmt=(1.3,)),
        "d": dict(natm_per_species=(2, 2), lmax=5, lorbl=(0, 1, 2), n_target=300, nk=1, seed=17,
                  cell=(9.0, 9.5, 10.0)),
    }
    return synth.make_small(**specs[key])


def iora_terms(cr, seed, scale=1e-2):
    """Seeded 'iora*' overlap terms for a crystal: h1aa symmetric in (io1, io2) with I + h1aa
    positive definite, h1loa and h1lolo (h1lolo symmetric, as olprad makes it), and a Hermitian
    m2effig (m2effig(-G) = conj m2effig(G)): cfunig times 1 + 0.3 cos(0.9 |G|)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ao, na = cr.apwordmax, cr.natm
    h = rng.uniform(-scale, scale, (ao, ao, cr.lmaxapw + 1, na))
    h1aa = 0.5 * (h + h.transpose(1, 0, 2, 3))
    nlo = cr.oalo.shape[1]
    h1loa = rng.uniform(-scale, scale, (ao, nlo, na))
    q = rng.uniform(-scale, scale, (nlo, nlo, na))
    h1lolo = 0.5 * (q + q.transpose(1, 0, 2))
    m2effig = cr.cfunig[:cr.ngvec] * (1.0 + 0.3 * np.cos(0.9 * cr.gc[:cr.ngvec]))
    return dict(h1aa=h1aa, h1loa=h1loa, h1lolo=h1lolo, m2effig=m2effig)
