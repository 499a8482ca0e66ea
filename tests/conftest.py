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
