import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def ref(oracle):
    if not oracle.have_ref():
        pytest.skip("reference build oracle/_ref/libafmref.so not present")
    return oracle.ref()


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return torch.device("cuda", 0)


@pytest.fixture(scope="session")
def afm():
    import afmibins_amd
    afmibins_amd.load()
    return afmibins_amd


_ctx_cache = {}


@pytest.fixture
def ctx_of(afm, dev):
    def get(n):
        if n not in _ctx_cache:
            _ctx_cache[n] = afm.Context(n, 0)
        return _ctx_cache[n]
    return get
