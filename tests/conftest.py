import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "difx-svn2git-try3_amd")
for p in (PKG, os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    import fxcorr
    return fxcorr.load_oracle()


@pytest.fixture(scope="session")
def gpu():
    """(lib, ctx) for the HIP library on device 0; fails loudly if not built."""
    import ctypes as C
    import fxcorr
    lib = fxcorr.load_gpu_lib()
    ctx = C.c_void_p()
    fxcorr.check(lib, lib.fxg_ctx_create(0, C.byref(ctx)), "fxg_ctx_create")
    yield lib, ctx
    lib.fxg_ctx_destroy(ctx)
