"""The C-ABI library loads on the CPU and exports every entry point that
include/wcns_amd.h declares (no compute calls: there may be no GPU here)."""
import ctypes as C
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "wcns_amd.h")
LIB = os.path.join(REPO, "wcns_amd", "libwcns_amd.so")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(wcns_\w+)\s*\(", txt, re.M)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from wcns_amd.solver import EXPORTS
    assert sorted(EXPORTS) == declared()


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_create_without_gpu_fails_cleanly():
    """On a host without a GPU, wcns_create reports an error instead of crashing."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from wcns_amd.solver import Params, load_library
    L = load_library()
    h = C.c_void_p()
    p = Params(moo=0.8, reynolds=1e5, tref=288.15, twall=-1, cfl=5, timedt_rate=1.1, efix=0.1, csrv=1,
               gama=1.4, prl=0.72, prt=0.9, nvis=0, nlamtur=-1, nscheme=41, nflux=1, ntmst=0)
    rc = L.wcns_create(C.byref(h), 0, C.byref(p))
    assert rc < 0
