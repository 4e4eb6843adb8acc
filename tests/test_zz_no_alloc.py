"""wcns_residual_noalloc (scratch in a fixed device array) gives the same
monitor as wcns_residual (scratch allocated at set-up), bit for bit, on two
calls in a row, the second on a smaller case; both through the C-ABI."""
import ctypes as C

import numpy as np
import pytest

from wcns_amd import cases
from wcns_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _both(s):
    L = s.L
    out = []
    for fn in (L.wcns_residual, L.wcns_residual_noalloc):
        rms, rmax = C.c_double(), C.c_double()
        loc = np.zeros(5, dtype=np.int32)
        rc = fn(s.h, C.byref(rms), C.byref(rmax), loc.ctypes.data_as(C.POINTER(C.c_int)))
        assert rc == 0, s.L.wcns_last_error(s.h)
        out.append((rms.value, rmax.value, loc.copy()))
    return out


def test_noalloc_residual_matches_residual():
    big = cases.multiblock_case(nx=(20, 24), ny=(18, 16), nz=(14,), nvis=1, wall=True)
    small = cases.cartesian_periodic_case(ni=24, nj=16, nk=12, seed=5)
    print("noalloc: multiblock 2x2x1 viscous, then 24x16x12 Euler", flush=True)
    sb = Solver(big, device=0)
    sb.step(3, history=False)
    (r0, m0, l0), (r1, m1, l1) = _both(sb)
    assert r0 > 0 and r0 == r1 and m0 == m1 and list(l0) == list(l1)
    ss = Solver(small, device=0)
    ss.step(2, history=False)
    (r0, m0, l0), (r1, m1, l1) = _both(ss)
    assert r0 > 0 and r0 == r1 and m0 == m1 and list(l0) == list(l1)
    # the first context again, after the smaller case used the shared array
    sb.step(1, history=False)
    (r0, m0, l0), (r1, m1, l1) = _both(sb)
    assert r0 == r1 and m0 == m1 and list(l0) == list(l1)
    ss.close()
    sb.close()
