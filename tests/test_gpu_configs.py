"""GPU library (through its C-ABI) against the CPU oracle on the configurations
and branches that the stage-by-stage tests do not reach:

* 64 uneven blocks with SA at Mach 0.3, 0.85 and 2.0 and Re 3e6 (the 64-block
  SA workload at reduced block sides, 12-20 nodes), and with SST at Re 1e7,
  Mach 0.85 (the forward part of the multi-step adjoint workload), streamwise
  inflow (BC 5) / outflow (BC 6), a no-slip wall, far field on top, 1:1
  interfaces, periodic span: 20 steps replayed from the HIP graph;
* the bench workload itself, the 256x128x64 bump channel at wall spacing
  1e-5 and growth 1.1: 3 steps (plain, then graph replays with the tuned
  sweep tiles) against the oracle;
* the FP32 implicit operator against the FP64 oracle, with a stated bound;
* the inviscid wall (BC 2 with nvis = 0) at Mach 2;
* forced failure paths: a repaired node and the WCNS positivity fallback, the
  too-many-repairs stop (WCNS_E_REPAIR), a Roe average with c^2 <= 0
  (WCNS_E_ROE) and a folded grid (WCNS_E_JACOBIAN), each with the oracle's
  counters or error on the same input.

FP64 tolerance: relative 1e-10 against each field's scale after 20 steps (the
turbulence models' pow/tanh/sqrt differ between the GPU and host libm by a few
ulps, which the steps carry along)."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from wcns_amd import cases

pytestmark = pytest.mark.gpu

THREADS = int(os.environ.get("OMP_NUM_THREADS", "8") or 8)
# FP32 implicit operator after 30 steps of the viscous bump channel: the
# largest relative difference of any primitive from the FP64 oracle
FP32_BOUND = 5e-6


def _oracle(case):
    o = O.OracleCase(case)
    o.set_threads(THREADS)
    return o


def _names(case):
    return ["r", "u", "v", "w", "p"] + {1: ["qke", "vist"], 2: ["qke", "qkw", "vist"]}.get(case.params.nlamtur, [])


def _max_rel(s, o, case, ghost=0):
    """Largest relative difference over the interior nodes, or with ghost > 0
    also over the ghost layers that a pencil stencil reads (nodes with at most
    one index outside 1..n; edge and corner ghosts are read by nothing)."""
    worst = 0.0
    for nb, blk in enumerate(case.blocks):
        idx = np.indices((blk.nk + 6, blk.nj + 6, blk.ni + 6)) - 2
        outside = sum(((idx[a] < 1) | (idx[a] > n)).astype(int) for a, n in enumerate((blk.nk, blk.nj, blk.ni)))
        depth = np.max(np.stack([np.maximum(1 - idx[a], idx[a] - n) for a, n in enumerate((blk.nk, blk.nj, blk.ni))]),
                       axis=0)
        mask = (outside == 0) if ghost == 0 else ((outside <= 1) & (depth <= ghost))
        for n in _names(case):
            a = s.field(nb, n)[mask]
            b = o.field(nb, n)[mask]
            e = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
            worst = max(worst, e)
    return worst


def _uneven(moo, model, reynolds):
    case = cases.uneven_blocks_case(nblk=(4, 4, 4), lo=12, hi=20, seed=11, nvis=1, wall=True, nflux=3,
                                    reynolds=reynolds, moo=moo, periodic_i=False, ilo_bc=5, ihi_bc=6)
    case.params.nlamtur = model
    return case


@pytest.mark.parametrize("moo,model,reynolds", [(0.3, 1, 3e6), (0.85, 1, 3e6), (2.0, 1, 3e6), (0.85, 2, 1e7)],
                         ids=["SA-M0.3", "SA-M0.85", "SA-M2.0", "SST-M0.85"])
def test_uneven_64_blocks_graph_20_steps(moo, model, reynolds):
    from wcns_amd.solver import Solver
    case = _uneven(moo, model, reynolds)
    print(f"64 uneven blocks, {case.nodes} nodes, model {model}, M {moo}", flush=True)
    s = Solver(case)
    s.step_graph(20)
    hs = s.residual_history(20)
    o = _oracle(case)
    ho = o.step(20)
    assert hs[3] == 20
    rel = np.abs(hs[0] - ho) / ho
    err = _max_rel(s, o, case)
    print(f"residual history rel err {rel.max():.2e}, fields {err:.2e}", flush=True)
    assert rel.max() <= 1e-10
    assert err <= 1e-10
    assert s.counter(0) == o.counter(0) and s.counter(1) == o.counter(1)
    o.close()
    s.close()


def test_bump_channel_full_size_against_oracle():
    """The bench workload itself: the 256x128x64 bump channel (bump 0.05 L,
    wall spacing 1e-5 L, growth 1.1, laminar Re 1e5, M 0.5, Roe, scheme 41),
    one plain step, the tile autotuner, then 2 steps from the HIP graph with
    the tuned sweep tiles, against 3 oracle steps: the residual monitor within 1e-10, the primitives within
    1e-10 of each field's scale, the fallback and repair counters equal."""
    from wcns_amd.solver import Solver
    case = cases.bump_channel_case(ni=256, nj=128, nk=64)
    print("bump 256x128x64: GPU", flush=True)
    s = Solver(case)
    s.step(1, history=False)
    s.autotune(reps=1)           # tile shapes timed on the stepped state
    s.step_graph(2)
    hs = s.residual_history(3)[0]
    print("bump 256x128x64: oracle", flush=True)
    o = _oracle(case)
    ho = o.step(3)
    rel = np.max(np.abs(hs - ho) / ho)
    e = _max_rel(s, o, case)
    print(f"monitor rel {rel:.2e}, primitives rel {e:.2e}", flush=True)
    assert rel <= 1e-10 and e <= 1e-10
    assert s.counter(0) == o.counter(0) and s.counter(1) == o.counter(1)
    s.close()
    o.close()


def test_fp32_operator_error_bound():
    """The FP32 implicit operator leaves the FP64 residual, weights and update
    alone; after 30 steps of a viscous bump channel its primitives differ from
    the FP64 oracle's by at most FP32_BOUND relative to each field's scale, and
    its residual monitor agrees to 1e-3 relative at every step."""
    from wcns_amd.solver import Solver
    case = cases.bump_channel_case(ni=32, nj=24, nk=12, dy_wall=1e-3, growth=1.15, reynolds=1e4)
    s = Solver(case, lhs_precision=1)
    o = _oracle(case)
    hs = s.step(30)
    ho = o.step(30)
    err = _max_rel(s, o, case, ghost=0)
    hrel = float(np.max(np.abs(hs - ho) / ho))
    print(f"fp32 operator: primitives rel err {err:.3e}, monitor rel err {hrel:.3e}", flush=True)
    assert err <= FP32_BOUND
    assert hrel <= 1e-3
    o.close()
    s.close()


def test_inviscid_wall_mach2():
    """Euler, slip wall (BC 2 with nvis = 0, BC.f90:1412-1486) at Mach 2, Roe,
    2 blocks: 10 steps against the oracle."""
    from wcns_amd.solver import Solver
    case = cases.multiblock_case(nx=(14, 12), ny=(12,), nz=(12,), nvis=0, wall=True, moo=2.0, nflux=3)
    s = Solver(case)
    o = _oracle(case)
    hs = s.step(10)
    ho = o.step(10)
    e0, e3 = _max_rel(s, o, case, ghost=0), _max_rel(s, o, case, ghost=3)
    print(f"inviscid wall: monitor {np.max(np.abs(hs - ho) / ho):.2e}, interior {e0:.2e}, ghosts {e3:.2e}", flush=True)
    assert np.max(np.abs(hs - ho) / ho) <= 1e-11
    assert e0 <= 1e-11 and e3 <= 1e-11
    o.close()
    s.close()


def _poke(targets, name, nodes, value):
    for t in targets:
        for nb, i, j, k in nodes:
            a = t.field(nb, name)
            a[k + 2, j + 2, i + 2] = value
            t.set_field(nb, name, a)


def test_forced_repair_and_wcns_fallback():
    """Two nodes above rmax_limit are repaired by the ordered neighbour mean
    (Sub2.f90:661-767) and two near-vacuum pressures send WCNS to its
    positivity fallback (WCNSE5.f90:2224-2245): same fields and counters."""
    from wcns_amd.solver import Solver
    case = cases.cartesian_periodic_case(ni=24, nj=16, nk=12, seed=3)
    s = Solver(case)
    o = _oracle(case)
    _poke((s, o), "r", [(0, 6, 6, 6), (0, 15, 9, 4)], 2.0e3)
    _poke((s, o), "p", [(0, 10, 4, 8), (0, 11, 4, 8)], 1.0e-7)
    hs = s.step(2)
    ho = o.step(2)
    print("repairs", s.counter(1), o.counter(1), "fallbacks", s.counter(0), o.counter(0), flush=True)
    assert o.counter(1) >= 2 and o.counter(0) > 0
    assert s.counter(1) == o.counter(1) and s.counter(0) == o.counter(0)
    assert np.max(np.abs(hs - ho) / ho) <= 1e-10
    assert _max_rel(s, o, case, ghost=3) <= 1e-10
    o.close()
    s.close()


def test_forced_abort_reports_repair_error():
    """More than n/64 abnormal nodes: update_limited stops (Sub2.f90:785-790);
    the library returns WCNS_E_REPAIR and the oracle its error code."""
    from wcns_amd.solver import Solver, WcnsError
    case = cases.cartesian_periodic_case(ni=24, nj=16, nk=12, seed=3)
    s = Solver(case)
    o = _oracle(case)
    for t in (s, o):
        a = t.field(0, "r")
        a[4:9, 4:11, 3:12] = 2.0e3
        t.set_field(0, "r", a)
    with pytest.raises(WcnsError) as ei:
        s.step(1)
    assert ei.value.code == -8
    with pytest.raises(RuntimeError):
        o.step(1)
    assert s.counter(1) == o.counter(1) > (24 * 16 * 12) // 64
    o.close()
    s.close()


def test_expired_sweep_wait_reports_error(monkeypatch):
    """The pipelined sweeps' waits for the upstream tiles' exchange words are
    bounded.  With the bound set from the host to one re-read
    (WCNS_SPIN_LIMIT=1) the downstream tiles give up at the start of the
    sweep, every wavefront still runs to the end of it, and the step returns
    WCNS_E_HIP; a context with the default bound then steps normally."""
    from wcns_amd.solver import Solver, WcnsError
    case = cases.cartesian_periodic_case(ni=32, nj=48, nk=24, seed=3)
    monkeypatch.setenv("WCNS_SPIN_LIMIT", "1")
    s = Solver(case)
    print("sweep layout", s.sweep_layout(), flush=True)
    with pytest.raises(WcnsError) as ei:
        s.step(1)
    assert ei.value.code == -3 and "wait expired" in str(ei.value)
    s.close()
    monkeypatch.delenv("WCNS_SPIN_LIMIT")
    s2 = Solver(case)
    o = _oracle(case)
    hs = s2.step(2)
    ho = o.step(2)
    assert np.max(np.abs(hs - ho) / ho) <= 1e-12
    o.close()
    s2.close()


def test_folded_grid_reports_jacobian_error():
    """A node moved past two of its i neighbours folds the cells around it:
    check_grid_derivative stops on the non-positive Jacobian
    (PreProcess.f90:2637-2666); the library returns WCNS_E_JACOBIAN from
    wcns_setup and names the count of such nodes, the oracle fails its set-up."""
    from wcns_amd.solver import Solver, WcnsError
    case = cases.cartesian_periodic_case(ni=24, nj=16, nk=12, seed=3)
    b = case.blocks[0]
    b.x[10, 8, 6] = b.x[13, 8, 6]
    with pytest.raises(WcnsError, match="non-positive Jacobian") as ei:
        Solver(case)
    assert ei.value.code == -6
    with pytest.raises(RuntimeError, match="-11"):
        O.OracleCase(case)


def test_roe_c2_nonpositive_reports_error():
    """A negative pressure node makes a Roe average with c^2 <= 0
    (flux.f90:178-181): the library returns WCNS_E_ROE, the oracle fails its
    inviscid stage."""
    from wcns_amd.solver import Solver, WcnsError
    case = cases.cartesian_periodic_case(ni=24, nj=16, nk=12, seed=3, nflux=3)
    s = Solver(case)
    o = _oracle(case)
    _poke((s, o), "p", [(0, 12, 8, 6)], -2.0)
    with pytest.raises(WcnsError) as ei:
        s.step(1)
    assert ei.value.code == -7
    with pytest.raises(RuntimeError):
        o.step(1)
    o.close()
    s.close()
