"""Throughput of the WCNS pseudo-time step on one MI355X.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload bump|cart|bump-sa|bump-euler|bump-slip|mb64-sa|mb64-sst]
                  [--lhs fp64|fp32] [--launch graph|plain]

Workloads (synthetic inputs generated by wcns_amd.cases, seeded):
  bump  256x128x64-node sinusoidal-bump channel (bump 0.05 L, wall spacing 1e-5 L,
        growth 1.1), laminar Navier-Stokes Re 1e5, Mach 0.5, Roe flux, WCNS-E5
        (scheme 41), adiabatic no-slip wall, far field elsewhere, periodic span
  cart  256x128x64-node Cartesian box, i/k periodic, far field in j, Euler,
        Steger-Warming (the inviscid path on the same node count)

A step is the whole pseudo-time step of src/Main.f90:85-109: ghost fill, speed
of sound/temperature, spectral radii and local dt, viscous and inviscid
residual, interface averaging, LU-SGS + SGS, limited update, interface
primitive averaging and the residual monitor.  Mcell-iter/s = nodes x steps /
seconds / 1e6, inputs resident in HBM before timing starts.

The timed steps are replayed from a captured HIP graph (--launch graph, the
default) or launched one by one (--launch plain).  The per-kernel times come
from a few further plain steps with HIP events recorded on the library's own
stream; the dominant kernel's roofline entry is its algorithmic HBM bytes per
launch over its mean launch time, against 8 TB/s.

"cpu_baseline" is the C oracle on all host cores; "cpu_reference" times the
reference's own serial step routines (built from its sources into
oracle/_ref/libref_step.so) on one core, on the same workload.

The line's "equal_drop" object is the metric's equal-residual-drop
comparison: the steady workload bump-euler (below) at CFL 20, run once with
the FP64 and once with the FP32 implicit operator until the monitor has
dropped 8 orders, graph replays of 100 steps between monitor reads; the wall
time to that drop of each and their ratio (--no-equal-drop skips it).

Extra workloads: bump-sa (the bump channel with the Spalart-Allmaras model),
bump-euler (the bump channel at wall spacing 1e-3 and growth 1.08, Euler,
with a slip top face instead of far field: the steady workload of --converge),
bump-slip (the laminar bump channel with the slip top), mb64-sa (64
uneven blocks, sides 24-96, SA, about 30 M nodes, inflow / outflow in i) and
mb64-sst (sides 30-120, SST, Re 1e7, M 0.85) and mb64-sst30 (the same at
sides 24-96, about 30 M nodes: the size whose multi-step adjoint fits one GPU
beside the solver); use --no-cpu for the 64-block cases.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per node for one launch of each single-kernel timing
# slot (one launch per block per step), FP64 words, each field counted once
# per pass.  Sweep operands are stored at the implicit operator's width.
def alg_bytes_per_node(slot, lhs_bytes):
    table = {
        # R += dF/dxi along one direction: read P 5, metric triple 3, R 5; write R 5
        "rhs_inv_i": 8 * 18, "rhs_inv_j": 8 * 18, "rhs_inv_k": 8 * 18,
        # viscous half-node fluxes along one direction: read u v w T 4, the 8
        # cross derivatives, 9 metrics, vol, mu; write 5 fluxes
        "rhs_vis_i": 8 * 28, "rhs_vis_j": 8 * 28, "rhs_vis_k": 8 * 28,
        # FLUX_DXYZ rows: read 5 fluxes and R 5, write R 5 (one launch per direction)
        "vis_rows": 8 * 15,
        # node derivatives: read u v w T, write 12
        "vis_deriv": 8 * 16,
        # sweeps (P 5, Geo 10, R 5, D 1, dQ in/out): 26 / 26 / 31 / 31 words
        "sweep_lu_l": lhs_bytes * 26, "sweep_lu_u": lhs_bytes * 26,
        "sweep_gs_l": lhs_bytes * 31, "sweep_gs_u": lhs_bytes * 31,
        # plane-major staging: read 17 FP64 fields, write them at the operator's width
        "sweep_stage": 8 * 17 + lhs_bytes * 17,
        # update: read P 5 + dQ 5, write P 5
        "update": 8 * 15,
    }
    return table.get(slot)


# Kernel of each single-kernel timing slot, as rocprofv3 names it (the i
# direction runs as instance 3: j pencils of the transposed copies).
SLOT_KERNEL = {"rhs_inv_i": "k_rhs_inv<3,", "rhs_inv_j": "k_rhs_inv<1,", "rhs_inv_k": "k_rhs_inv<2,",
               "rhs_vis_i": "k_vis_face<3>", "rhs_vis_j": "k_vis_face<1>", "rhs_vis_k": "k_vis_face<2>",
               "vis_rows": "k_vis_rows<",
               "vis_deriv": "k_vis_node_deriv", "sweep_lu_l": "k_sweep_pipe<0,", "sweep_lu_u": "k_sweep_pipe<1,",
               "sweep_gs_l": "k_sweep_pipe<2,", "sweep_gs_u": "k_sweep_pipe<3,", "sweep_stage": "k_skew_stage_all<",
               "update": "k_update"}
PROFILE_TAG = {"bump": "bump256x128x64", "cart": "cart256x128x64"}


def measured_traffic(workload, lhs, slot, sgs_paired=False):
    """HBM bytes per launch of the slot's kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of this workload (profiles/), or None.  A
    sweep slot names its exact instance (element type, paired SGS helpers)."""
    tag = PROFILE_TAG.get(workload)
    name = SLOT_KERNEL.get(slot)
    if not tag or not name:
        return None, None
    path = os.path.join(REPO, "profiles", f"{tag}_{lhs}_hbm_traffic.md")
    if not os.path.exists(path):
        return None, None
    kinds = {"sweep_lu_l": 0, "sweep_lu_u": 1, "sweep_gs_l": 2, "sweep_gs_u": 3}
    exact = slot in kinds
    if exact:
        kind = kinds[slot]
        name = (f"k_sweep_pipe<{kind},{'double' if lhs == 'fp64' else 'float'},"
                f"{'true' if sgs_paired and kind >= 2 else 'false'}>")
    key = name.replace(" ", "")
    col = None
    for line in open(path):
        cells = [c.strip().strip("`") for c in line.split("|")]
        if col is None and "HBM MB/launch (scaled)" in cells:
            col = cells.index("HBM MB/launch (scaled)")
            continue
        if col is None or len(cells) <= col:
            continue
        k = cells[1].replace(" ", "")
        if (k == key) if exact else k.startswith(key):
            try:
                return float(cells[col]) * 1e6, os.path.relpath(path, REPO)
            except ValueError:
                return None, None
    return None, None


# Whole-step algorithmic bytes per node (Euler / laminar Navier-Stokes, FP64):
# RHS 21 + LU 26 + 26 + SGS 31 + 31 + update 15 = 150 words.
STEP_BYTES_PER_NODE = 8 * 150


def build_case(workload):
    from wcns_amd import cases
    if workload == "bump":
        return cases.bump_channel_case(ni=256, nj=128, nk=64), \
            "bump-channel 256x128x64 laminar Re1e5 M0.5 Roe WCNS-E5"
    if workload == "bump-sa":
        case = cases.bump_channel_case(ni=256, nj=128, nk=64)
        case.params.nlamtur = 1
        return case, "bump-channel 256x128x64 Spalart-Allmaras Re1e5 M0.5 Roe WCNS-E5"
    if workload == "mb64-sa":
        case = cases.uneven_blocks_case(nblk=(4, 4, 4), lo=24, hi=96, seed=7, nvis=1, wall=True, nflux=3,
                                        reynolds=3.0e6, moo=0.6, periodic_i=False, ilo_bc=5, ihi_bc=6)
        case.params.nlamtur = 1
        return case, f"64 uneven blocks (sides 24-96, seed 7) {case.nodes} nodes, SA Re3e6 M0.6 Roe WCNS-E5"
    if workload == "mb64-sst":
        case = cases.uneven_blocks_case(nblk=(4, 4, 4), lo=30, hi=120, seed=7, nvis=1, wall=True, nflux=3,
                                        reynolds=1.0e7, moo=0.85, periodic_i=False, ilo_bc=5, ihi_bc=6)
        case.params.nlamtur = 2
        return case, f"64 uneven blocks (sides 30-120, seed 7) {case.nodes} nodes, SST Re1e7 M0.85 Roe WCNS-E5"
    if workload == "bump-euler":
        case = cases.bump_channel_case(ni=256, nj=128, nk=64, dy_wall=1e-3, growth=1.08, top_bc=3)
        case.params.nvis = 0
        return case, "bump-channel 256x128x64 (wall spacing 1e-3, growth 1.08) Euler M0.5 Roe WCNS-E5, slip top"
    if workload == "bump-slip":
        return cases.bump_channel_case(ni=256, nj=128, nk=64, top_bc=3), \
            "bump-channel 256x128x64 laminar Re1e5 M0.5 Roe WCNS-E5, slip top"
    if workload == "mb64-sst30":
        case = cases.uneven_blocks_case(nblk=(4, 4, 4), lo=24, hi=96, seed=7, nvis=1, wall=True, nflux=3,
                                        reynolds=1.0e7, moo=0.85, periodic_i=False, ilo_bc=5, ihi_bc=6)
        case.params.nlamtur = 2
        return case, f"64 uneven blocks (sides 24-96, seed 7) {case.nodes} nodes, SST Re1e7 M0.85 Roe WCNS-E5"
    if workload == "cart":
        return cases.cartesian_periodic_case(ni=256, nj=128, nk=64), \
            "cartesian-periodic 256x128x64 Euler M0.8 SW WCNS-E5"
    raise SystemExit(f"unknown workload {workload}")


def host_threads():
    """Host cores this process may use: the OpenMP thread setting of the
    environment when there is one, else the CPUs of its affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(case, budget_s=20.0):
    """The C oracle on a bounded sample of the same workload: whole steps of
    the full grid on all host cores (OpenMP over pencils and nodes, the sweeps
    hyperplane by hyperplane; bit-identical to its serial order) until about
    budget_s seconds have passed."""
    from oracle import oracle as O
    nthr = host_threads()
    o = O.OracleCase(case)
    o.set_threads(nthr)
    o.step(1)                        # first step outside the clock (allocation, first touch)
    t0 = time.perf_counter()
    n = 0
    while True:
        o.step(1)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 50:
            break
    dt = time.perf_counter() - t0
    o.close()
    return {"value": case.nodes * n / dt / 1e6, "unit": "Mcell-iter/s", "cores": nthr, "kind": "port",
            "sample": f"{n} full step(s) of the same {case.nodes}-node workload after one untimed step, "
                      f"{dt:.1f} s on {nthr} host threads (oracle/wcns_oracle.c, gcc -O2 -fopenmp)"}


def cpu_reference(workload, budget_s=10.0):
    """The reference's own step routines, compiled from its sources in their
    serial configuration (oracle/_ref/libref_step.so: flang -O2, one rank, one
    core), on the same workload: whole steps of boundary_all, initial_q_dq_c_t,
    timestep_tgh, r_h_s_vis, r_h_s_invis, l_h_s_tgh and residual_curve after
    one untimed step, until about budget_s seconds have passed.  It runs in a
    child process, since the reference ends its process on a fatal condition
    (Fortran STOP).  None where that library is not built or the run fails."""
    import subprocess
    from oracle import ref_step as R
    if not R.available():
        return None
    try:
        p = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-reference-child", workload,
                            "--cpu-reference-budget", str(budget_s)], capture_output=True, text=True,
                           timeout=20 * budget_s + 120)
    except subprocess.TimeoutExpired:
        return None
    lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
    return json.loads(lines[-1]) if p.returncode == 0 and lines else None


def cpu_reference_child(workload, budget_s):
    from oracle import ref_step as R
    case, _ = build_case(workload)
    if case.params.nlamtur > 0:
        print("{}")          # its SA / SST steps need the oracle's wall distance: not timed here
        return
    stages = ["boundary_all", "initial_q_dq_c_t", "timestep"] + (["rhs_vis"] if case.params.nvis == 1 else []) + \
        ["rhs_invis", "lhs"]
    r = R.RefStep(case)

    def step():
        for st in stages:
            r.stage(st)
        r.residual()
    step()
    t0 = time.perf_counter()
    n = 0
    while True:
        step()
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 20:
            break
    dt = time.perf_counter() - t0
    print(json.dumps({"value": case.nodes * n / dt / 1e6, "unit": "Mcell-iter/s", "cores": 1, "kind": "reference",
                      "sample": f"{n} full step(s) of the same {case.nodes}-node workload after one untimed step, "
                                f"{dt:.1f} s on one host core (the reference's serial step routines, flang -O2)"}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="bump")
    ap.add_argument("--lhs", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the run with the other implicit-operator precision")
    ap.add_argument("--no-tune", action="store_true", help="keep the default sweep tile shape")
    ap.add_argument("--launch", default="graph", choices=["graph", "plain"],
                    help="timed steps replayed from a captured HIP graph, or plain launches")
    ap.add_argument("--breakdown", action="store_true", help="print the per-kernel table to stderr")
    ap.add_argument("--cfl", type=float, default=None, help="override the workload's CFL number")
    ap.add_argument("--converge", type=int, default=0, metavar="STEPS",
                    help="run STEPS steps with the FP64 and with the FP32 implicit operator and report the "
                         "residual drop of each against time (the metric's equal-drop comparison)")
    ap.add_argument("--no-equal-drop", action="store_true",
                    help="skip the FP64 / FP32 runs to an 8-order drop on the steady Euler bump channel")
    ap.add_argument("--adjoint", action="store_true",
                    help="time the residual, its tangent and adjoint and the force gradient instead "
                         "(128x96x64 wall block, SA with frozen mu_t, M 0.7)")
    ap.add_argument("--multistep", type=int, default=0, metavar="STEPS",
                    help="time the reverse-mode wall-force sensitivity through STEPS whole steps of an inviscid "
                         "workload (default bump-euler), checkpoints every --every steps")
    ap.add_argument("--every", type=int, default=10, help="checkpoint interval of --multistep")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (calibrates HBM byte counters)")
    ap.add_argument("--cpu-reference-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-reference-budget", type=float, default=10.0, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_reference_child:
        return cpu_reference_child(args.cpu_reference_child, args.cpu_reference_budget)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(a multi-GPU run would shard the grid by block, one rank per GPU)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n}))
        return
    from wcns_amd.solver import Solver
    if args.adjoint:
        return adjoint_bench(args)
    if args.multistep:
        return multistep_bench(args)
    if args.converge:
        return converge_bench(args)

    case, wl_name = build_case(args.workload)
    if args.cfl:
        case.params.cfl = args.cfl
        wl_name += f" CFL {args.cfl:g}"
    s = Solver(case, device=0, lhs_precision=1 if args.lhs == "fp32" else 0)
    s.step(args.warmup, history=False)
    tile = None
    if not args.no_tune:
        tile = s.autotune(reps=2)    # sweep tile shape, timed on the warmed-up state
    if args.launch == "graph":
        s.step_graph(2)              # capture outside the clock
    s.sync()
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    if args.launch == "graph":
        s.step_graph(args.steps)
    else:
        s.step(args.steps, history=False)
    s.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    rms, rmax, loc = s.residual_curve()

    # per-kernel times: plain launches of the same step with HIP events on the
    # library's stream (graph replays carry no events)
    kt = max(2, min(args.steps, 10))
    s.timing(True)
    s.step(kt, history=False)
    s.sync()
    tim = s.timings()
    s.timing(False)

    sec = t1 - t0
    value = case.nodes * args.steps / sec / 1e6
    # dominant single-kernel slot: the one with the most time per step
    lhs_bytes = 4 if args.lhs == "fp32" else 8
    nodes = case.nodes
    priced = {k: v for k, v in tim.items() if alg_bytes_per_node(k, lhs_bytes) and v[1]}
    roof = None
    if priced:
        dom = max(priced, key=lambda k: priced[k][0])
        dms, dlaunch = priced[dom]
        # the launches of a slot in one step cover every node once (one per
        # block, or one batched launch for all blocks)
        per_launch = alg_bytes_per_node(dom, lhs_bytes) * nodes * kt / dlaunch
        avg_s = dms / dlaunch / 1e3
        ach = per_launch / avg_s / 1e9
        traffic, tsrc = measured_traffic(args.workload, args.lhs, dom, bool(s.sweep_layout()[2]))
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": tsrc, "kernel": dom,
                "alg_bytes_per_launch": per_launch, "avg_launch_ms": round(dms / dlaunch, 4)}
    if args.breakdown:
        tot = sum(v[0] for v in tim.values())
        for k, (ms, n) in sorted(tim.items(), key=lambda kv: -kv[1][0]):
            print(f"{k:12s} {ms / kt:9.3f} ms/step  {100 * ms / max(tot, 1e-9):5.1f}%  launches {n}",
                  file=sys.stderr)
    layout = s.sweep_layout()
    # the same workload with the other implicit-operator precision, timed the
    # same way: the mixed-vs-FP64 pair of the metric (the equal-drop history of
    # the two is bench.py --converge)
    other = None
    if not args.no_compare:
        s.close()
        olhs = "fp32" if args.lhs == "fp64" else "fp64"
        s2 = Solver(case, device=0, lhs_precision=1 if olhs == "fp32" else 0)
        s2.step(args.warmup, history=False)
        if not args.no_tune:
            s2.autotune(reps=2)
        if args.launch == "graph":
            s2.step_graph(2)
        s2.sync()
        torch.cuda.synchronize()
        u0 = time.perf_counter()
        if args.launch == "graph":
            s2.step_graph(args.steps)
        else:
            s2.step(args.steps, history=False)
        s2.sync()
        torch.cuda.synchronize()
        sec2 = time.perf_counter() - u0
        v2 = case.nodes * args.steps / sec2 / 1e6
        lay2 = s2.sweep_layout()
        other = {"lhs": olhs, "value": round(v2, 2), "ms_per_step": round(1e3 * sec2 / args.steps, 3),
                 "fp32_lhs_over_fp64": round(v2 / value if olhs == "fp32" else value / v2, 4),
                 "sweep_tile": list(lay2[:2]), "sgs_paired_helpers": lay2[2],
                 "residual_rms_end": s2.residual_curve()[0]}
        s2.close()
    cpu = cref = None
    if not args.no_cpu:
        cpu = cpu_baseline(case)
        cref = cpu_reference(args.workload)
    drop = None if args.no_equal_drop else equal_drop_line(args)
    line = {
        "metric": "Mcell-iter/s (RHS+LU-SGS), mixed vs FP64 at equal residual drop; % HBM peak",
        "value": round(value, 2), "unit": "Mcell-iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * sec / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.lhs == "fp64" else "f64+f32-lhs",
        "data": "synthetic (seeded generator wcns_amd.cases)",
        "config": {"workload": wl_name, "nodes": nodes, "blocks": len(case.blocks), "lhs": args.lhs,
                   "launch": args.launch, "sweep_tile": list(tile[:2]) if tile else None,
                   "sgs_paired_helpers": layout[2], "residual_rms_end": rms},
        "other_lhs_precision": other,
        "equal_drop": drop,
        "roofline": roof, "cpu_baseline": cpu, "cpu_reference": cref,
        "step_hbm_frac": round(STEP_BYTES_PER_NODE * nodes * args.steps / sec / (HBM_PEAK_GBS * 1e9), 5),
        "stage_ms_per_step": {k: round(v[0] / kt, 4) for k, v in tim.items()},
    }
    print(json.dumps(line))
    s.close()


def drop_run(case, lhs, max_steps, stop_at=None, tune=True, chunk=100):
    """Residual drop (orders of magnitude of the RMS monitor below its first
    value) against steps and seconds for one implicit-operator precision:
    one plain step, then graph replays of `chunk` steps between monitor
    reads, until `max_steps` or until the drop reaches `stop_at`."""
    import math
    import torch
    from wcns_amd.solver import Solver
    s = Solver(case, device=0, lhs_precision=1 if lhs == "fp32" else 0)
    s.step(1, history=False)
    r0 = s.residual_curve()[0]
    if tune:
        s.autotune(reps=2)
    hist, t_run, done = [], 0.0, 1
    while done < max_steps:
        n = min(chunk, max_steps - done)
        s.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.step_graph(n)
        s.sync()
        t_run += time.perf_counter() - t0
        done += n
        rms = s.residual_curve()[0]
        hist.append((done, round(t_run, 4), round(math.log10(r0 / rms), 3) if rms > 0 else None))
        print(f"converge {lhs}: {done} steps, {t_run:.1f} s, drop {hist[-1][2]}", file=sys.stderr, flush=True)
        if stop_at is not None and hist[-1][2] is not None and hist[-1][2] >= stop_at:
            break
    s.close()
    return {"rms_first": r0, "history": hist, "ms_per_step": round(1e3 * t_run / (done - 1), 3)}


def equal_drop(runs, goal=8.0):
    """The drop both precisions reach (`goal` orders when both do, else the
    deepest common one) and the step and second at which each reached it."""
    deepest = min(max(h[2] for h in runs[k]["history"] if h[2] is not None) for k in runs)
    target = min(goal, deepest)
    reach = {}
    for k, r in runs.items():
        hit = next((h for h in r["history"] if h[2] is not None and h[2] >= target), None)
        reach[k] = {"steps": hit[0], "seconds": hit[1]} if hit else None
    speed = reach["fp64"]["seconds"] / reach["fp32"]["seconds"] if reach["fp64"] and reach["fp32"] else None
    return target, deepest, reach, speed


def converge_bench(args):
    """Residual drop against steps and seconds, FP64 and FP32 implicit
    operators, over args.converge steps of the workload."""
    case, wl_name = build_case(args.workload)
    if args.cfl:
        case.params.cfl = args.cfl
        wl_name += f" CFL {args.cfl:g}"
    runs = {lhs: drop_run(case, lhs, args.converge, tune=not args.no_tune) for lhs in ("fp64", "fp32")}
    target, deepest, reach, speed = equal_drop(runs)
    print(json.dumps({"converge": {"workload": wl_name, "steps": args.converge, "equal_drop_orders": target,
                                   "deepest_common_drop": deepest,
                                   "reached": reach, "fp32_over_fp64_time_speedup": speed, "runs": runs}}))


def equal_drop_line(args, goal=8.0, max_steps=6000):
    """The metric's equal-drop comparison on the steady Euler bump channel
    (256x128x64, slip top, CFL 20): each precision runs until its monitor
    has dropped `goal` orders; the wall time to that drop, FP64 over FP32."""
    case, wl_name = build_case("bump-euler")
    case.params.cfl = 20.0
    runs = {lhs: drop_run(case, lhs, max_steps, stop_at=goal, tune=not args.no_tune) for lhs in ("fp64", "fp32")}
    target, _, reach, speed = equal_drop(runs, goal)
    return {"workload": wl_name + " CFL 20", "drop_orders": target, "reached": reach,
            "fp32_over_fp64_time_speedup": round(speed, 4) if speed else None,
            "ms_per_step": {k: r["ms_per_step"] for k, r in runs.items()}}


def adjoint_bench(args):
    """ms per call of wcns_adj_residual / _jvp / _vjp / _force (with gradient)
    on the adjoint workload, inputs resident on the device."""
    import torch
    from wcns_amd import adjoint as A, cases
    from wcns_amd.solver import Solver
    case = cases.bump_channel_case(ni=128, nj=96, nk=64, dy_wall=1e-4, growth=1.08, reynolds=3e6, moo=0.7)
    case.params.nlamtur = 1
    s = Solver(case)
    s.step(max(args.warmup, 1), history=False)
    s.boundary_all()
    s.sync()
    P, X = A.state_tensors(s, 0)
    g = torch.Generator().manual_seed(1)
    dP = (torch.randn(P.shape, generator=g, dtype=torch.float64) * 1e-3).cuda()
    dX = (torch.randn(X.shape, generator=g, dtype=torch.float64) * 1e-5).cuda()
    W = torch.randn((5,) + tuple(X.shape[1:]), generator=g, dtype=torch.float64).cuda()
    omega = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0]
    calls = {"residual": lambda: A.residual(s, 0, P, X), "jvp": lambda: A.jvp(s, 0, P, X, dP, dX),
             "vjp": lambda: A.vjp(s, 0, P, X, W), "force_grad": lambda: A.force(s, 0, P, X, omega, grad=True)}
    out = {}
    for name, f in calls.items():
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            f()
        torch.cuda.synchronize()
        out[name] = round(1e3 * (time.perf_counter() - t0) / args.steps, 3)
    print(json.dumps({"adjoint_ms_per_call": out, "nodes": case.nodes, "calls": args.steps,
                      "config": "bump-channel 128x96x64 SA (frozen mu_t) Re3e6 M0.7 Roe WCNS-E5, FP64"}))
    s.close()


def multistep_bench(args):
    """Seconds of wcns_adj_multistep: STEPS forward steps with a checkpoint
    every --every steps, J = the wall's pressure cfy, and the reverse pass
    (each segment replayed from its checkpoint, then every step's adjoint),
    against the same STEPS plain solver steps; with the HBM the checkpoints
    and work arrays take."""
    import torch
    from wcns_amd import adjoint as A
    from wcns_amd.solver import Solver
    wl = args.workload if args.workload != "bump" else "bump-euler"
    case, wl_name = build_case(wl)
    s = Solver(case)
    s.step(max(args.warmup, 1), history=False)
    P0 = A.pool_state(s)
    om = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(args.multistep, history=False)
    s.sync()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    J, g, used = A.multistep_gradient(s, P0, args.multistep, args.every, om)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    print(json.dumps({"multistep_adjoint": {"workload": wl_name, "nodes": case.nodes, "steps": args.multistep,
                                            "checkpoint_every": args.every, "J_cfy": J,
                                            "grad_norm": float(g.norm()), "seconds": round(t_adj, 3),
                                            "plain_steps_seconds": round(t_fwd, 3),
                                            "ratio_to_plain_steps": round(t_adj / t_fwd, 2),
                                            "hbm_gb_checkpoints_and_work": round(used, 3)}}))
    s.close()


if __name__ == "__main__":
    main()
