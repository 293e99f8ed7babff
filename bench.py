"""Throughput of SNAC's explicit time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one explicit time step of the whole subdomain (strain rate, stress,
force gather, node update, geometry, dt) on synthetic input of the named
workload.  Default workload: Mohr-Coulomb extension, 256 x 64 x 128 hex
elements (x, y vertical, z) over 400 x 100 x 200 km, FP64.

Prints one JSON line: element-steps/s over exactly K timed steps (after W
warm-up steps, device synchronised on both sides), the roofline of the
dominant kernel from HIP events recorded live on the library's stream, and the
CPU oracle timed on this host on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def make_problem(name, elements=""):
    from snac_amd import problem as P
    kw = {}
    if elements:
        kw["elements"] = tuple(int(v) for v in elements.lower().split("x"))
    if name not in P.WORKLOADS:
        raise SystemExit(f"unknown workload {name}")
    return P.WORKLOADS[name](**kw)


def problem_is_plastic(problem):
    from snac_amd import _abi
    return any(m.rheology & (_abi.RHEOLOGY_PLASTIC | _abi.RHEOLOGY_VISCOPLASTIC) for m in problem.materials)


def cpu_baseline(problem, seconds, threads, max_steps=200, st_steps=1):
    """The CPU oracle (C restatement of the reference step, gcc -O2
    -ffp-contract=off, OpenMP over the element and node loops) timed on this
    host: TimeStepZero, one warm-up step, then whole steps on `threads`
    threads until `seconds` of wall time have passed (at least two), then
    `st_steps` steps on one thread.  Returns ((rate, seconds, steps) for the
    threads, (rate, seconds, steps) for one thread)."""
    from oracle.oracle import Oracle
    o = Oracle(problem, threads=threads)
    if problem.hydrostatic:
        o.hydrostatic_ic()
    o.timestep_zero()
    o.step(1)
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        o.step(1)
        steps += 1
        if steps >= 2 and time.perf_counter() - t0 >= seconds:
            break
    t1 = time.perf_counter()
    Oracle.set_threads(1)
    t2 = time.perf_counter()
    o.step(st_steps)
    t3 = time.perf_counter()
    o.close()
    E = problem.n_elements
    return (E * steps / (t1 - t0), t1 - t0, steps), (E * st_steps / (t3 - t2), t3 - t2, st_steps)


def cpu_bc_scan_cost(steps=200):
    """The reference applies velocity BCs through a linear search of the BC
    index table for every node (VariableCondition.c:459-463, O(N B) per
    step); the oracle's default is an O(1) per-node table.  Both timed on one
    thread on the 16 x 8 x 16 elastic box."""
    from oracle.oracle import Oracle
    from snac_amd import problem as P
    out = {}
    for faithful in (False, True):
        p = P.elastic_box_compression()
        o = Oracle(p, threads=1)
        o.set_bc_scan(faithful)
        o.timestep_zero()
        t0 = time.perf_counter()
        o.step(steps)
        out["linear_scan_ms_per_step" if faithful else "table_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / steps
        o.close()
    out["workload"] = "elastic_box_compression (16, 8, 16), one thread"
    return out


def adjoint_bench(make, param, steps, interval, flags=0):
    """Forward time of `steps` plain steps vs the adjoint's reverse sweep
    (segment replays + reverse kernels) on the same inputs."""
    import numpy as np
    from snac_amd import Context
    from snac_amd.autograd import top_nodes
    p = make()
    c = Context(p).Construct().Build().Initialise()
    c.step(2)
    c.synchronize()
    c.Destroy()
    c = Context(p).Construct().Build().Initialise()
    c.synchronize()
    t0 = time.perf_counter()
    c.step(steps)
    c.synchronize()
    plain_ms = 1e3 * (time.perf_counter() - t0)
    c.Destroy()
    c = Context(p).Construct().Build().Initialise()
    top = top_nodes(p)
    x0 = c.state(fields=("coord",))["coord"]
    c.adjoint_begin(param, interval, flags)
    c.adjoint_forward(steps)
    x = c.state(fields=("coord",))["coord"]
    cb = np.zeros((p.n_nodes, 3))
    cb[top, 1] = x[top, 1] - x0[top, 1]
    t0 = time.perf_counter()
    g = c.adjoint_reverse(cb)
    sweep_ms = 1e3 * (time.perf_counter() - t0)
    t = c.adjoint_times()
    c.adjoint_end()
    c.Destroy()
    return {
        "workload": p.name, "elements": list(p.elements), "steps": steps, "checkpoint_interval": interval,
        "forward_ms": plain_ms, "recorded_forward_ms": t["forward_ms"], "replay_ms": t["replay_ms"],
        "reverse_ms": t["reverse_ms"], "sweep_wall_ms": sweep_ms,
        "ratio": (t["replay_ms"] + t["reverse_ms"]) / plain_ms,
        "ratio_reverse_kernels": t["reverse_ms"] / plain_ms,
        "grad_norm": float(np.linalg.norm(g)),
        "flags": flags,
    }


def _ensemble_times(probs, steps, graph_steps):
    """Seconds for `steps` steps of the members in graph launches; then with
    per-member launches from one host thread, each member on its own stream
    (members overlap on the GPU as far as its hardware queues allow); then
    serialised, each member's steps finished before the next member starts;
    and the autotuned element variants."""
    from snac_amd import Context
    from snac_amd.ensemble import Ensemble
    members = [Context(p).Construct().Build().Initialise() for p in probs]
    ens = Ensemble(members, graph_steps=graph_steps, autotune=True)
    variants = [t["elem_variant"] for t in ens.tuning]
    ens.step(graph_steps).synchronize()
    t0 = time.perf_counter()
    ens.step(steps).synchronize()
    t_ens = time.perf_counter() - t0
    ens.destroy()
    t0 = time.perf_counter()
    for m in members:
        m.step(steps)
    for m in members:
        m.synchronize()
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    for m in members:
        m.step(steps)
        m.synchronize()
    t_ser = time.perf_counter() - t0
    for m in members:
        m.Destroy()
    return t_ens, t_seq, t_ser, variants


def ensemble_bench(n_models, steps, graph_steps=10):
    """Aggregate element-steps/s of an ensemble of uneven elasto-plastic
    models: one graph launch per graph_steps steps (members on parallel
    branches) against the same members launched member by member (on their
    own streams, and serialised); FP64, then the same members with FP32 bulk
    arithmetic and FP64 state."""
    import copy
    from snac_amd import _abi
    from snac_amd.ensemble import elastoplastic_ensemble
    probs = elastoplastic_ensemble(n_models=n_models)
    total_e = sum(p.n_elements for p in probs)
    t_ens, t_seq, t_ser, variants = _ensemble_times(probs, steps, graph_steps)
    out = {"models": n_models, "elements_total": total_e, "steps": steps, "graph_steps": graph_steps,
           "value": total_e * steps / t_ens, "unit": "element-steps/s",
           "sequential_value": total_e * steps / t_seq, "speedup_vs_sequential": t_seq / t_ens,
           "serialised_value": total_e * steps / t_ser, "speedup_vs_serialised": t_ser / t_ens,
           "elem_variants": {"tet_per_lane": variants.count(0), "element_per_lane": variants.count(1),
                             "element_per_lane_lowreg": variants.count(2), "tet_per_wave": variants.count(3)}}
    mixed = [copy.deepcopy(p) for p in probs]
    for p in mixed:
        p.precision = _abi.PRECISION_MIXED
    m_ens, m_seq, m_ser, m_variants = _ensemble_times(mixed, steps, graph_steps)
    out["mixed"] = {"value": total_e * steps / m_ens, "sequential_value": total_e * steps / m_seq,
                    "serialised_value": total_e * steps / m_ser,
                    "vs_fp64": t_ens / m_ens,
                    "elem_variants": {"tet_per_lane": m_variants.count(0),
                                      "tet_per_wave": len(m_variants) - m_variants.count(0)}}
    out["adjoint"] = ensemble_adjoint_bench(probs, steps)
    return out


def ensemble_adjoint_bench(probs, steps, interval=25, snapshots=8):
    """Forward + adjoint of every member (cohesion gradient of a top-surface
    misfit, corner clamp, binomial checkpointing: blocks of `interval` steps,
    `snapshots` states beside the first): members on their own streams from
    one host thread each, against the members run one after the other."""
    import numpy as np
    from snac_amd import Context, _abi
    from snac_amd.autograd import top_nodes
    from snac_amd.ensemble import adjoint_gradients
    rng = np.random.default_rng(7)
    targets = [1e-3 * rng.standard_normal(len(top_nodes(p))) for p in probs]
    total_e = sum(p.n_elements for p in probs)
    members = [Context(p).Construct().Build().Initialise() for p in probs]
    t0 = time.perf_counter()
    res = adjoint_gradients(members, _abi.PARAM_COHESION, steps, targets, interval, _abi.ADJ_CORNER_CLAMP,
                            snapshots=snapshots)
    t_con = time.perf_counter() - t0
    for m in members:
        m.Destroy()
    t_seq = 0.0
    for k, p in enumerate(probs):
        c = Context(p).Construct().Build().Initialise()
        t0 = time.perf_counter()
        adjoint_gradients([c], _abi.PARAM_COHESION, steps, [targets[k]], interval, _abi.ADJ_CORNER_CLAMP, threads=1,
                          snapshots=snapshots)
        t_seq += time.perf_counter() - t0
        c.Destroy()
    return {"steps": steps, "checkpoint_interval": interval, "snapshots": snapshots,
            "state_bytes_per_element": (snapshots + 1) * 560 + interval * 750,
            "forward_plus_adjoint_element_steps_per_s": total_e * steps / t_con,
            "sequential_element_steps_per_s": total_e * steps / t_seq, "speedup_vs_sequential": t_seq / t_con,
            "finite_gradients": all(bool(np.all(np.isfinite(g))) for _, g in res)}


def mixed_c3_bench(steps=200, warmup=10):
    """The mixed-precision workload of the metric: Maxwell viscoelastic with
    a log-normal viscosity field, 192 x 48 x 192 elements over 300 x 75 x
    300 km.  FP64 and FP32-bulk/FP64-state runs of the same inputs, each
    autotuned, warmed up and timed over `steps` steps; the error is the
    largest top-surface uplift difference relative to the largest uplift
    after those steps, and likewise for stress and velocity."""
    import numpy as np
    from snac_amd import Context, _abi
    from snac_amd import problem as P
    from snac_amd.autograd import top_nodes
    res = {}
    for prec in ("fp64", "mixed"):
        p = P.maxwell_viscoelastic()
        p.precision = _abi.PRECISION_MIXED if prec == "mixed" else _abi.PRECISION_FP64
        c = Context(p).Construct().Build().Initialise()
        x0 = c.state(fields=("coord",))["coord"]
        tune = c.autotune()
        c.step(warmup)
        c.synchronize()
        t0 = time.perf_counter()
        c.step(steps)
        c.synchronize()
        dt = time.perf_counter() - t0
        st = c.state(fields=("coord", "velocity", "stress"))
        c.Destroy()
        top = top_nodes(p)
        res[prec] = {"rate": p.n_elements * steps / dt, "ms_per_step": 1e3 * dt / steps,
                     "uplift": st["coord"][top, 1] - x0[top, 1], "st": st, "variant": tune["elem_variant"]}
    a, b = res["fp64"], res["mixed"]
    rel = lambda x, y: float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
    return {"workload": "maxwell_viscoelastic", "elements": [192, 48, 192], "steps": steps + warmup,
            "timed_steps": steps, "fp64_value": a["rate"], "mixed_value": b["rate"],
            "unit": "element-steps/s", "mixed_vs_fp64": b["rate"] / a["rate"],
            "fp64_ms_per_step": a["ms_per_step"], "mixed_ms_per_step": b["ms_per_step"],
            "error_vs_fp64": {"topography": rel(b["uplift"], a["uplift"]),
                              "stress": rel(b["st"]["stress"], a["st"]["stress"]),
                              "velocity": rel(b["st"]["velocity"], a["st"]["velocity"])},
            "error_on_fp64_schedule": mixed_error_on_schedule(steps + warmup),
            "elem_variants": {"fp64": a["variant"], "mixed": b["variant"]}}


def mixed_error_on_schedule(steps, seg=50):
    """The mixed run's arithmetic error alone: the mixed context replays the
    FP64 run's own dt, speed of sound and damping branch of every node
    component (snh_record_schedule / snh_set_schedule, in segments of `seg`
    steps), so that no discrete decision of the damped map differs; relative
    to the FP64 run's largest uplift, stress and velocity."""
    import numpy as np
    from snac_amd import Context, _abi
    from snac_amd import problem as P
    from snac_amd.autograd import top_nodes
    ctx = {}
    for prec in ("fp64", "mixed"):
        p = P.maxwell_viscoelastic()
        p.precision = _abi.PRECISION_MIXED if prec == "mixed" else _abi.PRECISION_FP64
        ctx[prec] = Context(p).Construct().Build().Initialise()
    x0 = ctx["fp64"].state(fields=("coord",))["coord"]
    for s0 in range(0, steps, seg):
        n = min(seg, steps - s0)
        ctx["mixed"].set_schedule(*ctx["fp64"].record_schedule(n, damping=True))
        ctx["mixed"].Execute(n)
    a = ctx["fp64"].state(fields=("coord", "velocity", "stress"))
    b = ctx["mixed"].state(fields=("coord", "velocity", "stress"))
    for c in ctx.values():
        c.Destroy()
    top = top_nodes(p)
    rel = lambda x, y: float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
    return {"steps": steps, "topography": rel(b["coord"][top, 1] - x0[top, 1], a["coord"][top, 1] - x0[top, 1]),
            "displacement": rel(b["coord"] - x0, a["coord"] - x0),
            "stress": rel(b["stress"], a["stress"]), "velocity": rel(b["velocity"], a["velocity"])}


def _pmc_child_cmd(args):
    cmd = [sys.executable, os.path.abspath(__file__), "--steps", "3", "--warmup", "1", "--no-cpu", "--no-adjoint",
           "--no-ensemble", "--no-traffic", "--no-mixed-c3", "--workload", args.workload, "--precision", args.precision]
    if args.elements:
        cmd += ["--elements", args.elements]
    return cmd


def _pmc_run(counters, args):
    """Per-dispatch counter values {kernel name: [value per dispatch]} of a
    short bench (3 steps) under rocprofv3 --pmc, or None."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        sys.stderr.write("bench.py: rocprofv3 not found; roofline.traffic left empty\n")
        return None
    d = tempfile.mkdtemp(prefix="snac_pmc_")
    cmd = ["timeout", "-s", "KILL", "150", prof, "--pmc"] + list(counters) + ["-d", d, "-o", "pmc", "-f", "csv",
                                                                             "--"] + _pmc_child_cmd(args)
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=170)
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            sys.stderr.write(f"bench.py: rocprofv3 {counters} gave no counters (rc {r.returncode})\n")
            return None
        per = {}
        for fn in files:
            with open(fn) as f:
                for row in csv.DictReader(f):
                    key = (row["Dispatch_Id"], row["Kernel_Name"])
                    per.setdefault(key, {}).setdefault(row["Counter_Name"], 0.0)
                    per[key][row["Counter_Name"]] += float(row["Counter_Value"])
        out = {}
        for (_, name), vals in per.items():
            out.setdefault(name, []).append(vals)
        return out
    except (OSError, subprocess.SubprocessError, KeyError, ValueError) as exc:
        sys.stderr.write(f"bench.py: rocprofv3 {counters} failed: {exc}\n")
        return None
    finally:
        shutil.rmtree(d, ignore_errors=True)


def _kernel_rows(res, prefix):
    """dispatches of the most-dispatched kernel whose name starts with prefix
    (the variant the timed steps ran; the autotuner launches the others a
    few times each)"""
    names = [n for n in res if n.split("(")[0].split("<")[0].strip().endswith(prefix)]
    if not names:
        return None, None
    name = max(names, key=lambda n: len(res[n]))
    return name, res[name]


def hbm_traffic(args, kernel_prefixes=("k_elem", "k_node", "k_geo")):
    """HBM bytes per launch of the step's kernels, from FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) scaled by what the same counters read for
    the copy probe k_copy of known size (1 GiB read, 1 GiB written per
    launch), and the VALU issue fraction from SQ_INSTS_VALU against the
    chip's 256 wave-instructions per cycle over the kernel's cycles
    (GRBM_GUI_ACTIVE summed over the 8 XCDs)."""
    out = {}
    res = {c: _pmc_run([c], args) for c in ("FETCH_SIZE", "WRITE_SIZE")}
    if not all(res.values()):
        return None
    cal = {}
    for c in res:
        _, rows = _kernel_rows(res[c], "k_copy")
        if not rows:
            sys.stderr.write(f"bench.py: no k_copy dispatch among {sorted(res[c])[:8]}\n")
            return None
        cal[c] = (1 << 30) / (1024.0 * sum(r[c] for r in rows) / len(rows))
    valu = _pmc_run(["SQ_INSTS_VALU", "GRBM_GUI_ACTIVE"], args)
    for pre in kernel_prefixes:
        name_f, rf = _kernel_rows(res["FETCH_SIZE"], pre)
        name_w, rw = _kernel_rows(res["WRITE_SIZE"], pre)
        if not rf or not rw:
            continue
        fetch = 1024.0 * sum(r["FETCH_SIZE"] for r in rf) / len(rf) * cal["FETCH_SIZE"]
        write = 1024.0 * sum(r["WRITE_SIZE"] for r in rw) / len(rw) * cal["WRITE_SIZE"]
        k = {"kernel": name_f, "read_bytes": fetch, "write_bytes": write, "bytes": fetch + write}
        if valu:
            name_v, rv = _kernel_rows(valu, pre)
            if rv:
                insts = sum(r["SQ_INSTS_VALU"] for r in rv) / len(rv)
                cycles = sum(r["GRBM_GUI_ACTIVE"] for r in rv) / len(rv) / 8.0
                k["valu_issue_frac"] = insts / (256.0 * cycles) if cycles > 0 else None
        out[pre] = k
    out["calibration"] = {"fetch_bytes_per_kib": 1024.0 * cal["FETCH_SIZE"],
                          "write_bytes_per_kib": 1024.0 * cal["WRITE_SIZE"],
                          "probe": "k_copy, 1 GiB read + 1 GiB written per launch"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000,
                    help="timed steps (default: the Mohr-Coulomb extension's full 2000-step run)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="mohr_coulomb_extension")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="length of the CPU oracle sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores given to this job")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--topo-kappa", type=float, default=0.0,
                    help="top-surface diffusivity (the reference's default input uses 5e-6; 0 = off)")
    ap.add_argument("--elements", default="", help="override the mesh, e.g. 256x64x128")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward timing")
    ap.add_argument("--adjoint-steps", type=int, default=500)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 64-model ensemble line")
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter runs behind roofline.traffic")
    ap.add_argument("--no-mixed-c3", action="store_true", help="skip the mixed-precision Maxwell line")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only the 64-model ensemble line (for profiling it), printed as JSON")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the step runs one subdomain on one GPU; --gpus {args.gpus} is not supported "
              "(the split over several GPUs is designed, not implemented).", file=sys.stderr)
        sys.exit(2)
    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_bench(64, args.ensemble_steps)}))
        return

    # counter runs first, in child processes, before this process touches the GPU
    traffic = None if args.no_traffic else hbm_traffic(args)

    from snac_amd import Context, _abi, metrics
    problem = make_problem(args.workload, args.elements)
    problem.precision = _abi.PRECISION_MIXED if args.precision == "mixed" else _abi.PRECISION_FP64
    problem.topo_kappa = args.topo_kappa
    E = problem.n_elements

    ctx = Context(problem).Construct().Build().Initialise()
    tune = ctx.autotune()
    ctx.step(args.warmup)
    ctx.synchronize()

    ctx.profile(True)
    t0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.synchronize()
    t1 = time.perf_counter()
    kt = ctx.kernel_times()
    ctx.profile(False)
    sc = ctx.scalars()
    # host round trip of the state (what a dump or a host remesh costs): the
    # step itself never leaves HBM
    fields = ("coord", "velocity", "stress") + (("plastic_strain",) if problem_is_plastic(problem) else ())
    t2 = time.perf_counter()
    st = ctx.state(fields=fields)
    t3 = time.perf_counter()
    ctx.set_state(**st)
    ctx.synchronize()
    t4 = time.perf_counter()
    xfer_bytes = sum(a.nbytes for a in st.values())
    ctx.Destroy()

    wall = t1 - t0
    value = E * args.steps / wall
    ms_per_step = 1e3 * wall / args.steps

    # dominant kernel and its roofline
    kinds = ("elem", "node", "geo", "fin") + (("heat_elem", "heat_node") if problem.thermal is not None else ())
    per = {k: kt[k + "_ms"] / max(1, kt[k + "_n"]) for k in kinds}
    dom = max((k for k in kinds if k != "fin"), key=lambda k: per[k])
    dom_bytes = {"elem": metrics.elem_kernel_bytes_per_element(problem) * E,
                 "node": metrics.node_kernel_bytes_per_element(problem) * E,
                 "geo": 24.0 * problem.n_nodes,
                 "heat_elem": metrics.heat_elem_kernel_bytes_per_element(problem) * E,
                 "heat_node": metrics.heat_node_kernel_bytes_per_element(problem) * E}[dom]
    achieved = dom_bytes / (per[dom] * 1e-3) / 1e9
    lib = _abi.load_library()
    bw = C.c_double(0.0)
    rc = lib.snh_measure_copy_bandwidth(1 << 30, 20, C.byref(bw))
    measured_peak = bw.value if rc == 0 else None
    step_bytes = metrics.step_bytes_per_element(problem)
    kernel_sum_ms = sum(kt[k + "_ms"] for k in kinds) / args.steps

    out = {
        "metric": "element-steps/s",
        "value": value,
        "unit": "element-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": problem.name, "elements": list(problem.elements),
                   "nodes": list(problem.nodes), "precision": args.precision, "topo_kappa": args.topo_kappa},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_" + dom,
            "achieved": achieved,
            "peak": metrics.HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / metrics.HBM_PEAK_GBPS,
            "traffic": (traffic or {}).get("k_" + dom, {}).get("bytes"),
            "bytes_per_launch": dom_bytes,
            "avg_launch_ms": per[dom],
            "measured_copy_peak": measured_peak,
        },
        "step_roofline": {
            "bytes_per_element_step": step_bytes,
            "achieved": value * step_bytes / 1e9,
            "frac": value * step_bytes / 1e9 / metrics.HBM_PEAK_GBPS,
            "kernel_ms_per_step": {k: per[k] for k in per},
            "kernel_sum_ms_per_step": kernel_sum_ms,
        },
        "final_dt": sc["dt"],
        "host_transfer": {
            "fields": list(fields), "bytes": xfer_bytes,
            "download_ms": 1e3 * (t3 - t2), "upload_ms": 1e3 * (t4 - t3),
            "value_with_state_round_trip_every_100_steps":
                E * 100 / (100 * wall / args.steps + (t4 - t2)),
        },
        "autotune": tune,
    }
    if traffic and "k_" + dom in traffic:
        out["roofline"]["valu_issue_frac"] = traffic["k_" + dom].get("valu_issue_frac")
    if traffic:
        out["counters"] = traffic
        # HBM bytes the counters saw per element-step over the step's kernels
        # (k_fin, one block, is left out), beside the compulsory bytes
        measured = sum(traffic[k]["bytes"] for k in ("k_elem", "k_node", "k_geo") if k in traffic)
        out["step_roofline"]["measured_bytes_per_element_step"] = measured / E
        out["step_roofline"]["measured_over_compulsory"] = measured / E / step_bytes

    if not args.no_adjoint and args.precision == "fp64":
        from snac_amd import problem as P
        out["adjoint"] = [
            adjoint_bench(lambda: P.maxwell_viscoelastic(elements=(64, 32, 64)), _abi.PARAM_LOG_VISCOSITY,
                          args.adjoint_steps, 50),
            adjoint_bench(lambda: P.mohr_coulomb_shear(elements=(64, 32, 64)), _abi.PARAM_COHESION,
                          args.adjoint_steps, 50, _abi.ADJ_CORNER_CLAMP),
        ]

    if not args.no_mixed_c3 and args.precision == "fp64":
        out["mixed_c3"] = mixed_c3_bench()

    if not args.no_ensemble and args.precision == "fp64":
        out["ensemble"] = ensemble_bench(64, args.ensemble_steps)

    if not args.no_cpu and args.cpu_seconds > 0:
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        (cval, csec, csteps), (sval, ssec, ssteps) = cpu_baseline(problem, args.cpu_seconds, threads)
        out["cpu_baseline"] = {
            "value": cval, "unit": "element-steps/s", "cores": threads, "kind": "port",
            "sample": f"{problem.name} {problem.elements} full mesh, {csteps} timed steps after "
                      f"TimeStepZero + 1 warm-up step ({csec:.1f} s), CPU oracle gcc -O2 -ffp-contract=off "
                      f"OpenMP on {cpu_model()}",
            "nproc": os.cpu_count(),
            "single_thread": {"value": sval, "steps": ssteps, "seconds": ssec},
            "bc_application": cpu_bc_scan_cost(),
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
