#!/usr/bin/env python3
"""Benchmark of the explicit internal-force step on one MI355X.

Default (``--form plain``), the workload the headline metric is quoted on:
200 x 100 x 100 structured _hexahedron_8 grid (2,000,000 elements, 8 Gauss
points each = 16,000,000 element quadrature points), interior nodes jittered
by U(-0.1, 0.1) h (seed 1), compressible neo-Hookean E = 1, nu = 0.3,
rho = 1, x = 0 face clamped, initial velocity 1e-3 sin(pi x / L) U(-1, 1)^3
(seed 1), dt = 0.5 dt_crit, plain FP64 path.

One step = one SolidMechanicsModel::solveStep (predictor, internal force:
gradient -> neo-Hookean PK2 -> (B_L + B_NL)^T S -> integrate -> gather
assembly, residual, lumped solve, corrector).  Inputs are resident in HBM
before the timed region.  Prints one JSON line.

Other forms (each prints its own line):
  --form precision  1000 x 100 x 100 cubes x 6 Tet4 (60 M), per-element
                    E ~ logU(0.5, 2), nu ~ U(0.2, 0.35), linear elastic:
                    FP32 form timed, plain FP64 timed beside it
  --form batch      512 independent Tet4 boxes, 5e3-4e5 elements
                    (log-uniform), aspect U(1, 8), linear isotropic hardening
                    with per-instance parameters, FP32 form, one persistent
                    launch per substep in a hipGraph, autotuned; FP64 batch
                    timed beside it
  --form serve      Poisson stream of solve requests served by continuous
                    batching (akantu2_amd.serving)

    python bench.py [--form F] [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 (vector and matrix, FMA = 2 flops)
FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 (dense matrix; packed vector)
FP32_VALU_TFLOPS = 78.6        # MI355X FP32 vector without packed instructions
# Floating-point operations per Gauss point of the Hex8 neo-Hookean element
# pass in the reference's operation order (adds, multiplies, the division
# and the logarithm of the law counted once each): dN/dX columns 120,
# grad u 135, neo-Hookean PK2 119, (B_L + B_NL)^T S 624, times w detJ 24,
# Gauss-point sum 21.  J^-1 and w det J are read from the table that init
# fills, as the reference reads its precomputed shape derivatives, and the
# dN/dX columns of the gradient serve the force terms too.
HEX8_NEOHOOKEAN_FLOP_PER_QP = 1043
SHAPE = (200, 100, 100)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_hex8_neohookean_f64.json")
# executed flops per Gauss point of the element passes, counted with
# rocprofv3 (SQ_INSTS_VALU_FLOPS_FP32 / _FP64 x 64 lanes)
FLOPS_FILE = os.path.join(ROOT, "profiles", "flops_hex8_neohookean_200x100x100.json")
PL = "plastic_linear_isotropic_hardening"
TIMING_STEPS = 100  # steps of the event-timed per-kernel pass
RAMP_STEPS = 40     # untimed steps first: the GPU clocks rise after the idle set-up


def build_hex8_box(shape=SHAPE):
    from akantu2_amd import mesh as M
    mesh = M.box_mesh("_hexahedron_8", *shape, jitter=0.1, seed=1)
    blocked = M.clamped_face_x0(mesh)
    v0 = M.initial_velocity(mesh, 1e-3, seed=1)
    return mesh, blocked, v0


def copy_bandwidth():
    from akantu2_amd import _native
    copy_bytes = 1 << 30
    copy_ms = _native.calibration_copy(copy_bytes, repeats=20)
    return 2 * copy_bytes / (copy_ms * 1e-3) / 1e9


def rel_errors(a, b):
    """max-norm and L2 relative error of a against the reference b."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    d = a - b
    mx = float(np.abs(d).max() / max(np.abs(b).max(), 1e-300))
    l2 = float(np.linalg.norm(d) / max(np.linalg.norm(b), 1e-300))
    return {"max": mx, "l2": l2}


# --------------------------------------------------------------------------
# plain form (headline)
# --------------------------------------------------------------------------
def host_threads():
    """Host cores the CPU baseline may use (at most 16: a one-GPU host share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_plain(shape=(20, 100, 100), steps=10):
    """The CPU oracle (plain C, FP64, reference dataflow with precomputed
    shape derivatives) timed on a bounded sample of the same workload: the
    first 20 cells in x of the 200 x 100 x 100 grid, on one thread and on
    the host's cores (OpenMP element and DOF loops; results identical)."""
    import oracle
    mesh, blocked, v0 = build_hex8_box(shape)
    o = oracle.OracleModel(oracle.HEX8, mesh.nodes, mesh.conn, "neohookean")
    o.init_full()
    o.blocked[:] = blocked
    o.v[:] = v0
    o.set_time_step(0.5 * o.stable_dt())
    runs = {}
    nt = host_threads()
    for threads, nsteps in ((1, steps), (nt, 4 * steps)):
        oracle.set_threads(threads)
        t0 = time.perf_counter()
        o.solve_steps(nsteps)
        runs[threads] = (mesh.nb_elements * 8 * nsteps / (time.perf_counter() - t0),
                         time.perf_counter() - t0, nsteps)
    oracle.set_threads(1)
    v1, s1, n1 = runs[1]
    vn, sn, nn_ = runs[nt]
    return {"value": vn, "unit": "elem*qp/s", "cores": nt, "kind": "port",
            "sample": f"{nn_} steps of a {shape[0]}x{shape[1]}x{shape[2]} Hex8 slab "
                      f"({mesh.nb_elements} elements) of the same jittered neo-Hookean "
                      f"workload, oracle/akantu_oracle.c, {nt} OpenMP threads",
            "seconds": sn,
            "single_core": {"value": v1, "cores": 1, "steps": n1, "seconds": s1}}


def run_plain(args):
    from akantu2_amd import SolidMechanicsModel

    steps = args.steps or 1000
    shape = tuple(int(x) for x in args.shape.split(",")) if args.shape else SHAPE
    t = time.perf_counter()
    mesh, blocked, v0 = build_hex8_box(shape)
    model = SolidMechanicsModel(mesh)
    model.addMaterial("neohookean", rho=1.0, E=1.0, nu=0.3)
    model.setPrecision(args.precision)
    model.initFull()
    model.getBlockedDOFs()[:] = blocked
    model.getVelocity()[:] = v0
    dt = 0.5 * model.getStableTimeStep()
    model.setTimeStep(dt)
    setup_s = time.perf_counter() - t
    print(f"# setup {setup_s:.1f} s, {mesh.nb_elements} elements, "
          f"{mesh.nb_nodes} nodes, dt {dt:.6g}, {args.precision}", file=sys.stderr,
          flush=True)

    schedule = "plain"
    if args.pipeline > 1 and model.setPipeline(args.pipeline):
        schedule = f"pipelined, {args.pipeline} parts"
    # The copy calibration first: run between the kernel-timing pass and the
    # timed steps, it left the element passes of the timed steps 7 % slower.
    copy_gbs = copy_bandwidth()
    # After the idle set-up the element passes of the first ~20 steps run up
    # to 25 % slower while the GPU clocks ramp: RAMP_STEPS untimed steps, then
    # the per-kernel times, a separate pass of TIMING_STEPS steps with HIP
    # events around every launch (plain schedule), so that the timed loop
    # below carries no event overhead.
    model.solveSteps(RAMP_STEPS)
    model.synchronize()
    model.enableKernelTiming(True)
    model.solveSteps(TIMING_STEPS)
    model.synchronize()
    kern = {}
    names = {0: "element_pass", 1: "gather_update", 2: "predictor"}
    for k, name in names.items():
        ms, n = model.kernelTiming(k)
        if n:
            kern[name] = {"avg_ms": ms / n, "launches": n, "bytes": model.kernelBytes(k)}
    model.enableKernelTiming(False)
    model.solveSteps(args.warmup)
    model.synchronize()
    t0 = time.perf_counter()
    model.solveSteps(steps)
    model.synchronize()
    elapsed = time.perf_counter() - t0
    nqp = mesh.nb_elements * 8
    value = nqp * steps / elapsed
    ms_step = 1e3 * elapsed / steps
    el = kern["element_pass"]
    hbm_gbs = el["bytes"] / (el["avg_ms"] * 1e-3) / 1e9
    if args.precision == "fp64":
        flops = HEX8_NEOHOOKEAN_FLOP_PER_QP * nqp
        tflops = flops / (el["avg_ms"] * 1e-3) / 1e12
        traffic = None
        if os.path.exists(TRAFFIC_FILE):
            with open(TRAFFIC_FILE) as fh:
                tr = json.load(fh)
            if tr.get("shape") == list(shape):
                traffic = tr.get("element_pass_bytes")
        # The dominant kernel (the element pass, ~85 % of the step) is bound
        # by FP64 arithmetic, not by HBM: it forms the dN/dX columns from the
        # stored J^-1 every step instead of streaming them, and its table
        # reads run at a third of the HBM bandwidth.  MI355X's FP64 peak is
        # the same 78.6 TFLOP/s on the matrix cores and the vector ALUs; the
        # reference operation order is FMA-free, so its own ceiling is half.
        roofline = {"bound": "fp64_valu", "achieved": tflops, "peak": FP64_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": tflops / FP64_PEAK_TFLOPS,
                    "traffic": traffic,
                    "kernel": "k_hex8_strict<neo-Hookean> (FP64, reference operation order)",
                    "compute": "FP64 VALU, FMA-free by construction (ceiling 39.3 TFLOP/s)",
                    "algorithmic_flops_per_launch": flops,
                    "algorithmic_bytes_per_launch": el["bytes"],
                    "avg_kernel_ms": el["avg_ms"],
                    "hbm": {"achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": hbm_gbs / HBM_PEAK_GBS,
                            "measured_copy_gbs": copy_gbs}}
    elif args.precision == "fp32" and os.path.exists(FLOPS_FILE):
        # The sum-factorised FP32 element pass is bound by FP32 vector issue:
        # its flops are the ones the counters saw per Gauss point.
        with open(FLOPS_FILE) as fh:
            fq = json.load(fh)["f32_sf"]["flops_per_qp"]
        flops = fq * nqp
        tflops = flops / (el["avg_ms"] * 1e-3) / 1e12
        roofline = {"bound": "fp32_valu", "achieved": tflops, "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": tflops / FP32_PEAK_TFLOPS, "traffic": None,
                    "kernel": "k_hex8_bpa<neo-Hookean> (FP32, sum-factorised, block partials)",
                    "compute": "FP32 VALU; packed FP32 instructions are not used, so the "
                               f"ceiling is {FP32_VALU_TFLOPS} TFLOP/s "
                               f"({tflops / FP32_VALU_TFLOPS:.0%} of it)",
                    "flops_per_qp": fq, "algorithmic_flops_per_launch": flops,
                    "algorithmic_bytes_per_launch": el["bytes"],
                    "avg_kernel_ms": el["avg_ms"],
                    "hbm": {"achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": hbm_gbs / HBM_PEAK_GBS,
                            "measured_copy_gbs": copy_gbs}}
    else:
        roofline = {"bound": "hbm", "achieved": hbm_gbs, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": hbm_gbs / HBM_PEAK_GBS, "traffic": None,
                    "kernel": f"k_hex8_fast<neo-Hookean> ({args.precision})",
                    "algorithmic_bytes_per_launch": el["bytes"],
                    "avg_kernel_ms": el["avg_ms"], "measured_copy_gbs": copy_gbs}
    gu = kern["gather_update"]
    gu["achieved_gbs"] = gu["bytes"] / (gu["avg_ms"] * 1e-3) / 1e9
    gu["frac_hbm"] = gu["achieved_gbs"] / HBM_PEAK_GBS
    out = {
        "metric": "elem*qp stress+force updates/s",
        "value": value,
        "unit": "elem*qp/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "fp64_fast": "f64"}[args.precision],
        "data": "synthetic",
        "config": {"workload": f"Hex8 {shape[0]}x{shape[1]}x{shape[2]} jittered grid, "
                               "compressible neo-Hookean, explicit central difference "
                               "with lumped mass, "
                               + {"fp64": "plain FP64", "fp32": "FP32 form",
                                  "fp64_fast": "FP64 FMA form"}[args.precision],
                   "elements": mesh.nb_elements, "nodes": mesh.nb_nodes,
                   "quadrature_points": nqp, "dt": dt, "schedule": schedule},
        "roofline": roofline,
        "kernels": kern,
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline_plain()
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"error": repr(e)}
    if args.precision == "fp64" and not args.no_forms:
        print("# forms: FP32 form, 512-instance batch, serving stream", file=sys.stderr,
              flush=True)
        out["forms"] = run_forms(args, model, mesh, blocked, v0, dt, value,
                                 RAMP_STEPS + TIMING_STEPS + args.warmup + steps)
        out["form_targets"] = form_targets(out)
    return out


# --------------------------------------------------------------------------
# the three forms beside the headline (default run)
# --------------------------------------------------------------------------
def run_forms(args, plain, mesh, blocked, v0, dt, plain_value, plain_steps):
    """The precision, batching and serving forms measured in the same run as
    the plain FP64 line, each on its own workload, each also against the
    plain FP64 arithmetic on that same workload:
    - precision: the headline workload (200 x 100 x 100 Hex8 neo-Hookean) in
      the FP32 form, the same kernel-timing / warmup / timed steps as the
      plain line, its u, v, f_int errors against the plain model after the
      same number of steps (the plain path is bit-pinned to the oracle), then
      its throughput over max(steps, 200) more steps;
    - batch: the 512 Tet4 plasticity instances (5e3-4e5 elements, seed 4) in
      the FP32 form and in FP64 (the plain path's arithmetic), autotuned,
      hipGraph, the same timed step count;
    - serve: a Poisson stream of 24 requests at 8/s from the serving
      workload's distribution (Tet4 / Hex8, 1e3-2e6 elements, 200-5,000
      steps), continuous batching, FP32 form and the same stream in FP64:
      throughput over the makespan and over the device-busy time, p50 / p99
      time-to-result."""
    from akantu2_amd import SolidMechanicsModel
    from akantu2_amd import serving as S
    steps = args.steps or 1000
    forms = {}
    t = time.perf_counter()
    m = SolidMechanicsModel(mesh)
    m.addMaterial("neohookean", rho=1.0, E=1.0, nu=0.3)
    m.setPrecision("fp32")
    m.initFull()
    m.getBlockedDOFs()[:] = blocked
    m.getVelocity()[:] = v0
    m.setTimeStep(dt)
    m.solveSteps(RAMP_STEPS)
    m.synchronize()
    m.enableKernelTiming(True)
    m.solveSteps(TIMING_STEPS)
    m.synchronize()
    kern = {}
    for k, name in {0: "element_pass", 1: "gather_update"}.items():
        ms, n = m.kernelTiming(k)
        by = m.kernelBytes(k)
        kern[name] = {"avg_ms": ms / n, "bytes": by, "achieved_gbs": by / (ms / n * 1e-3) / 1e9}
    m.enableKernelTiming(False)
    m.solveSteps(args.warmup + steps)
    m.synchronize()
    assert RAMP_STEPS + TIMING_STEPS + args.warmup + steps == plain_steps
    err = {"u": rel_errors(m.getDisplacement(), plain.getDisplacement()),
           "v": rel_errors(m.getVelocity(), plain.getVelocity()),
           "f_int": rel_errors(m.getInternalForce(), plain.getInternalForce())}
    # the form's throughput over at least 200 steps (a 20-step window of a
    # 0.25 ms step is a few milliseconds of wall clock, within the noise),
    # after a few untimed steps: the first call after the getters above
    # releases their host copies
    timed = max(steps, 200)
    m.solveSteps(max(args.warmup, 1))
    m.synchronize()
    t0 = time.perf_counter()
    m.solveSteps(timed)
    m.synchronize()
    el = time.perf_counter() - t0
    value = mesh.nb_elements * 8 * timed / el
    forms["precision"] = {
        "workload": "the headline workload in the FP32 form (FP64 compensated nodal update)",
        "value": value, "unit": "elem*qp/s", "ms_per_step": 1e3 * el / timed,
        "timed_steps": timed, "vs_plain_fp64": value / plain_value, "kernels": kern,
        "error_vs_plain_fp64_after_steps": plain_steps, "error": err,
        "setup_s": t0 - t}
    del m
    print(f"# forms: precision {1e3 * el / timed:.3f} ms/step", file=sys.stderr, flush=True)

    specs = batch_instances(512)
    batch = {}
    for prec in ("fp32", "fp64"):
        t = time.perf_counter()
        b, ids = _fill_batch(prec, specs)
        st = b.stats()
        setup = time.perf_counter() - t
        el, bk, best = _time_batch(b, 3, steps, tune=True)
        b.close()
        batch[prec] = {"value": st["quadrature_points"] * steps / el,
                       "ms_per_step": 1e3 * el / steps, "kernels": bk, "launch": best,
                       "hbm_frac": {k: v["achieved_gbs"] / HBM_PEAK_GBS for k, v in bk.items()},
                       "setup_s": setup}
        print(f"# forms: batch {prec} {1e3 * el / steps:.3f} ms/step, setup {setup:.1f} s",
              file=sys.stderr, flush=True)
    forms["batch"] = dict(batch["fp32"])
    forms["batch"].update({
        "workload": "512 Tet4 boxes of 5e3-4e5 elements (log-uniform, seed 4), aspect "
                    "U(1, 8), linear isotropic hardening with per-instance E, nu, sigma_y, "
                    "h, FP32 form, one tile per workgroup, in a hipGraph, autotuned",
        "instances": st["live"], "elements": st["elements"], "unit": "elem*qp/s",
        "steps": steps, "vs_plain_fp64": batch["fp32"]["value"] / plain_value,
        "fp64_same_workload": batch["fp64"],
        "vs_fp64_same_workload": batch["fp32"]["value"] / batch["fp64"]["value"]})

    served = {}
    for prec in ("fp32", "fp64"):
        reqs = S.synthetic_requests(24, rate=8.0, seed=5)
        for r in reqs:
            r.materialize()
        server = S.Server(precision=prec, hbm_bytes=16 << 30)
        rep = server.run(reqs)
        server.close()
        tt = np.array([r.time_to_result for r in rep.done])
        served[prec] = {
            "value": rep.throughput, "unit": "elem*qp/s",
            "busy_throughput": sum(r.work for r in rep.done) / rep.busy if rep.busy else None,
            "latency_s": {"p50": float(np.percentile(tt, 50)),
                          "p99": float(np.percentile(tt, 99)), "max": float(tt.max())},
            "makespan_s": rep.makespan, "done": len(rep.done), "failed": len(rep.failed),
            "max_live": rep.max_live, "chunks": rep.chunks}
        print(f"# forms: serve {prec} p50 {served[prec]['latency_s']['p50']:.2f} s",
              file=sys.stderr, flush=True)
    forms["serve"] = dict(served["fp32"])
    forms["serve"].update({
        "workload": "Poisson stream of 24 requests at 8/s, Tet4 or Hex8 boxes of 1e3-2e6 "
                    "elements (log-uniform), elastic or neo-Hookean, 200-5,000 steps, FP32 "
                    "form, continuous batching with HBM-footprint admission",
        "vs_plain_fp64": served["fp32"]["busy_throughput"] / plain_value,
        "fp64_same_stream": served["fp64"],
        "vs_fp64_same_stream": served["fp32"]["busy_throughput"] /
                               served["fp64"]["busy_throughput"]})
    return forms


def form_targets(out):
    """The target stated for the forms: the FP32 batched / served form at >= 3x
    the plain FP64 path's element*qp throughput, <= 1e-5 relative
    displacement error, >= 70 % of peak HBM bandwidth.  Ratios are given
    against the headline plain line and against the plain FP64 arithmetic on
    each form's own workload (served forms: device-busy throughput)."""
    f = out["forms"]
    return {"target_ratio": 3.0, "target_hbm_frac": 0.70, "target_u_error": 1e-5,
            "plain_fp64_value": out["value"],
            "ratios_vs_headline": {k: f[k]["vs_plain_fp64"]
                                   for k in ("precision", "batch", "serve")},
            "ratios_same_workload": {"precision": f["precision"]["vs_plain_fp64"],
                                     "batch": f["batch"]["vs_fp64_same_workload"],
                                     "serve": f["serve"]["vs_fp64_same_stream"]},
            "precision_u_error": f["precision"]["error"]["u"]["max"],
            "batch_hbm_frac": f["batch"]["hbm_frac"],
            "precision_hbm_frac": {k: v["achieved_gbs"] / HBM_PEAK_GBS
                                   for k, v in f["precision"]["kernels"].items()},
            "serve_latency_s": f["serve"]["latency_s"]}


# --------------------------------------------------------------------------
# precision form: 60 M Tet4, per-element E and nu
# --------------------------------------------------------------------------
def build_tet4_box(shape, seed=3):
    from akantu2_amd import mesh as M
    mesh = M.box_mesh("_tetrahedron_4", *shape, jitter=0.1, seed=seed)
    rng = np.random.default_rng(seed)
    E = np.exp(rng.uniform(np.log(0.5), np.log(2.0), mesh.nb_elements))
    nu = rng.uniform(0.2, 0.35, mesh.nb_elements)
    return mesh, M.clamped_face_x0(mesh), M.initial_velocity(mesh, 1e-3, seed), E, nu


def _timed_model(mesh, blocked, v0, E, nu, precision, warmup, steps, assembly=True):
    from akantu2_amd import SolidMechanicsModel
    m = SolidMechanicsModel(mesh)
    m.addMaterial("elastic", rho=1.0, E=1.0, nu=0.3, E_el=E, nu_el=nu)
    m.setPrecision(precision)
    m.setBlockAssembly(assembly)
    m.initFull()
    m.getBlockedDOFs()[:] = blocked
    m.getVelocity()[:] = v0
    m.setTimeStep(0.5 * m.getStableTimeStep())
    m.solveSteps(warmup)
    m.synchronize()
    t0 = time.perf_counter()
    m.solveSteps(steps)
    m.synchronize()
    el = time.perf_counter() - t0
    m.enableKernelTiming(True)
    m.solveSteps(10)
    m.synchronize()
    kern = {}
    for k, name in {0: "element_pass", 1: "gather_update"}.items():
        ms, n = m.kernelTiming(k)
        b = m.kernelBytes(k)
        kern[name] = {"avg_ms": ms / n, "bytes": b,
                      "achieved_gbs": b / (ms / n * 1e-3) / 1e9}
    return m, el, kern


def run_precision(args):
    steps = args.steps or 2000
    shape = tuple(int(x) for x in args.shape.split(",")) if args.shape else (1000, 100, 100)
    t = time.perf_counter()
    mesh, blocked, v0, E, nu = build_tet4_box(shape)
    print(f"# mesh {time.perf_counter() - t:.1f} s, {mesh.nb_elements} Tet4",
          file=sys.stderr, flush=True)
    res, models = {}, {}
    # the plain FP64 path sums each node's first block run in the element
    # pass (prefix partials, same bits); the opt-in FP64 block partials sum
    # every run (other rounding)
    for key, prec, assembly in (("fp32", "fp32", True), ("fp64", "fp64", True),
                                ("fp64_block_partials", "fp64", "block_partials")):
        m, el, kern = _timed_model(mesh, blocked, v0, E, nu, prec, args.warmup, steps,
                                   assembly)
        res[key] = {"value": mesh.nb_elements * steps / el, "ms_per_step": 1e3 * el / steps,
                    "kernels": kern, "assembly": m.assembly()}
        print(f"# {key}: {1e3 * el / steps:.3f} ms/step", file=sys.stderr, flush=True)
        models[key] = m
    # all ran warmup + steps + 10 steps: the FP32 form and the FP64 block
    # partials against the plain FP64 path (bit-identical to the oracle at
    # this size) on the full mesh
    g64 = models["fp64"]

    def errors_vs_plain(g):
        return {"after_steps": args.warmup + steps + 10,
                "u": rel_errors(g.getDisplacement(), g64.getDisplacement()),
                "v": rel_errors(g.getVelocity(), g64.getVelocity()),
                "f_int": rel_errors(g.getInternalForce(), g64.getInternalForce())}
    full_error = errors_vs_plain(models["fp32"])
    bp_error = errors_vs_plain(models["fp64_block_partials"])
    del g64, models
    el = res["fp32"]["kernels"]["element_pass"]
    out = {
        "metric": "elem*qp stress+force updates/s", "value": res["fp32"]["value"],
        "unit": "elem*qp/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": res["fp32"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"Tet4 {shape[0]}x{shape[1]}x{shape[2]} cubes x 6 jittered, "
                               "linear elastic with per-element E, nu, FP32 form "
                               "(FP64 compensated nodal update)",
                   "elements": mesh.nb_elements, "nodes": mesh.nb_nodes},
        "roofline": {"bound": "hbm", "achieved": el["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": el["achieved_gbs"] / HBM_PEAK_GBS,
                     "traffic": None, "kernel": "k_tet4_bpa<elastic> (FP32, block-partial assembly)",
                     "algorithmic_bytes_per_launch": el["bytes"],
                     "avg_kernel_ms": el["avg_ms"]},
        "forms": res,
        "fp32_over_fp64": res["fp32"]["value"] / res["fp64"]["value"],
        "fp32_error_vs_plain_fp64_full_mesh": full_error,
        "fp64_block_partials_error_vs_plain_fp64_full_mesh": bp_error,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_precision()
    return out


def cpu_baseline_precision(shape=(60, 10, 10), steps=200):
    """The oracle on a 36,000-Tet4 box of the same workload (timed, 1 thread),
    and the FP32 form's per-field error against it on that box."""
    import oracle
    mesh, blocked, v0, E, nu = build_tet4_box(shape)
    o = oracle.OracleModel(oracle.TET4, mesh.nodes, mesh.conn, "elastic", E_el=E, nu_el=nu)
    o.init_full()
    o.blocked[:] = blocked
    o.v[:] = v0
    dt = 0.5 * o.stable_dt()
    o.set_time_step(dt)
    t0 = time.perf_counter()
    o.solve_steps(steps)
    el = time.perf_counter() - t0
    from akantu2_amd import SolidMechanicsModel
    g = SolidMechanicsModel(mesh)
    g.addMaterial("elastic", rho=1.0, E=1.0, nu=0.3, E_el=E, nu_el=nu)
    g.setPrecision("fp32")
    g.initFull()
    g.getBlockedDOFs()[:] = blocked
    g.getVelocity()[:] = v0
    g.setTimeStep(dt)
    g.solveSteps(steps)
    err = {"u": rel_errors(g.getDisplacement(), o.u), "v": rel_errors(g.getVelocity(), o.v),
           "f_int": rel_errors(g.getInternalForce(), o.f_int)}
    return {"value": mesh.nb_elements * steps / el, "unit": "elem*qp/s", "cores": 1,
            "kind": "port", "seconds": el,
            "sample": f"{steps} steps of a {shape[0]}x{shape[1]}x{shape[2]} x 6 Tet4 box "
                      "of the same workload, oracle/akantu_oracle.c, 1 thread",
            "fp32_error_vs_oracle": err}


# --------------------------------------------------------------------------
# batching form: 512 Tet4 plasticity instances
# --------------------------------------------------------------------------
ETYPES = {"tet4": "_tetrahedron_4", "hex8": "_hexahedron_8"}


def batch_instances(n=512, seed=4, scale=1.0, etype="tet4", law=PL):
    """Instance descriptions: cells, material, seed (scale shrinks the sizes
    for a quick run).  The batching workload is Tet4 with linear isotropic
    hardening; etype / law give the same stream of sizes and aspects for
    another element type or law."""
    from akantu2_amd import mesh as M
    rng = np.random.default_rng(seed)
    nel = np.exp(rng.uniform(np.log(5e3), np.log(4e5), n)) * scale
    aspect = rng.uniform(1.0, 8.0, n)
    E = rng.uniform(0.5, 2.0, n)
    nu = rng.uniform(0.2, 0.35, n)
    sy = rng.uniform(1e-3, 5e-3, n)
    hh = rng.uniform(0.0, 0.1, n)
    out = []
    for k in range(n):
        cells = M.cells_for(ETYPES[etype], max(nel[k], 6.0), aspect[k])
        mat = dict(E=float(E[k]), nu=float(nu[k]))
        if law == PL:
            mat.update(sigma_y=float(sy[k]), h=float(hh[k]))
        out.append(dict(cells=cells, mat=mat, seed=1000 + k, etype=etype, law=law))
    return out


def instance_mesh(spec):
    from akantu2_amd import mesh as M
    mesh = M.box_mesh(ETYPES[spec.get("etype", "tet4")], *spec["cells"])
    return mesh, M.clamped_face_x0(mesh), M.initial_velocity(mesh, 1e-2, spec["seed"])


def _fill_batch(precision, specs):
    from akantu2_amd.batch import Batch
    nn = sum((c[0] + 1) * (c[1] + 1) * (c[2] + 1) + 255 for c in (s["cells"] for s in specs))
    ncell = sum(c[0] * c[1] * c[2] for c in (s["cells"] for s in specs))
    hex8 = specs[0].get("etype") == "hex8"
    law = specs[0].get("law", PL)
    b = Batch(precision, nodes=nn, tets=0 if hex8 else 6 * ncell, hexes=ncell if hex8 else 0,
              plastic=law == PL)
    ids = []
    for s in specs:
        mesh, blocked, v0 = instance_mesh(s)
        ids.append(b.add(mesh, law, blocked=blocked, velocity=v0, dt_factor=0.5, **s["mat"]))
    return b, ids


def _time_batch(b, warmup, steps, tune):
    best = b.autotune(repeats=2) if tune else b.launch()
    b.solveStep(warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.solveStep(steps)
    b.synchronize()
    el = time.perf_counter() - t0
    b.enableKernelTiming(True)
    b.solveStep(10)
    b.synchronize()
    kern = {}
    for k, name in {0: "element_pass", 1: "gather_update"}.items():
        ms, n = b.kernelTiming(k)
        by = b.kernelBytes(k)
        kern[name] = {"avg_ms": ms / n, "bytes": by, "achieved_gbs": by / (ms / n * 1e-3) / 1e9}
    b.enableKernelTiming(False)
    return el, kern, best


def run_batch(args):
    steps = args.steps or 500
    law = args.law or (PL if args.etype == "tet4" else "neohookean")
    specs = batch_instances(args.instances, scale=args.scale, etype=args.etype, law=law)
    t = time.perf_counter()
    res = {}
    for prec in ("fp32", "fp64"):
        b, ids = _fill_batch(prec, specs)
        st = b.stats()
        print(f"# {prec} batch: {st['live']} instances, {st['elements']} {args.etype}, "
              f"{st['nodes']} nodes, setup {time.perf_counter() - t:.1f} s",
              file=sys.stderr, flush=True)
        el, kern, best = _time_batch(b, args.warmup, steps, tune=True)
        res[prec] = {"value": st["quadrature_points"] * steps / el,
                     "ms_per_step": 1e3 * el / steps, "kernels": kern, "launch": best}
        print(f"# {prec}: {1e3 * el / steps:.3f} ms/step {best}", file=sys.stderr, flush=True)
        if prec == "fp32" and not args.no_cpu_baseline:
            # the smallest instance, for the error check against the oracle
            k = int(np.argmin([np.prod(s["cells"]) for s in specs]))
            gpu_u = b.getDisplacement(ids[k])
            gpu_steps = b.steps(ids[k])
        b.close()
        t = time.perf_counter()
    el = res["fp32"]["kernels"]["element_pass"]
    out = {
        "metric": "elem*qp stress+force updates/s", "value": res["fp32"]["value"],
        "unit": "elem*qp/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": res["fp32"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(specs)} independent {ETYPES[args.etype]} boxes "
                               "(5e3-4e5 elements log-uniform x scale, aspect U(1,8)), "
                               + ("linear isotropic hardening with per-instance E, nu, "
                                  "sigma_y, h" if law == PL else
                                  f"{law} with per-instance E, nu")
                               + ", FP32 form, one tile per workgroup, in a hipGraph, autotuned",
                   "instances": len(specs), "scale": args.scale,
                   "elements": int(st["elements"]),
                   "quadrature_points": int(st["quadrature_points"])},
        "roofline": batch_roofline(args.etype, law, el, int(st["quadrature_points"])),
        "forms": res,
        "fp32_over_fp64": res["fp32"]["value"] / res["fp64"]["value"],
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batch(specs[k], gpu_u, gpu_steps)
    return out


def batch_roofline(etype, law, el, nq):
    """The FP32 batch's element pass against its bound.  Tet4 (one Gauss
    point per element: its node data, plastic history and table entries
    outweigh its arithmetic) is priced on HBM; the sum-factorised Hex8 neo-Hookean
    body on FP32 vector issue, with the flops per Gauss point the counters
    saw on the single-model pass that runs the same body."""
    kernel = (f"k_batch_{'tet4_bpa' if etype == 'tet4' else 'hex8_bpa'}"
              f"<{'plastic' if law == PL else law}>")
    hbm = {"achieved": el["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": el["achieved_gbs"] / HBM_PEAK_GBS}
    if etype == "hex8" and law == "neohookean" and os.path.exists(FLOPS_FILE):
        with open(FLOPS_FILE) as fh:
            fq = json.load(fh)["f32_sf"]["flops_per_qp"]
        flops = fq * nq
        tflops = flops / (el["avg_ms"] * 1e-3) / 1e12
        return {"bound": "fp32_valu", "achieved": tflops, "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": tflops / FP32_PEAK_TFLOPS, "traffic": None,
                "kernel": kernel + " (FP32, sum-factorised, block partials)",
                "compute": "FP32 VALU; packed FP32 instructions are not used, so the "
                           f"ceiling is {FP32_VALU_TFLOPS} TFLOP/s "
                           f"({tflops / FP32_VALU_TFLOPS:.0%} of it)",
                "flops_per_qp": fq, "algorithmic_flops_per_launch": flops,
                "algorithmic_bytes_per_launch": el["bytes"],
                "avg_kernel_ms": el["avg_ms"], "hbm": hbm}
    return dict(bound="hbm", **hbm, traffic=None, kernel=kernel,
                algorithmic_bytes_per_launch=el["bytes"], avg_kernel_ms=el["avg_ms"])


def cpu_baseline_batch(spec, gpu_u, steps):
    """The oracle on one instance of the batch (the smallest), timed on one
    thread, and the FP32 batch's displacement error against it."""
    import oracle
    mesh, blocked, v0 = instance_mesh(spec)
    hex8 = spec.get("etype") == "hex8"
    law = spec.get("law", PL)
    o = oracle.OracleModel(oracle.HEX8 if hex8 else oracle.TET4, mesh.nodes, mesh.conn, law,
                           **spec["mat"])
    o.init_full()
    o.blocked[:] = blocked
    o.v[:] = v0
    o.set_time_step(0.5 * o.stable_dt())
    t0 = time.perf_counter()
    o.solve_steps(steps)
    el = time.perf_counter() - t0
    nqp = 8 if hex8 else 1
    return {"value": mesh.nb_elements * nqp * steps / el, "unit": "elem*qp/s", "cores": 1,
            "kind": "port", "seconds": el,
            "sample": f"{steps} steps of the smallest batch instance ({mesh.nb_elements} "
                      f"{ETYPES[spec.get('etype', 'tet4')]}, {law}), oracle/akantu_oracle.c, "
                      "1 thread",
            "fp32_u_error_vs_oracle": rel_errors(gpu_u, o.u)}


# --------------------------------------------------------------------------
# serving form
# --------------------------------------------------------------------------
def run_serve(args):
    from akantu2_amd import serving as S
    reqs = S.synthetic_requests(args.requests, rate=args.rate, seed=5, scale=args.scale)
    t = time.perf_counter()
    for r in reqs:
        r.materialize()
    print(f"# {len(reqs)} requests generated in {time.perf_counter() - t:.1f} s",
          file=sys.stderr, flush=True)
    server = S.Server(precision="fp32", hbm_bytes=int(args.hbm_gib * (1 << 30)))
    rep = server.run(reqs)
    tt = np.array([r.time_to_result for r in rep.done])
    out = {
        "metric": "elem*qp stress+force updates/s", "value": rep.throughput,
        "unit": "elem*qp/s", "n_gpus": 1, "steps": int(sum(r.steps for r in rep.done)),
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"Poisson stream of {len(reqs)} requests at {args.rate}/s, "
                               "Tet4 or Hex8 boxes of 1e3-2e6 elements (log-uniform x "
                               "scale), elastic or neo-Hookean, 200-5000 steps, FP32 form, "
                               "continuous batching with HBM-footprint admission",
                   "requests": len(reqs), "rate_per_s": args.rate, "scale": args.scale},
        "latency_s": {"p50": float(np.percentile(tt, 50)), "p99": float(np.percentile(tt, 99)),
                      "max": float(tt.max())},
        "makespan_s": rep.makespan, "chunks": rep.chunks, "failed": len(rep.failed),
        "max_live": rep.max_live,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_serve(rep)
    return out


def cpu_baseline_serve(rep):
    """The oracle on the smallest served request (timed, 1 thread) and the
    served result's error against it."""
    import oracle
    from akantu2_amd.batch import law_id
    r = min(rep.done, key=lambda q: q.mesh.nb_elements * q.steps)
    law = {0: "elastic", 2: "neohookean"}[law_id(r.law)]
    o = oracle.OracleModel(r.mesh.etype, r.mesh.nodes, r.mesh.conn, law, **r.material)
    o.init_full()
    o.blocked[:] = r.blocked
    o.v[:] = r.velocity
    o.set_time_step(0.5 * o.stable_dt())
    t0 = time.perf_counter()
    o.solve_steps(r.steps)
    el = time.perf_counter() - t0
    return {"value": r.mesh.nb_elements * r.mesh.nqp * r.steps / el, "unit": "elem*qp/s",
            "cores": 1, "kind": "port", "seconds": el,
            "sample": f"the smallest served request ({r.mesh.nb_elements} elements, "
                      f"{r.steps} steps, {law}), oracle/akantu_oracle.c, 1 thread",
            "u_error_vs_oracle": rel_errors(r.result["displacement"], o.u)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--form", choices=["plain", "precision", "batch", "serve"],
                    default="plain")
    ap.add_argument("--precision", choices=["fp64", "fp32", "fp64_fast"], default="fp64",
                    help="plain form only")
    ap.add_argument("--pipeline", type=int, default=0,
                    help="plain form: element parts of the pipelined schedule (0 = off)")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="plain form: skip the FP32 / batch / serving legs")
    ap.add_argument("--shape", type=str, default=None,
                    help="override cells per direction, e.g. 20,10,10")
    ap.add_argument("--instances", type=int, default=512, help="batch form")
    ap.add_argument("--etype", choices=["tet4", "hex8"], default="tet4",
                    help="batch form: element type of the instances")
    ap.add_argument("--law", type=str, default=None,
                    help="batch form: material law (default: plasticity for Tet4, "
                         "neohookean for Hex8)")
    ap.add_argument("--scale", type=float, default=1.0,
                    help="batch / serve forms: multiply the instance sizes")
    ap.add_argument("--requests", type=int, default=64, help="serve form")
    ap.add_argument("--rate", type=float, default=4.0, help="serve form: requests/s")
    ap.add_argument("--hbm-gib", type=float, default=16.0,
                    help="serve form: device memory the server may use")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1 "
                          "(sharding over several GPUs is designed, not implemented)"}))
        sys.exit(2)
    run = {"plain": run_plain, "precision": run_precision, "batch": run_batch,
           "serve": run_serve}[args.form]
    print(json.dumps(run(args)), flush=True)


if __name__ == "__main__":
    main()
