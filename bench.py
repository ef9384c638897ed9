"""Benchmark of the MELODY 2D explicit step on one MI355X.

Workload: biaxial compression of 1000 polydisperse deformable grains
(~200 MLS nodes each, ~209k nodes, ~1.04M Gauss points, neo-Hookean,
penalty Mohr-Coulomb mu = 0.5, rigid walls, density x100), generated with
fixed seeds on the host and handed to the library as a binary model
container.  A "step" is one Euler_step (internal forces, contact
re-projection and laws, BCs, integration, kinematics) plus the contact-update
cadence (a proximity update every 10 steps), all on the device.

Printed: one JSON line with whole-job node-steps/s, the roofline of the
dominant kernel (algorithmic bytes over its HIP-event time in the timed
region), and the CPU baseline (the FP64 OpenMP oracle on a bounded sample
of the same workload, on this host's cores).
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(kernel, info, fp32=False):
    """Bytes a launch must move at minimum, each array counted once.
    N = nodes, G = Gauss points, nnz = Gauss-point support entries,
    NB = border nodes, nc = contact elements."""
    N, G, nnz = info.n_nodes, info.n_gauss, info.nnz_gauss
    NB, nc = info.n_border_nodes, info.n_contacts
    row = info.nnz_d2d / max(N, 1)
    # internal forces: per support entry dphi/dx, dphi/dy (16 B in FP64, 8 B
    # in the FP32-mixed form) + node index and destination slot (4); per
    # Gauss point c0 (8) and region (4); per node the displacement
    # parameters read (16), the internal force written (16), node ids (8)
    forces = (12 if fp32 else 20) * nnz + 12 * G + 40 * N
    # finish: per node internal force read (16), masses and inverse masses
    # (32), damping force (16), velocity and displacement parameters read (32),
    # spreading offset (4), new displacement / velocity / acceleration
    # parameters written (48); per border node its bucket (count, four
    # records: 84), six point forces read (48) and two written (16), its
    # kinematics (X, Y read, 8 doubles written: 80) and DOF_TO_DISP row
    # (12 B per entry)
    finish = 148 * N + NB * (84 + 48 + 16 + 80 + 12 * row)
    contacts = 256 * nc     # element state read and written, point-force records
    if kernel == "block_force_update":
        return int(forces)
    if kernel == "step_forces":
        return int(forces + contacts)
    if kernel == "step_finish":
        return int(finish)
    return None


def measured_traffic(kernel, info, fp32=False):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary, used only when it was measured on this workload with the
    current kernel sources; otherwise None."""
    import hashlib
    path = os.path.join(ROOT, "profiles", "%s%s_1000grains.json" % (kernel, "32" if fp32 else ""))
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        rec = json.load(f)
    h = hashlib.sha256()
    for src in rec.get("sources", []):
        with open(os.path.join(ROOT, "melody2d_amd", "csrc", src), "rb") as f:
            h.update(f.read())
    if rec.get("src_sha256_16") != h.hexdigest()[:16] or rec.get("nodes") != info.n_nodes or \
            rec.get("support_entries") != info.nnz_gauss:
        return None, None
    return rec["traffic_bytes"], "profiles/step_kernels_1000grains.md"


def cpu_baseline(model_path, n_nodes, steps, threads):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    lib.oracle_run.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                               ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(ctypes.c_double)]
    lib.oracle_last_error.restype = ctypes.c_char_p
    sec = ctypes.c_double()
    r = lib.oracle_run(os.fsencode(model_path), -1, steps, None, None, 0, threads, ctypes.byref(sec))
    if r != 0:
        raise RuntimeError(lib.oracle_last_error().decode())
    return n_nodes * steps / sec.value, sec.value


def serve_leg(args, native, tmpdir):
    """Continuous batching through the job server: Poisson arrivals of
    independent granular jobs of log-uniform size, half with a cohesive
    contact law, admitted under an HBM budget and run concurrently; the same
    arrival schedule in FP64 and in the FP32-mixed form.  Latency is
    completion minus arrival (queueing, loading and stepping)."""
    import math
    import numpy as np
    from melody2d_amd.preprocessor import configs
    from melody2d_amd.preprocessor.binary import write_binary
    rng = np.random.default_rng(4242)
    lo, hi = args.serve_grains
    t0 = time.time()
    jobs = []
    for k in range(args.serve_jobs):
        g = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        mu = float(rng.uniform(0.2, 0.6))
        m = configs.c2(n_grains=g, seed=3000 + k, mu=mu)
        if k % 2:
            # cohesive: peak traction 1 MPa, linear softening over 10 um
            m = configs._set_law(m, "CZMlinear", [1e13, 1e6, 1e6 / 1e13 + 1e-5, 0.0, 1.0, 2e-6])
        p = os.path.join(tmpdir, "job%03d.bin" % k)
        write_binary(m, p)
        jobs.append((p, m.n_nodes))
    gaps = rng.exponential(1.0 / args.serve_rate, args.serve_jobs)
    arrivals = np.cumsum(gaps) - gaps[0]
    log("serve: generated %d jobs (%d..%d nodes) in %.1fs" % (
        len(jobs), min(n for _, n in jobs), max(n for _, n in jobs), time.time() - t0))
    out = {}
    for prec, name in ((native.FP64, "fp64"), (native.FP32_MIXED, "fp32_mixed")):
        with native.Server(hbm_budget=int(args.serve_budget_gb * 1e9), steps_per_slice=args.serve_slice,
                           graph_replay=bool(args.serve_graphs)) as srv:
            ids = [srv.submit(p, args.serve_steps, precision=prec, arrival=float(a))
                   for (p, _), a in zip(jobs, arrivals)]
            left = srv.run(timeout=600.0)
            infos = [srv.job(j) for j in ids]
            failed = [(j, srv.job_error(j)) for j, inf in zip(ids, infos) if inf.state != native.JOB_DONE]
            if left or failed:
                raise RuntimeError("serve (%s): %d unfinished, failed %s" % (name, left, failed[:3]))
        lat = np.array([inf.finished - inf.arrival for inf in infos])
        load = np.array([inf.admitted - inf.arrival for inf in infos])
        span = max(inf.finished for inf in infos) - min(inf.arrival for inf in infos)
        work = sum(inf.n_nodes * inf.steps_done for inf in infos)
        out[name] = {"p50_latency_s": round(float(np.percentile(lat, 50)), 4),
                     "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
                     "p50_wait_s": round(float(np.percentile(load, 50)), 4),
                     "node_steps_per_s": round(work / span, 1), "makespan_s": round(span, 3)}
        log("serve %s: p50 %.3fs p99 %.3fs, %.3e node-steps/s over %.2fs" % (
            name, out[name]["p50_latency_s"], out[name]["p99_latency_s"], work / span, span))
    for p, _ in jobs:
        os.remove(p)
    out.update({"jobs": len(jobs), "steps_per_job": args.serve_steps,
                "nodes_min": min(n for _, n in jobs), "nodes_max": max(n for _, n in jobs),
                "nodes_total": sum(n for _, n in jobs), "arrival_rate_per_s": args.serve_rate,
                "hbm_budget_gb": args.serve_budget_gb, "steps_per_slice": args.serve_slice,
                "graph_replay": bool(args.serve_graphs),
                "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                "workload": "Poisson arrivals of biaxial granular jobs, grains log-uniform in [%d, %d] "
                            "(~210 nodes each), Coulomb mu ~ U[0.2, 0.6], every second job CZMlinear"
                            % (lo, hi)})
    return out


def run_serve_child(args):
    """The serving leg in a child process of its own (its contexts, streams
    and graphs come and go by the dozen), so that whatever happens there
    cannot take the headline figures with it."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--serve-only", "--serve-jobs", str(args.serve_jobs),
           "--serve-steps", str(args.serve_steps), "--serve-grains", str(args.serve_grains[0]),
           str(args.serve_grains[1]), "--serve-rate", str(args.serve_rate), "--serve-slice", str(args.serve_slice),
           "--serve-budget-gb", str(args.serve_budget_gb), "--serve-graphs", str(args.serve_graphs)]
    # more hardware queues let more of the jobs' streams run at once
    env = dict(os.environ, GPU_MAX_HW_QUEUES=str(max(1, min(32, args.serve_hw_queues))))
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, timeout=900, env=env)
    except subprocess.TimeoutExpired:
        log("serve: child timed out")
        return {"error": "timed out"}
    lines = r.stdout.decode().strip().splitlines()
    if r.returncode != 0 or not lines:
        log("serve: child exited with %d" % r.returncode)
        return {"error": "exit status %d" % r.returncode}
    return json.loads(lines[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=["granular", "plate"], default="granular",
                    help="granular: biaxial compression of --grains grains (the headline); plate: the "
                         "perforated RPIM plate of --plate-nodes nodes, internal forces only")
    ap.add_argument("--grains", type=int, default=1000)
    ap.add_argument("--plate-nodes", type=int, default=1000000)
    ap.add_argument("--cpu-steps", type=int, default=2000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--breakdown", action="store_true", help="also time every kernel in a separate window")
    ap.add_argument("--error-step", type=int, default=3000,
                    help="FP64 step at which the FP32-mixed force error is evaluated")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies of MIB MiB (scale reference for HBM byte counters)")
    ap.add_argument("--serve-jobs", type=int, default=16, help="jobs of the serving leg (0: skip it)")
    ap.add_argument("--serve-steps", type=int, default=5000)
    ap.add_argument("--serve-grains", type=int, nargs=2, default=[20, 200])
    ap.add_argument("--serve-rate", type=float, default=8.0, help="job arrivals per second")
    ap.add_argument("--serve-slice", type=int, default=100, help="steps per slice of a running job")
    ap.add_argument("--serve-budget-gb", type=float, default=250.0)
    ap.add_argument("--serve-graphs", type=int, default=0,
                    help="1: the jobs replay captured graphs (their launches then overlap less)")
    ap.add_argument("--serve-hw-queues", type=int, default=16,
                    help="hardware queues of the serving child process (GPU_MAX_HW_QUEUES, at most 32)")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving leg and print its JSON object (the child process of a full run)")
    args = ap.parse_args()
    if args.serve_only:
        from melody2d_amd import native
        d = tempfile.mkdtemp(prefix="mel_serve_")
        print(json.dumps(serve_leg(args, native, d)), flush=True)
        os.rmdir(d)
        return
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied %d MiB 5 times" % args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (samples and grains would be "
              "sharded across GPUs by HBM footprint; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    from melody2d_amd import native
    from melody2d_amd.preprocessor import c2, configs
    from melody2d_amd.preprocessor.binary import write_binary

    t0 = time.time()
    plate = args.workload == "plate"
    if plate:
        model = configs.c3(n_nodes=args.plate_nodes)
        args.serve_jobs = 0
    else:
        model = c2(n_grains=args.grains)
    log("generated %d bodies, %d nodes, %d Gauss points in %.1fs" % (len(model.bodies), model.n_nodes,
                                                                     model.n_gauss, time.time() - t0))
    tmpdir = tempfile.mkdtemp(prefix="mel_bench_")
    path = os.path.join(tmpdir, "c2.bin")
    t0 = time.time()
    write_binary(model, path)
    log("wrote %.0f MB in %.1fs" % (os.path.getsize(path) / 1e6, time.time() - t0))

    def uniaxial(ctx):
        # the plate's displacement and velocity parameters: the uniaxial
        # field of a 10/s strain rate after 1000 steps (RPIM interpolates,
        # so parameters are displacements)
        b = model.bodies[0]
        rate = 10.0
        eps = rate * 1000 * model.dt
        x, y = b.X[:, 0], b.X[:, 1]
        nu = model.materials[0][2][4]
        for f, v in (("x_disp_param", eps * x), ("y_disp_param", -nu * eps * y),
                     ("x_vel_param", rate * x), ("y_vel_param", -nu * rate * y)):
            ctx.set_node_field(0, f, v)

    def run(precision):
        name = "fp64" if precision == native.FP64 else "fp32-mixed"
        t0 = time.time()
        ctx = native.Context(path, binary=True, precision=precision)
        info = ctx.get_info()
        log("[%s] loaded in %.1fs: nodes %d, gauss %d, nnz %d, contacts %d" % (
            name, time.time() - t0, info.n_nodes, info.n_gauss, info.nnz_gauss, info.n_contacts))
        if plate:
            uniaxial(ctx)
        ctx.euler_steps(args.warmup)
        ctx.synchronize()
        # value window: the steps as a user runs them (runs between contact
        # updates replayed from captured graphs)
        t0 = time.perf_counter()
        ctx.euler_steps(args.steps)
        ctx.synchronize()
        t1 = time.perf_counter()
        sec = t1 - t0
        info = ctx.get_info()
        value = info.n_nodes * args.steps / sec
        log("[%s] timed %d steps in %.3fs: %.3e node-steps/s, %.3f ms/step" % (
            name, args.steps, sec, value, 1e3 * sec / args.steps))
        # launch-duration window: the same number of steps with HIP events
        # around every launch of the step's dominant kernel (launch A of the
        # two-launch step, or the internal-force kernel of the launch
        # sequence), recorded on the library's stream; per-launch events
        # switch graph replay off, so this window runs launch by launch
        ctx.set_kernel_timing(["step_forces", "step_finish", "block_force_update"])
        ctx.synchronize()
        t2 = time.perf_counter()
        ctx.euler_steps(args.steps)
        ctx.synchronize()
        t3 = time.perf_counter()
        kt = ctx.kernel_timing()
        ctx.set_kernel_timing(False)
        log("[%s] event window: %d steps in %.3fs (%.3f ms/step launch by launch)" % (
            name, args.steps, t3 - t2, 1e3 * (t3 - t2) / args.steps))
        roof = None
        dominant = next(k for k in ("step_forces", "block_force_update") if kt[k][1])
        ms, n = kt[dominant]
        if n:
            avg_s = ms / n * 1e-3
            b = algorithmic_bytes(dominant, info, fp32=precision != native.FP64)
            achieved = b / avg_s / 1e9
            traffic, src = measured_traffic(dominant, info, fp32=precision != native.FP64)
            roof = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "avg_launch_us": round(avg_s * 1e6, 2), "algorithmic_bytes": b}
            if src:
                roof["traffic_source"] = src
            log("[%s] kernel %s %.2f us/launch, %.1f GB/s" % (name, dominant, avg_s * 1e6, achieved))
        breakdown = None
        if args.breakdown:
            ctx.set_kernel_timing(True)
            ctx.euler_steps(min(args.steps, 50))
            kt2 = ctx.kernel_timing()
            ctx.set_kernel_timing(False)
            breakdown = {k: round(v[0] / max(v[1], 1) * 1e3, 2) for k, v in kt2.items() if v[1]}
            log("[%s] per-launch us: %s" % (name, json.dumps(breakdown)))
        return ctx, info, value, sec, roof, breakdown

    ctx64, info, value, sec, roof, breakdown = run(native.FP64)
    ctx32, _, value32, sec32, roof32, breakdown32 = run(native.FP32_MIXED)

    # relative internal-force error of the FP32-mixed form, from the FP64
    # state: the FP64 run continues to step --error-step (until then the
    # walls have not reached the grains, whose internal forces are pure
    # rounding residue of the reference formula), then both forms take one
    # step from the FP64 displacement parameters
    t0 = time.time()
    done = args.warmup + 2 * args.steps + (min(args.steps, 50) if args.breakdown else 0)
    if args.error_step > done:
        ctx64.euler_steps(args.error_step - done)
    counts = [b.n_nodes for b in model.bodies]
    rigid = [b.rigid for b in model.bodies]
    fields = ["x_disp_param", "y_disp_param", "x_vel_param", "y_vel_param"]
    for bi, n in enumerate(counts):
        if rigid[bi]:
            continue
        for f in fields:
            ctx32.set_node_field(bi, f, ctx64.node_field(bi, f, n))
    ctx64.euler_steps(1)
    ctx32.euler_steps(1)
    num = den = 0.0
    for bi, n in enumerate(counts):
        if rigid[bi]:
            continue
        for f in ("x_internal_force", "y_internal_force"):
            a = ctx64.node_field(bi, f, n)
            d = ctx32.node_field(bi, f, n) - a
            num += float(d @ d)
            den += float(a @ a)
    rel_err = (num / den) ** 0.5 if den > 0 else None
    log("fp32-mixed internal-force error vs fp64 from the same state at step %d: %.3e (%.1fs)" % (
        max(args.error_step, done), rel_err, time.time() - t0))
    ctx64.close()
    ctx32.close()

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        if plate:
            # one body: only its internal forces spread over the threads
            args.cpu_steps = min(args.cpu_steps, 10)
        t0 = time.time()
        v, sc = cpu_baseline(path, info.n_nodes, args.cpu_steps, threads)
        log("cpu baseline: %.3e node-steps/s over %d steps (%.1fs stepping, %.1fs total)" % (
            v, args.cpu_steps, sc, time.time() - t0))
        cpu = {"value": round(v, 1), "unit": "node-steps/s", "cores": threads, "kind": "port",
               "sample": "%d Euler steps of the same %d-node workload after its initialisation, "
                         "FP64 OpenMP oracle" % (args.cpu_steps, info.n_nodes)}
    os.remove(path)
    os.rmdir(tmpdir)
    serving = run_serve_child(args) if args.serve_jobs > 0 else None

    out = {
        "metric": "node-steps/s, FP32-mixed vs FP64 at \u22641e-5 rel force err; p50/p99 job latency",
        "value": round(value, 1),
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * sec / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": ("perforated plate, %d RPIM nodes (30-node supports), %d holes, linear elastic, "
                                "uniaxial 10/s field, no contact (FP64 drop-in)" % (info.n_nodes, len(model.holes))
                                if plate else
                                "biaxial compression, %d deformable grains + 4 rigid walls (FP64 drop-in)" % args.grains),
                   "nodes": info.n_nodes, "gauss_points": info.n_gauss, "support_entries": info.nnz_gauss,
                   "contacts": info.n_contacts, "precision": "fp64", "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp32_mixed": {"value": round(value32, 1), "unit": "node-steps/s",
                       "ms_per_step": round(1e3 * sec32 / args.steps, 4),
                       "speedup_vs_fp64": round(value32 / value, 3),
                       "rel_internal_force_err_vs_fp64": rel_err,
                       "error_state_step": max(args.error_step, done),
                       "dtype": "f32 gradients/stresses/contributions, f64 state and sums",
                       "roofline": roof32},
    }
    if serving:
        out["serving"] = serving
    if breakdown:
        out["kernel_us"] = breakdown
        out["fp32_mixed"]["kernel_us"] = breakdown32
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
