"""Benchmark of the MELODY 2D explicit step on one MI355X.

Headline workload (`configs.c2`): biaxial compression of 1000 polydisperse
deformable grains (~200 MLS nodes each, ~208k nodes, ~1.04M Gauss points,
neo-Hookean, penalty Mohr-Coulomb mu = 0.5, rigid walls, density x100)
deposited into a pile, so ~1.8e4 contact elements act from the start;
generated with fixed seeds on the host and handed to the library as a
binary model container.  A "step" is one Euler_step (internal forces,
contact re-projection and laws, BCs, integration, kinematics) plus the
contact-update cadence (a proximity update every 10 steps), all on the
device, with the state resident in HBM.

Printed: one JSON line with whole-job node-steps/s (FP64 drop-in), the
roofline of the dominant kernel (algorithmic bytes over its HIP-event time
in the timed region) and the CPU baseline (the FP64 OpenMP oracle on a
bounded sample of the same workload, on this host's cores), and beside them:
  fp32_mixed  the same workload in the FP32-mixed form, with its force,
              displacement and energy errors against FP64 from one state;
  plate       the perforated 1M-node RPIM plate (`configs.c3`): internal forces
              of one large body, FP64 and FP32-mixed, roofline, CPU baseline;
  serving     Poisson arrivals of plate jobs (`configs.c5`) through the job
              server, p50/p99 latency and aggregate node-steps/s.
"""
import argparse
import json
import math
import os
import shutil
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FORCE_BAR = 1e-5


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


def measured_traffic(kernel, info, tag, fp32=False):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of workload `tag`, used only when it was measured on this
    workload with the current kernel sources; otherwise None."""
    import hashlib
    path = os.path.join(ROOT, "profiles", "%s%s_%s.json" % (kernel, "32" if fp32 else "", tag))
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
    return rec["traffic_bytes"], rec.get("summary", "profiles/")


def cpu_baseline(model_path, n_nodes, steps, threads):
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


def timed_run(native, path, precision, args, tag, prepare=None, options=None):
    """Load, warm up, then two windows of args.steps steps: the value window
    (steps as a user runs them) and the launch-duration window (HIP events
    around the dominant kernels on the library's stream)."""
    name = "fp64" if precision == native.FP64 else "fp32-mixed"
    t0 = time.time()
    ctx = native.Context(path, binary=True, precision=precision, options=options)
    info = ctx.get_info()
    log("[%s %s] loaded in %.1fs: nodes %d, gauss %d, nnz %d, contacts %d" % (
        tag, name, time.time() - t0, info.n_nodes, info.n_gauss, info.nnz_gauss, info.n_contacts))
    if prepare:
        prepare(ctx)
    ctx.euler_steps(args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.euler_steps(args.steps)
    ctx.synchronize()
    sec = time.perf_counter() - t0
    info = ctx.get_info()
    value = info.n_nodes * args.steps / sec
    log("[%s %s] timed %d steps in %.3fs: %.3e node-steps/s, %.4f ms/step" % (
        tag, name, args.steps, sec, value, 1e3 * sec / args.steps))
    ctx.set_kernel_timing(["step_forces", "step_finish", "block_force_update"])
    ctx.synchronize()
    ctx.euler_steps(args.steps)
    ctx.synchronize()
    kt = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    roof = None
    dominant = next((k for k in ("step_forces", "block_force_update") if kt[k][1]), None)
    if dominant:
        ms, n = kt[dominant]
        avg_s = ms / n * 1e-3
        fp32 = precision != native.FP64
        b = algorithmic_bytes(dominant, info, fp32=fp32)
        achieved = b / avg_s / 1e9
        traffic, src = measured_traffic(dominant, info, tag, fp32=fp32)
        roof = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "avg_launch_us": round(avg_s * 1e6, 2), "algorithmic_bytes": b}
        if src:
            roof["traffic_source"] = src
        log("[%s %s] kernel %s %.2f us/launch, %.1f GB/s (%.3f of peak)" % (
            tag, name, dominant, avg_s * 1e6, achieved, achieved / HBM_PEAK_GBS))
    breakdown = None
    if args.breakdown:
        ctx.set_kernel_timing(True)
        ctx.euler_steps(min(args.steps, 50))
        kt2 = ctx.kernel_timing()
        ctx.set_kernel_timing(False)
        breakdown = {k: round(v[0] / max(v[1], 1) * 1e3, 2) for k, v in kt2.items() if v[1]}
        log("[%s %s] per-launch us: %s" % (tag, name, json.dumps(breakdown)))
    return ctx, info, value, sec, roof, breakdown


def _norm_err(native, a, b, counts, rigid, fields):
    """||f_b - f_a|| / ||f_a|| over the deformable bodies' nodes."""
    num = den = 0.0
    for bi, n in enumerate(counts):
        if rigid[bi]:
            continue
        for f in fields:
            x = a.node_field(bi, f, n)
            d = b.node_field(bi, f, n) - x
            num += float(d @ d)
            den += float(x @ x)
    return (num / den) ** 0.5 if den > 0 else None


def fp32_errors(native, ctx64, ctx32, counts, rigid, steps_after):
    """Errors of the FP32-mixed form against FP64 from one FP64 state (the
    FP64 context's, loaded into the FP32 context): internal and contact
    forces of one evaluation (mel_eval_forces), then displacements (the
    increment of the displacement parameters) and kinetic energy after
    `steps_after` steps of both forms."""
    ctx32.load_state(ctx64.save_state())
    out = {}
    ctx64.eval_forces(native.EVAL_INTERNAL)
    ctx32.eval_forces(native.EVAL_INTERNAL)
    out["rel_internal_force_err"] = _norm_err(native, ctx64, ctx32, counts, rigid,
                                              ("x_internal_force", "y_internal_force"))
    ctx64.eval_forces(native.EVAL_CONTACT)
    ctx32.eval_forces(native.EVAL_CONTACT)
    out["rel_contact_force_err"] = _norm_err(native, ctx64, ctx32, counts, rigid,
                                             ("x_contact_force", "y_contact_force"))
    u0 = {(bi, f): ctx64.node_field(bi, f, n) for bi, n in enumerate(counts) if not rigid[bi]
          for f in ("x_disp_param", "y_disp_param")}
    _, w0, e0 = ctx64.scalars()
    ctx64.euler_steps(steps_after)
    ctx32.euler_steps(steps_after)
    num = den = 0.0
    for (bi, f), u in u0.items():
        d64 = ctx64.node_field(bi, f, counts[bi]) - u
        d32 = ctx32.node_field(bi, f, counts[bi]) - u
        num += float((d32 - d64) @ (d32 - d64))
        den += float(d64 @ d64)
    out["rel_displacement_err"] = (num / den) ** 0.5 if den > 0 else None
    _, w64, e64 = ctx64.scalars()
    _, w32, e32 = ctx32.scalars()
    out["rel_kinetic_energy_err"] = abs(e32[1] - e64[1]) / e64[1] if e64[1] > 0 else None
    dw64 = float(sum(w64 - w0))
    out["rel_work_err"] = abs(float(sum(w32 - w0)) - dw64) / abs(dw64) if dw64 else None
    out["steps_after"] = steps_after
    out["meets_force_bar"] = bool(out["rel_internal_force_err"] is not None and
                                  out["rel_internal_force_err"] <= FORCE_BAR and
                                  (out["rel_contact_force_err"] is None or out["rel_contact_force_err"] <= FORCE_BAR))
    return out


def granular_leg(args, native, tmpdir):
    from melody2d_amd.preprocessor import c2
    from melody2d_amd.preprocessor.binary import write_binary
    t0 = time.time()
    model = c2(n_grains=args.grains)
    log("[granular] generated %d bodies, %d nodes, %d Gauss points in %.1fs" % (
        len(model.bodies), model.n_nodes, model.n_gauss, time.time() - t0))
    path = os.path.join(tmpdir, "c2.bin")
    write_binary(model, path)
    tag = "%dgrains" % args.grains
    ctx64, info, value, sec, roof, bd = timed_run(native, path, native.FP64, args, tag)
    ctx32, info32, value32, sec32, roof32, bd32 = timed_run(native, path, native.FP32_MIXED, args, tag)
    counts = [b.n_nodes for b in model.bodies]
    rigid = [b.rigid for b in model.bodies]
    done = args.warmup + 2 * args.steps + (min(args.steps, 50) if args.breakdown else 0)
    if args.no_fp32_errors:
        errs = {"skipped": "--no-fp32-errors", "meets_force_bar": None, "rel_internal_force_err": None}
    else:
        if args.error_step > done:
            ctx64.euler_steps(args.error_step - done)
        state_info = ctx64.get_info()
        t0 = time.time()
        errs = fp32_errors(native, ctx64, ctx32, counts, rigid, args.error_steps_after)
        errs["state_step"] = max(args.error_step, done)
        errs["state_contacts"] = state_info.n_contacts
        log("[granular] fp32-mixed errors vs fp64 from the step-%d state (%d contacts): %s (%.1fs)" % (
            errs["state_step"], state_info.n_contacts, json.dumps(errs), time.time() - t0))
    ctx64.close()
    ctx32.close()
    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        t0 = time.time()
        v, sc = cpu_baseline(path, info.n_nodes, args.cpu_steps, threads)
        log("[granular] cpu baseline: %.3e node-steps/s over %d steps (%.1fs stepping, %.1fs total)" % (
            v, args.cpu_steps, sc, time.time() - t0))
        cpu = {"value": round(v, 1), "unit": "node-steps/s", "cores": threads, "kind": "port",
               "sample": "%d Euler steps of the same %d-node workload after its initialisation, "
                         "FP64 OpenMP oracle" % (args.cpu_steps, info.n_nodes)}
    os.remove(path)
    return dict(info=info, value=value, sec=sec, roof=roof, bd=bd, value32=value32, sec32=sec32, roof32=roof32,
                bd32=bd32, errs=errs, cpu=cpu)


def make_models(kind, args, out_dir):
    """Model containers built in a process of their own: their RPIM systems
    are solved by torch on the GPU, and torch brings a HIP runtime of its
    own, which must not share a process with the library's."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--make-models", kind, "--out-dir", out_dir,
           "--plate-nodes", str(args.plate_nodes), "--serve-jobs", str(args.serve_jobs),
           "--serve-nodes", str(args.serve_nodes[0]), str(args.serve_nodes[1])]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, timeout=900)
    if r.returncode != 0:
        raise RuntimeError("model generation (%s) exited with %d" % (kind, r.returncode))
    return json.loads(r.stdout.decode().strip().splitlines()[-1])


def make_models_main(args):
    import numpy as np
    from melody2d_amd.preprocessor import configs
    from melody2d_amd.preprocessor.binary import write_binary
    out = []
    t0 = time.time()
    if args.make_models == "plate":
        m = configs.c3(n_nodes=args.plate_nodes, device="cuda")
        p = os.path.join(args.out_dir, "c3.bin")
        write_binary(m, p)
        b = m.bodies[0]
        np.save(os.path.join(args.out_dir, "c3_X.npy"), b.X)
        out.append({"path": p, "nodes": m.n_nodes, "gauss": m.n_gauss, "holes": len(m.holes), "dt": m.dt,
                    "nu": m.materials[0][2][4], "X": os.path.join(args.out_dir, "c3_X.npy")})
    else:
        rng = np.random.default_rng(4242)
        lo, hi = args.serve_nodes
        for k in range(args.serve_jobs):
            n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
            m = configs.c5_job(n, seed=3000 + k, cohesive=bool(k % 2), device="cuda")
            p = os.path.join(args.out_dir, "job%03d.bin" % k)
            write_binary(m, p)
            out.append({"path": p, "nodes": m.n_nodes, "bodies": len(m.bodies), "holes": len(m.holes)})
    log("generated %d %s model(s) in %.1fs" % (len(out), args.make_models, time.time() - t0))
    print(json.dumps(out), flush=True)


def plate_leg(args, native, tmpdir):
    """Configuration C3: the perforated RPIM plate, internal forces of one large
    body (the whole step is timed: the internal-force launch and a small
    update), from the uniaxial 10/s field."""
    import numpy as np
    t0 = time.time()
    rec = make_models("plate", args, tmpdir)[0]
    log("[plate] generated %d nodes, %d Gauss points, %d holes in %.1fs" % (
        rec["nodes"], rec["gauss"], rec["holes"], time.time() - t0))
    path = rec["path"]
    X = np.load(rec["X"])
    os.remove(rec["X"])
    rate = 10.0
    eps = rate * 1000 * rec["dt"]
    nu = rec["nu"]
    x, y = X[:, 0], X[:, 1]

    def uniaxial(ctx):
        for f, v in (("x_disp_param", eps * x), ("y_disp_param", -nu * eps * y),
                     ("x_vel_param", rate * x), ("y_vel_param", -nu * rate * y)):
            ctx.set_node_field(0, f, np.ascontiguousarray(v))

    tag = "plate%dk" % (args.plate_nodes // 1000)
    a = argparse.Namespace(**vars(args))
    a.steps, a.warmup = args.plate_steps, 5
    ctx64, info, value, sec, roof, bd = timed_run(native, path, native.FP64, a, tag, uniaxial)
    ctx32, _, value32, sec32, roof32, bd32 = timed_run(native, path, native.FP32_MIXED, a, tag, uniaxial)
    if args.no_fp32_errors:
        errs = {"skipped": "--no-fp32-errors", "meets_force_bar": None}
    else:
        uniaxial(ctx64)
        errs = fp32_errors(native, ctx64, ctx32, [rec["nodes"]], [False], 200)
        errs.pop("rel_contact_force_err")
        errs["meets_force_bar"] = bool(errs["rel_internal_force_err"] <= FORCE_BAR)
        errs["state"] = "uniaxial 10/s field after 1000 steps' strain"
        log("[plate] fp32-mixed errors vs fp64: %s" % json.dumps(errs))
    ctx64.close()
    ctx32.close()
    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        t0 = time.time()
        v, sc = cpu_baseline(path, info.n_nodes, args.plate_cpu_steps, threads)
        log("[plate] cpu baseline: %.3e node-steps/s over %d steps (%.1fs stepping, %.1fs total)" % (
            v, args.plate_cpu_steps, sc, time.time() - t0))
        cpu = {"value": round(v, 1), "unit": "node-steps/s", "cores": threads, "kind": "port",
               "sample": "%d Euler steps of the same plate after its initialisation, FP64 OpenMP oracle "
                         "(one body: only its internal forces spread over the threads)" % args.plate_cpu_steps}
    os.remove(path)
    out = {"workload": "perforated plate, %d RPIM nodes (30-node supports), %d holes with their own borders, "
                       "linear elastic E = 70 GPa, uniaxial 10/s field, no contact" % (info.n_nodes, rec["holes"]),
           "nodes": info.n_nodes, "gauss_points": info.n_gauss, "support_entries": info.nnz_gauss,
           "value": round(value, 1), "unit": "node-steps/s", "ms_per_step": round(1e3 * sec / a.steps, 4),
           "dtype": "f64", "roofline": roof, "cpu_baseline": cpu,
           "fp32_mixed": {"value": round(value32, 1), "ms_per_step": round(1e3 * sec32 / a.steps, 4),
                          "speedup_vs_fp64": round(value32 / value, 3), "errors_vs_fp64": errs,
                          "roofline": roof32}}
    if args.breakdown:
        out["kernel_us"] = bd
        out["fp32_mixed"]["kernel_us"] = bd32
    return out


def serve_leg(args, native, tmpdir):
    """Continuous batching through the job server (`configs.c5`): Poisson
    arrivals of perforated-plate jobs of log-uniform size, every second one
    two plates bonded by a cohesive law, admitted under an HBM budget and
    run concurrently; the same arrival schedule in FP64 and in the
    FP32-mixed form.  Latency is completion minus arrival (queueing,
    loading and stepping)."""
    import numpy as np
    lo, hi = args.serve_nodes
    t0 = time.time()
    jobs = [(r["path"], r["nodes"]) for r in make_models("serve", args, tmpdir)]
    rng = np.random.default_rng(4243)
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
            left = srv.run(timeout=900.0)
            infos = [srv.job(j) for j in ids]
            failed = [(j, srv.job_error(j)) for j, inf in zip(ids, infos) if inf.state != native.JOB_DONE]
            if left or failed:
                raise RuntimeError("serve (%s): %d unfinished, failed %s" % (name, left, failed[:3]))
        lat = np.array([inf.finished - inf.arrival for inf in infos])
        wait = np.array([inf.admitted - inf.arrival for inf in infos])
        span = max(inf.finished for inf in infos) - min(inf.arrival for inf in infos)
        work = sum(inf.n_nodes * inf.steps_done for inf in infos)
        out[name] = {"p50_latency_s": round(float(np.percentile(lat, 50)), 4),
                     "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
                     "p50_wait_s": round(float(np.percentile(wait, 50)), 4),
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
                "workload": "Poisson arrivals of perforated RPIM plates in uniaxial tension at 10/s, nodes "
                            "log-uniform in [%d, %d], every second job two plates bonded by CZMlinear "
                            "(peak 1 MPa, softening over 10 um)" % (lo, hi)})
    return out


def batch_leg(args, native, tmpdir):
    """Configuration C4: independent periodic shear cells of uneven sizes in one
    context (every step of every sample in the same launches), FP32-mixed,
    launch configuration chosen by the autotuner on the batch's first
    samples; the default configuration timed beside it."""
    from melody2d_amd.preprocessor import configs
    from melody2d_amd.preprocessor.binary import write_binary
    t0 = time.time()
    sub = configs.c4_batch(n_samples=args.batch_tune_samples)
    sub_path = os.path.join(tmpdir, "c4_tune.bin")
    write_binary(sub, sub_path)
    best, tune_ms, tried = native.autotune(sub_path, precision=native.FP32_MIXED, steps=100)
    os.remove(sub_path)
    log("[batch] autotuned on %d samples (%d nodes): %s, %.4f ms/step, %d configurations (%.1fs)" % (
        args.batch_tune_samples, sub.n_nodes, json.dumps(best), tune_ms, tried, time.time() - t0))
    t0 = time.time()
    m = configs.c4_batch(n_samples=args.batch_samples)
    path = os.path.join(tmpdir, "c4.bin")
    write_binary(m, path)
    log("[batch] generated %d samples, %d bodies, %d nodes, %d Gauss points in %.1fs" % (
        args.batch_samples, len(m.bodies), m.n_nodes, m.n_gauss, time.time() - t0))
    tag = "c4_%d" % args.batch_samples
    ctx, info, value, sec, roof, bd = timed_run(native, path, native.FP32_MIXED, args, tag, options=best)
    used = ctx.options()
    ctx.close()
    ctx, _, value_def, sec_def, _, _ = timed_run(native, path, native.FP32_MIXED, args, tag + "_default")
    ctx.close()
    os.remove(path)
    grains = [g for g, _, _ in m.samples]
    out = {"workload": "%d independent periodic shear cells in one context, U{10..200} grains of U{100..300} "
                       "MLS nodes, Coulomb mu U[0.2, 0.6] per sample, top wall sheared at 0.05 m/s" % args.batch_samples,
           "samples": args.batch_samples, "grains": int(sum(grains)), "nodes": info.n_nodes,
           "gauss_points": info.n_gauss, "support_entries": info.nnz_gauss, "contacts": info.n_contacts,
           "value": round(value, 1), "unit": "node-steps/s", "ms_per_step": round(1e3 * sec / args.steps, 4),
           "dtype": "f32 gradients/stresses/contributions, f64 state, contacts and sums",
           "tuned_options": used, "tuned_on_samples": args.batch_tune_samples,
           "default_options_value": round(value_def, 1),
           "tuned_speedup_vs_default": round(value / value_def, 3), "roofline": roof}
    if args.breakdown:
        out["kernel_us"] = bd
    return out


def run_serve_child(args):
    """The serving leg in a child process of its own (its contexts, streams
    and graphs come and go by the dozen), so that whatever happens there
    cannot take the headline figures with it."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--serve-only", "--serve-jobs", str(args.serve_jobs),
           "--serve-steps", str(args.serve_steps), "--serve-nodes", str(args.serve_nodes[0]),
           str(args.serve_nodes[1]), "--serve-rate", str(args.serve_rate), "--serve-slice", str(args.serve_slice),
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
    ap.add_argument("--grains", type=int, default=1000)
    ap.add_argument("--plate-nodes", type=int, default=1000000)
    ap.add_argument("--plate-steps", type=int, default=50)
    ap.add_argument("--plate-cpu-steps", type=int, default=5)
    ap.add_argument("--no-plate", action="store_true", help="skip the plate leg")
    ap.add_argument("--plate-only", action="store_true",
                    help="run only the plate leg and print its JSON object (for profiling its kernels)")
    ap.add_argument("--cpu-steps", type=int, default=1000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--breakdown", action="store_true", help="also time every kernel in a separate window")
    ap.add_argument("--error-step", type=int, default=3000,
                    help="FP64 step whose state the FP32-mixed errors are evaluated from")
    ap.add_argument("--no-fp32-errors", action="store_true",
                    help="skip the FP32-mixed error evaluation (profiling passes time only the kernels)")
    ap.add_argument("--error-steps-after", type=int, default=500,
                    help="steps both forms run from that state for the displacement and energy errors")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies and 5 reductions of MIB MiB (scale references for HBM byte counters)")
    ap.add_argument("--batch-samples", type=int, default=64,
                    help="shear cells of the batched leg (0: skip it; the configuration's full batch is 512)")
    ap.add_argument("--batch-tune-samples", type=int, default=16,
                    help="samples the autotuner times candidate configurations on")
    ap.add_argument("--serve-jobs", type=int, default=12, help="jobs of the serving leg (0: skip it)")
    ap.add_argument("--serve-steps", type=int, default=20000)
    ap.add_argument("--serve-nodes", type=int, nargs=2, default=[5000, 500000])
    ap.add_argument("--serve-rate", type=float, default=2.0, help="job arrivals per second")
    ap.add_argument("--serve-slice", type=int, default=200, help="steps per slice of a running job")
    ap.add_argument("--serve-budget-gb", type=float, default=250.0)
    ap.add_argument("--serve-graphs", type=int, default=0,
                    help="1: the jobs replay captured graphs (their launches then overlap less)")
    ap.add_argument("--serve-hw-queues", type=int, default=16,
                    help="hardware queues of the serving child process (GPU_MAX_HW_QUEUES, at most 32)")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving leg and print its JSON object (the child process of a full run)")
    ap.add_argument("--make-models", choices=["plate", "serve"],
                    help="only build the plate or the serving jobs into --out-dir (a child process of a full run)")
    ap.add_argument("--out-dir")
    args = ap.parse_args()
    if args.make_models:
        make_models_main(args)
        return
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
        # a read-only pass (a reduction reads every byte once, writes one
        # value): the read counter's scale without a write stream beside it
        s = torch.zeros((), dtype=torch.float64, device="cuda")
        for _ in range(5):
            s += a.sum()
        torch.cuda.synchronize()
        log("copied %d MiB 5 times, summed it 5 times (%.0f)" % (args.copy_calibration, float(s)))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (samples and grains would be "
              "sharded across GPUs by HBM footprint).", file=sys.stderr)
        sys.exit(2)

    from melody2d_amd import native
    tmpdir = tempfile.mkdtemp(prefix="mel_bench_")
    if args.plate_only:
        print(json.dumps(plate_leg(args, native, tmpdir)), flush=True)
        shutil.rmtree(tmpdir, ignore_errors=True)
        return
    g = granular_leg(args, native, tmpdir)
    plate = None
    if not args.no_plate:
        try:
            plate = plate_leg(args, native, tmpdir)
        except Exception as e:   # the headline line is printed whatever happens here
            log("plate leg failed: %r" % (e,))
            plate = {"error": repr(e)}
    batch = None
    if args.batch_samples > 0:
        try:
            batch = batch_leg(args, native, tmpdir)
        except Exception as e:
            log("batch leg failed: %r" % (e,))
            batch = {"error": repr(e)}
    shutil.rmtree(tmpdir, ignore_errors=True)
    serving = run_serve_child(args) if args.serve_jobs > 0 else None

    info = g["info"]
    errs = g["errs"]
    out = {
        "metric": "node-steps/s, FP32-mixed vs FP64 at ≤1e-5 rel force err; p50/p99 job latency",
        "value": round(g["value"], 1),
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * g["sec"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "biaxial compression, %d deformable grains deposited into a pile + 4 rigid walls "
                               "(FP64 drop-in)" % args.grains,
                   "nodes": info.n_nodes, "gauss_points": info.n_gauss, "support_entries": info.nnz_gauss,
                   "contacts": info.n_contacts, "precision": "fp64", "parallelism": "single GPU"},
        "roofline": g["roof"],
        "cpu_baseline": g["cpu"],
        "fp32_mixed": {"value": round(g["value32"], 1), "unit": "node-steps/s",
                       "ms_per_step": round(1e3 * g["sec32"] / args.steps, 4),
                       "speedup_vs_fp64": round(g["value32"] / g["value"], 3),
                       "meets_force_bar": errs["meets_force_bar"],
                       "rel_internal_force_err_vs_fp64": errs["rel_internal_force_err"],
                       "errors_vs_fp64": errs,
                       "dtype": "f32 gradients/stresses/contributions, f64 state, contacts and sums",
                       "roofline": g["roof32"]},
    }
    if plate:
        out["plate"] = plate
    if batch:
        out["batch"] = batch
    if serving:
        out["serving"] = serving
    if g["bd"]:
        out["kernel_us"] = g["bd"]
        out["fp32_mixed"]["kernel_us"] = g["bd32"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
