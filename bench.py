"""Benchmark: Helmholtz CG throughput on one MI355X.

Workloads (--workload):

  ensemble_512 (default) -- 512 independent periodic Helmholtz solves
      (lambda M + K) x = f on [0, 2pi]^2, nEls ~ U{8,16,32,64}^2,
      GLL order p ~ U{4,6,8}, lambda = 10^U[0,3] (26.8 M unknowns in all),
      right-hand sides f = M g with g a seeded random field with an isotropic
      k^-3 spectrum.  One step = all 512 solves to a true relative residual of
      1e-10, batched: FP64 CG preconditioned by fast diagonalisation with its
      transforms on the matrix cores (--precision fp64_bf16c, the default:
      FP64 transforms for the first application, BF16 operands for the
      corrections; fp64_fp32c: FP32 corrections; fp64, fp32_ir, bf16_ir:
      FP64, FP32 or BF16 operands throughout;
      --pc jacobi: the reference's Jacobi-PCG with FP32 or BF16 inner
      iterations and FP64 reliable updates).  The plain FP64 GPU path
      (each member alone, FP64 Jacobi-PCG, the reference's algorithm) is timed
      on the same members once and reported beside it, with the other
      precisions / preconditioners batched, and a served two-layer QG stream of
      16 requests for the metric's p50/p99 latency.
  helmholtz_128 -- one solve on 128x128 elements at p = 8 (1 048 576
      unknowns), lambda = 100, same kind of right-hand side.
  qg2_stream -- served two-layer QG step requests (Poisson arrivals at
      2000/s, m^2 elements with m log-uniform in [16, 96] at p = 6, 1..50 steps
      each) on the request server; value = unknowns of all inversions / wall
      time, with p50/p99 latency and steps/s in the config, beside the same
      stream served with plain FP64 inversions.

value = unknowns solved / second (GDOF/s to solution), with every input
resident in HBM before the timed region.  The JSON line carries the roofline
of the dominant kernel, timed with HIP events on the stream it runs on, and a
CPU baseline: the CPU oracle's assembled-CSR Jacobi-PCG (the reference's
algorithm) on all host threads (and on one), on a bounded sample, scaled to
the full workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Helmholtz CG GDOF/s at 1e-10 rel. residual vs FP64; p50/p99 latency under load"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector
# MI355X dense matrix-core peaks by operand type (TFLOP/s)
MFMA_PEAK_TFLOPS = {"fp64": 78.6, "fp64_fp32c": 78.6, "fp64_bf16c": 78.6, "fp32_ir": 157.3, "bf16_ir": 2500.0}
DTYPE = {"fp64": "f64", "fp64_fp32c": "f64+f32", "fp64_bf16c": "f64+bf16", "fp32_ir": "f32+f64",
         "bf16_ir": "bf16+f64"}
# the element type of the probed GEMM (for fp64_*c: the first application's, FP64)
GEMM_T = {"fp64": "double", "fp64_fp32c": "double", "fp64_bf16c": "double", "fp32_ir": "float", "bf16_ir": "bf16"}


def prec_code(name):
    import se_amd
    return {"fp64": se_amd.PREC_FP64, "fp64_fp32c": se_amd.PREC_FP64_FP32C, "fp64_bf16c": se_amd.PREC_FP64_BF16C,
            "fp32_ir": se_amd.PREC_FP32_IR, "bf16_ir": se_amd.PREC_BF16_IR}[name]


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="ensemble_512")
    p.add_argument("--precision", default=None, choices=["fp64", "fp64_fp32c", "fp64_bf16c", "fp32_ir", "bf16_ir"],
                   help="default: fp64_bf16c (FDM: first application on the FP64 matrix cores, corrections "
                        "with BF16 operands) for ensemble_512, fp32_ir for "
                        "qg2_stream, fp64 for helmholtz_*")
    p.add_argument("--pc", default="default", choices=["default", "jacobi", "fdm"],
                   help="preconditioner: default = FDM for ensemble_512 and the mixed precisions, Jacobi (the "
                        "reference's algorithm) for fp64 on helmholtz_* and qg2_stream")
    p.add_argument("--members", type=int, default=512, help="ensemble size")
    p.add_argument("--detj", default="physical", choices=["physical", "literal"])
    p.add_argument("--cpu-iters", type=int, default=12, help="CG iterations per CPU baseline sample")
    p.add_argument("--cpu-members", type=int, default=16, help="ensemble members in the CPU sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison")
    p.add_argument("--probe-reps", type=int, default=50)
    p.add_argument("--check-every", type=int, default=32,
                   help="mixed-precision solves: CG iterations between host checks (one graph replay)")
    p.add_argument("--requests", type=int, default=32, help="qg2_stream: number of requests")
    p.add_argument("--stream-requests", type=int, default=16,
                   help="ensemble_512: requests of the served-stream leg that gives the latency figures (0: none)")
    p.add_argument("--budget", type=float, default=200e9, help="qg2_stream: server memory budget (bytes)")
    return p.parse_args()


def apply_bytes(m):
    # algorithmic bytes of one FP64 operator apply: x read, y written, 8 B each
    # per unknown (the kernel works on the unknown grid: no map is read)
    return 16.0 * m.n


def apply_flops(m):
    # per element: 4 contractions of (N+1)^3 multiply-adds, ~30 flops of metric
    # per quadrature point, mass term; once per element
    n1 = m.N + 1
    return m.nElsTotal * (8.0 * n1 ** 3 + 32.0 * n1 * n1)


def cpu_model():
    """The host CPU's model string (as lscpu reports it) and its logical CPU count."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_sample(cfg, f, iters_gpu, sample_iters, threads=1):
    """Oracle CSR Jacobi-PCG (reference algorithm) on `threads` host threads
    (row-parallel SpMV and vector updates, fixed-block dot products): seconds
    per iteration on this system, scaled to the reference algorithm's
    iteration count iters_gpu."""
    from oracle import oracle as O
    O.set_threads(threads)
    om = O.OracleMesh(cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], cfg["periodic"], cfg["bc"])
    t0 = time.perf_counter()
    oh = O.Helm(om, cfg["c_mass"], cfg["c_lap"], cfg["detj_mode"])
    t_asm = time.perf_counter() - t0
    oh.pcg(f, rtol=1e-30, maxit=1)
    t0 = time.perf_counter()
    _, info = oh.pcg(f, rtol=1e-30, maxit=sample_iters)
    t_it = (time.perf_counter() - t0) / max(info["iters"], 1)
    return dict(t_it=t_it, t_solve=t_it * iters_gpu, t_asm=t_asm, iters=info["iters"],
                nnz=oh.L.orc_helm_nnz(oh.h), n=om.n)


def host_threads():
    """Threads for the all-core CPU baseline: the CPUs this process may run on."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def timed(ctx, stream, steps, fn):
    """Run fn() `steps` times between HIP events on the library's stream."""
    import torch
    torch.cuda.synchronize()
    ctx.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    res = []
    t_wall = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        res.append(fn())
    ev1.record(stream)
    ev1.synchronize()
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps, time.perf_counter() - t_wall, res


def base_line(a, value, ms, dtype, config, roofline, cpu):
    return {"metric": METRIC, "value": value, "unit": "GDOF/s", "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config,
            "roofline": roofline, "cpu_baseline": cpu}


COPY_BYTES = 1 << 30  # yardstick copy: 1 GiB each way, far beyond the 256 MB last-level cache


def copy_yardstick(ctx, reps=5):
    """A plain device copy of known size, timed like the roofline kernel: the
    attainable streaming bandwidth, and the reference for the byte counters."""
    us = ctx.copy_probe(COPY_BYTES, reps)
    return {"kernel": "k_copy_probe", "bytes_read": COPY_BYTES, "bytes_written": COPY_BYTES,
            "us_per_launch": us, "achieved": 2.0 * COPY_BYTES / (us * 1e-6) / 1e9}


def pmc_traffic(roof, profile):
    """HBM bytes per launch of the roofline kernel from its committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by
    what the counters read for the copy yardstick); None when the summary is
    absent or was taken on another kernel."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    fields = {}
    for line in open(path):
        if ":" in line and not line.startswith(("|", "#", " ")):
            k, v = line.split(":", 1)
            fields[k.strip("- `")] = v.strip().strip("`")
    if fields.get("kernel") != roof["kernel"] or "traffic_bytes_per_launch" not in fields:
        return None
    return float(fields["traffic_bytes_per_launch"])


def mixed_roofline(kernel, pr):
    """pr: us_dir, us_upd, bytes_dir, bytes_upd of the inner iteration kernels."""
    ach = pr["bytes_dir"] / (pr["us_dir"] * 1e-6) / 1e9
    upd = pr["bytes_upd"] / (pr["us_upd"] * 1e-6) / 1e9
    return {"kernel": kernel, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": None, "us_per_launch": pr["us_dir"],
            "bytes_per_launch": pr["bytes_dir"],
            "update_kernel": {"us_per_launch": pr["us_upd"], "bytes_per_launch": pr["bytes_upd"],
                              "achieved": upd, "frac": upd / HBM_PEAK_GBS}}


def pc_code(a, prec_name, fdm_default=False):
    """--pc, or by default FDM for the mixed precisions (and for FP64 where
    fdm_default), Jacobi (the reference's algorithm) for FP64 otherwise."""
    import se_amd
    if a.pc == "jacobi":
        return se_amd.PC_JACOBI
    if a.pc == "fdm" or prec_name != "fp64" or fdm_default:
        return se_amd.PC_FDM
    return se_amd.PC_JACOBI


def fdm_roofline(pr, prec_name, kernel):
    """pr: per-pass us / flops of the FDM transforms and us / bytes of the FP64 apply."""
    us, fl = sum(pr["us_pass"]), sum(pr["flops_pass"])
    ach = fl / (us * 1e-6) / 1e12
    peak = MFMA_PEAK_TFLOPS[prec_name]
    apply_gbs = pr["bytes_apply"] / (pr["us_apply"] * 1e-6) / 1e9
    return {"kernel": kernel, "bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
            "frac": ach / peak, "traffic": None, "us_per_launch": us / 4.0, "flops_per_launch": fl / 4.0,
            "passes": [{"us": u, "flops": f, "tflops": f / (u * 1e-6) / 1e12} for u, f in zip(pr["us_pass"],
                                                                                             pr["flops_pass"])],
            "fp64_apply": {"us_per_launch": pr["us_apply"], "bytes_per_launch": pr["bytes_apply"],
                           "achieved_gbs": apply_gbs, "frac_hbm": apply_gbs / HBM_PEAK_GBS}}


def run_ensemble(a, prec_name):
    import numpy as np
    import torch

    import se_amd
    from se_amd import workloads as W
    prec = prec_code(prec_name)
    pc = pc_code(a, prec_name, fdm_default=True)
    cfgs = W.ensemble_512(count=a.members, detj_mode=1 if a.detj == "physical" else 0)
    ctx = se_amd.Context(0)
    dev = torch.device("cuda:0")
    hs, fs, xs = [], [], []
    for c in cfgs:
        m = se_amd.Mesh(ctx, c["nEls"], c["N"], c["lo"], c["hi"], c["periodic"], c["bc"])
        h = se_amd.Helm(m, c["c_mass"], c["c_lap"], c["detj_mode"])
        g = torch.from_numpy(W.make_rhs(c, m.gll()[0])).to(dev)
        f = torch.empty(m.n, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        se_amd.check(se_amd.lib().se_helm_rhs_dev(h.h, g.data_ptr(), f.data_ptr()), "rhs")
        hs.append(h)
        fs.append(f)
        xs.append(torch.zeros(m.n, dtype=torch.float64, device=dev))
    ctx.synchronize()
    ntot = sum(h.mesh.n for h in hs)
    b = se_amd.Batch(ctx, hs)
    fp, xp = [f.data_ptr() for f in fs], [x.data_ptr() for x in xs]
    stream = torch.cuda.ExternalStream(ctx.stream)

    def step(p=prec, c=pc, x=xp):
        return b.pcg_dev(fp, x, precision=p, rtol=1e-10, check_every=a.check_every, preconditioner=c)

    for _ in range(a.warmup):
        step()
    ms, t_wall, res = timed(ctx, stream, a.steps, step)
    infos = res[-1]
    ok = all(i["status"] == 0 and i["true_rel_residual"] <= 1e-10 for r in res for i in r)
    its = np.array([i["iters"] for i in infos])
    value = ntot / (ms * 1e-3) / 1e9
    T = GEMM_T[prec_name]
    if pc == se_amd.PC_FDM:
        # dominant kernel: the FDM transforms on the matrix cores (four passes
        # over all members, timed back to back after the last solve)
        roof = fdm_roofline(b.fdm_probe(reps=a.probe_reps), prec_name, "k_fdm_gemm<%s>" % T)
        roof["traffic"] = pmc_traffic(roof, "k_fdm_gemm__ensemble_512_%s_pmc.md" % prec_name.split("_")[0])
    else:
        # dominant kernel: the direction + separable operator kernel of the
        # largest order group (the probe continues the last solve's recurrences)
        groups = b.probe(prec, reps=a.probe_reps)
        gmax = max(groups, key=lambda g: g["bytes_dir"])
        roof = mixed_roofline("k_tcg_dir_apply_batch<%d, %s, NORMAL>" % (gmax["order"], T), gmax)
        roof["groups"] = groups
        roof["traffic"] = pmc_traffic(roof, "k_tcg_dir_apply_batch__ensemble_512_%s_pmc.md" % prec_name.split("_")[0])
    roof["copy_yardstick"] = copy_yardstick(ctx)
    config = {"workload": "ensemble_512" if a.members == 512 else "ensemble_%d" % a.members,
              "members": len(hs), "unknowns": ntot, "precision": prec_name,
              "preconditioner": "fdm" if pc == se_amd.PC_FDM else "jacobi", "detj": a.detj, "rtol": 1e-10,
              "iters_min_median_max": [int(its.min()), int(np.median(its)), int(its.max())],
              "reliable_updates_max": max(i["outer_iters"] for i in infos), "all_converged": ok,
              "max_true_rel_residual": max(i["true_rel_residual"] for i in infos), "wall_s_timed": t_wall}
    if pc != se_amd.PC_FDM:
        config.update({"tuned_workgroups_per_cu": dict(zip(("direction", "update"), b.launch_shape())),
                       "check_every": a.check_every})
    if not a.no_fp64:
        # the plain FP64 GPU path on the same members, one after another
        xs64 = [torch.zeros_like(x) for x in xs]

        def step64():
            return [h.pcg_dev(f, x.data_ptr(), rtol=1e-10) for h, f, x in zip(hs, fp, xs64)]
        ms64, _, res64 = timed(ctx, stream, 1, step64)
        v64 = ntot / (ms64 * 1e-3) / 1e9
        diff = max(float(torch.linalg.norm(x - y) / torch.linalg.norm(y)) for x, y in zip(xs, xs64))
        config.update({"fp64_gpu_value": v64, "fp64_gpu_ms": ms64, "speedup_vs_fp64_gpu": value / v64,
                       "fp64_all_converged": all(i["status"] == 0 for i in res64[-1]),
                       "max_rel_diff_vs_fp64_solution": diff})
        # the other legs of the same batch, to separate the gains: batching
        # (FP64 FDM batched), precision (this line's FDM at another
        # precision) and the preconditioner (the inner Jacobi-PCG mixed solve)
        legs = {}
        for lab, p2, c2 in (("fp64_fdm_batched", se_amd.PREC_FP64, se_amd.PC_FDM),
                            ("fp64_fp32c_fdm_batched", se_amd.PREC_FP64_FP32C, se_amd.PC_FDM),
                            ("fp64_bf16c_fdm_batched", se_amd.PREC_FP64_BF16C, se_amd.PC_FDM),
                            ("fp32_fdm_batched", se_amd.PREC_FP32_IR, se_amd.PC_FDM),
                            ("bf16_fdm_batched", se_amd.PREC_BF16_IR, se_amd.PC_FDM),
                            ("fp32_jacobi_batched", se_amd.PREC_FP32_IR, se_amd.PC_JACOBI)):
            if p2 == prec and c2 == pc:
                continue
            xs2 = [torch.zeros_like(x) for x in xs]
            xp2 = [x.data_ptr() for x in xs2]
            step(p2, c2, xp2)  # warm (setup of that precision)
            ms2, _, r2 = timed(ctx, stream, 1, lambda: step(p2, c2, xp2))
            legs[lab] = {"ms": ms2, "value": ntot / (ms2 * 1e-3) / 1e9,
                         "all_converged": all(i["status"] == 0 and i["true_rel_residual"] <= 1e-10 for i in r2[-1]),
                         "iters_max": max(i["iters"] for i in r2[-1])}
        config["other_legs"] = legs
    if a.stream_requests > 0:
        # the metric's latency half: the served two-layer QG stream (its first
        # requests, Poisson arrivals at 2000/s, 16^2..96^2 elements at p = 6),
        # served with this line's precision and preconditioner, by a server of
        # a process that has served them once before (as a long-lived server
        # has: the line operators' eigen-decompositions are then known)
        sreqs = stream_requests(a.stream_requests)
        serve_stream(sreqs, prec, a.budget, pc, prec_name + " warm-up")
        sr = serve_stream(sreqs, prec, a.budget, pc, prec_name)
        ss = stream_summary(sr)
        config.update({"latency_p50_s": ss["latency_p50_s"], "latency_p99_s": ss["latency_p99_s"],
                       "steps_per_s": ss["steps_per_s"]})
        ss.update({"requests": a.stream_requests, "rate_per_s": 2000.0, "warm_up_runs": 1})
        config["stream"] = ss
    cpu = None
    if not a.no_cpu_baseline:
        iters_ref = [i["iters"] for i in res64[-1]] if not a.no_fp64 else [i["iters"] for i in infos]
        cpu = ensemble_cpu(a, cfgs, hs, fs, iters_ref)
    return base_line(a, value, ms, DTYPE[prec_name], config, roof, cpu)


def ensemble_cpu(a, cfgs, hs, fs, iters_ref):
    """Every (512 / cpu_members)-th member: oracle CSR Jacobi-PCG time per
    iteration x the reference algorithm's iteration count for that member
    (the plain FP64 GPU path's, the same algorithm), scaled to all members;
    on all host threads, and on one (the serial reference)."""
    step = max(1, len(cfgs) // a.cpu_members)
    nth = host_threads()
    out = {}
    for th in (nth, 1):
        t_all, n_all, samples = 0.0, 0, 0
        for i in range(0, len(cfgs), step):
            s = cpu_sample(cfgs[i], fs[i].cpu().numpy(), iters_ref[i], a.cpu_iters, threads=th)
            t_all += s["t_solve"]
            n_all += s["n"]
            samples += 1
        out[th] = n_all / t_all / 1e9
    model, ncpu = cpu_model()
    return {"value": out[nth], "unit": "GDOF/s", "cores": nth, "kind": "port", "cpu_model": model,
            "logical_cpus": ncpu, "single_thread_value": out[1],
            "sample": "members 0, %d, 2x%d, ... (%d of %d): %d Jacobi-PCG iterations each of the assembled "
                      "PETSc-pattern CSR (the reference's algorithm) on %d threads (and on 1), scaled to each "
                      "member's iteration count with that algorithm; CSR assembly not counted"
                      % (step, step, samples, len(cfgs), a.cpu_iters, nth)}


def run_single(a, prec_name):
    import torch

    import se_amd
    from se_amd import workloads as W
    parts = a.workload.split("_")
    ne = int(parts[1]) if len(parts) > 1 else 128
    N = int(parts[2][1:]) if len(parts) > 2 else 8
    cfg = W.helmholtz_128(nEls=ne, N=N, detj_mode=1 if a.detj == "physical" else 0)
    prec = prec_code(prec_name)
    pc = pc_code(a, prec_name)
    ctx = se_amd.Context(0)
    m = se_amd.Mesh(ctx, cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], cfg["periodic"], cfg["bc"])
    h = se_amd.Helm(m, cfg["c_mass"], cfg["c_lap"], cfg["detj_mode"])
    rhs_field = W.make_rhs(cfg, m.gll()[0])
    dev = torch.device("cuda:0")
    d_rhs = torch.from_numpy(rhs_field).to(dev)
    d_f = torch.empty(m.n, dtype=torch.float64, device=dev)
    d_x = torch.empty(m.n, dtype=torch.float64, device=dev)
    d_y = torch.empty(m.n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    se_amd.check(se_amd.lib().se_helm_rhs_dev(h.h, d_rhs.data_ptr(), d_f.data_ptr()), "rhs")
    ctx.synchronize()
    stream = torch.cuda.ExternalStream(ctx.stream)

    def solve(p=prec, x=d_x, c=pc):
        return h.pcg_dev(d_f.data_ptr(), x.data_ptr(), rtol=1e-10, precision=p, check_every=a.check_every,
                         preconditioner=c)

    for _ in range(a.warmup):
        solve()
    ms, t_wall, res = timed(ctx, stream, a.steps, solve)
    info = res[-1]
    ok = all(i["status"] == 0 and i["true_rel_residual"] <= 1e-10 for i in res)
    value = m.n / (ms * 1e-3) / 1e9
    config = {"workload": cfg["workload"], "nEls": list(cfg["nEls"]), "N": cfg["N"], "lambda": cfg["c_mass"],
              "n": m.n, "precision": prec_name, "preconditioner": "fdm" if pc == se_amd.PC_FDM else "jacobi",
              "detj": a.detj, "rtol": 1e-10, "iters": info["iters"],
              "outer_iters": info["outer_iters"], "true_rel_residual": info["true_rel_residual"],
              "all_converged": ok, "n_iter_per_s": m.n * info["iters"] / (ms * 1e-3), "wall_s_timed": t_wall}
    T = GEMM_T[prec_name]
    if pc == se_amd.PC_FDM:
        roof = fdm_roofline(h.fdm_probe(reps=a.probe_reps), prec_name, "k_fdm_gemm<%s>" % T)
    elif prec == se_amd.PREC_FP64:
        # dominant kernel: the FP64 operator apply, timed on its stream
        for _ in range(5):
            h.apply_dev(d_x.data_ptr(), d_y.data_ptr())

        def app():
            h.apply_dev(d_x.data_ptr(), d_y.data_ptr())
        ms_apply, _, _ = timed(ctx, stream, a.probe_reps, app)
        t_apply = ms_apply * 1e-3
        ach = apply_bytes(m) / t_apply / 1e9
        fl = apply_flops(m)
        roof = {"kernel": "k_apply_lines<%d, false> (FP64 operator apply)" % m.N, "bound": "hbm",
                "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": None, "us_per_launch": t_apply * 1e6, "bytes_per_launch": apply_bytes(m),
                "fp64_tflops": fl / t_apply / 1e12, "fp64_frac": fl / t_apply / 1e12 / FP64_PEAK_TFLOPS}
    else:
        pr = h.probe(prec, reps=a.probe_reps)
        roof = mixed_roofline("k_tcg_dir_apply<%d, %s, NORMAL>" % (m.N, T), pr)
    if cfg["workload"] == "helmholtz_128" and pc != se_amd.PC_FDM:
        roof["traffic"] = pmc_traffic(roof, "%s__helmholtz_128_%s_pmc.md" % (
            roof["kernel"].split("<")[0], prec_name.split("_")[0]))
    roof["copy_yardstick"] = copy_yardstick(ctx)
    if not (prec == se_amd.PREC_FP64 and pc == se_amd.PC_JACOBI) and not a.no_fp64:
        # the plain FP64 GPU path: the reference's algorithm (Jacobi-PCG) in FP64
        d_x64 = torch.empty_like(d_x)
        ms64, _, r64 = timed(ctx, stream, 1, lambda: solve(se_amd.PREC_FP64, d_x64, se_amd.PC_JACOBI))
        config.update({"fp64_gpu_value": m.n / (ms64 * 1e-3) / 1e9, "fp64_gpu_ms": ms64,
                       "speedup_vs_fp64_gpu": ms64 / ms,
                       "max_rel_diff_vs_fp64_solution":
                           float(torch.linalg.norm(d_x - d_x64) / torch.linalg.norm(d_x64))})
        legs = {}
        for lab, p2, c2 in (("fp64_fdm", se_amd.PREC_FP64, se_amd.PC_FDM),
                            ("fp32_fdm", se_amd.PREC_FP32_IR, se_amd.PC_FDM),
                            ("bf16_fdm", se_amd.PREC_BF16_IR, se_amd.PC_FDM),
                            ("fp32_jacobi", se_amd.PREC_FP32_IR, se_amd.PC_JACOBI)):
            if p2 == prec and c2 == pc:
                continue
            d_x2 = torch.empty_like(d_x)
            solve(p2, d_x2, c2)
            ms2, _, r2 = timed(ctx, stream, 1, lambda: solve(p2, d_x2, c2))
            legs[lab] = {"ms": ms2, "value": m.n / (ms2 * 1e-3) / 1e9, "iters": r2[-1]["iters"],
                         "converged": r2[-1]["status"] == 0 and r2[-1]["true_rel_residual"] <= 1e-10}
        config["other_legs"] = legs
    cpu = None
    if not a.no_cpu_baseline:
        # the reference's algorithm: Jacobi-PCG, as many iterations as the plain FP64 GPU path needs
        it_ref = info["iters"] if (prec == se_amd.PREC_FP64 and pc == se_amd.PC_JACOBI) else (
            r64[-1]["iters"] if "fp64_gpu_ms" in config else info["iters"])
        nth = host_threads()
        s = cpu_sample(cfg, d_f.cpu().numpy(), it_ref, a.cpu_iters, threads=nth)
        s1 = cpu_sample(cfg, d_f.cpu().numpy(), it_ref, a.cpu_iters, threads=1)
        model, ncpu = cpu_model()
        cpu = {"value": m.n / s["t_solve"] / 1e9, "unit": "GDOF/s", "cores": nth, "kind": "port",
               "cpu_model": model, "logical_cpus": ncpu, "single_thread_value": m.n / s1["t_solve"] / 1e9,
               "sample": "%d Jacobi-PCG iterations of the assembled PETSc-pattern CSR (%d nnz) on the same "
                         "system, %d threads (and 1), scaled to the reference algorithm's %d-iteration solve; "
                         "CSR assembly %.1f s not counted" % (s["iters"], s["nnz"], nth, it_ref, s["t_asm"]),
               "seconds_per_iteration": s["t_it"], "seconds_per_iteration_1_thread": s1["t_it"]}
    return base_line(a, value, ms, DTYPE[prec_name], config, roof, cpu)


def stream_requests(n):
    """The served stream's first n requests with their seeded initial fields."""
    import numpy as np

    import se_amd
    from se_amd import workloads as W
    reqs = W.qg2_stream(n_requests=n)
    ctx = se_amd.Context(0)
    abc = {}
    for r in reqs:
        if r["N"] not in abc:
            m = se_amd.Mesh(ctx, (2, 2), r["N"], (0, 0), (1, 1), (0, 0), (0, 0, 0, 0))
            abc[r["N"]] = m.gll()[0]
            m.close()
        X, Y = W.mesh_coords(r["nEls"], r["N"], r["lo"], r["hi"], abc[r["N"]])
        r["init"] = W.qg2_initial(r, X, Y)
    ctx.close()
    return reqs


def serve_stream(reqs, prec, budget, pc=0, tag=""):
    """Submit the requests at their Poisson arrival times to a fresh server and
    wait for all; wall time, per-request latency (submit to done), steps."""
    import numpy as np

    import se_amd
    sv = se_amd.Server(memory_budget=budget, precision=prec, rtol=1e-10, preconditioner=pc)
    t0 = time.perf_counter()
    ids = []
    for r in reqs:  # submit at the arrival times
        wait = r["t_arrive"] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        ids.append(sv.submit(r, *r["init"]))
    res = []
    for k, i in enumerate(ids):
        res.append(sv.wait(i, fields=False)[0])
        if (k + 1) % 8 == 0 or k + 1 == len(ids):
            print("[qg2_stream %s] %d/%d requests done, %.1f s" % (tag, k + 1, len(ids), time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    st = sv.stats()
    sv.close()
    lat = np.array([x["t_done"] - x["t_submit"] for x in res])
    ok = all(x["status"] == 0 for x in res)
    return dict(wall=wall, lat=lat, ok=ok, steps=sum(x["steps_done"] for x in res),
                unknowns=sum(x["unknowns"] for x in res), ticks=st["ticks"],
                iters=sum(x["inner_iters"] for x in res))


def stream_summary(r):
    import numpy as np
    return {"steps_per_s": r["steps"] / r["wall"], "latency_p50_s": float(np.percentile(r["lat"], 50)),
            "latency_p99_s": float(np.percentile(r["lat"], 99)), "wall_s": r["wall"],
            "value": r["unknowns"] / r["wall"] / 1e9, "steps_served": r["steps"], "all_converged": r["ok"],
            "ticks": r["ticks"]}


def run_stream(a, prec_name):
    """Served two-layer QG step requests arriving as a Poisson process."""
    import se_amd
    prec = prec_code(prec_name)
    pc = pc_code(a, prec_name)
    reqs = stream_requests(a.requests)
    for _ in range(a.warmup):
        serve_stream(reqs, prec, a.budget, pc, prec_name)
    runs = [serve_stream(reqs, prec, a.budget, pc, prec_name) for _ in range(a.steps)]
    r = min(runs, key=lambda x: x["wall"])
    value = r["unknowns"] / r["wall"] / 1e9
    config = {"workload": "qg2_stream_%d" % a.requests, "requests": a.requests, "rate_per_s": 2000.0,
              "mesh": "m^2 elements, m log-uniform in [16, 96], p = 6, x-periodic channel", "precision": prec_name,
              "preconditioner": "fdm" if pc == se_amd.PC_FDM else "jacobi",
              "rtol": 1e-10, "memory_budget_bytes": a.budget, "inner_iterations": r["iters"]}
    config.update(stream_summary(r))
    if not a.no_fp64 and not (prec == se_amd.PREC_FP64 and pc == se_amd.PC_JACOBI):
        r64 = serve_stream(reqs, se_amd.PREC_FP64, a.budget, se_amd.PC_JACOBI, "fp64")
        config.update({"fp64_gpu_value": r64["unknowns"] / r64["wall"] / 1e9,
                       "fp64_steps_per_s": r64["steps"] / r64["wall"],
                       "fp64_latency_p50_s": stream_summary(r64)["latency_p50_s"],
                       "fp64_latency_p99_s": stream_summary(r64)["latency_p99_s"],
                       "speedup_vs_fp64_gpu": r64["wall"] / r["wall"]})
    return base_line(a, value, r["wall"] * 1e3, DTYPE[prec_name], config, None, None)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU only: se_amd has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    if a.workload.startswith("ensemble"):
        prec = a.precision or "fp64_bf16c"
        if prec.startswith("fp64") and a.pc == "jacobi":
            raise SystemExit("batched FP64 solves take the FDM preconditioner (--pc fdm); the plain FP64 "
                             "Jacobi path is timed member by member inside every ensemble line")
        out = run_ensemble(a, prec)
    elif a.workload.startswith("qg2_stream"):
        out = run_stream(a, a.precision or "fp32_ir")
    elif a.workload.startswith("helmholtz"):
        out = run_single(a, a.precision or "fp64")
    else:
        raise SystemExit("unknown workload %s" % a.workload)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
