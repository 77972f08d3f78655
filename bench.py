"""Benchmark: Helmholtz CG throughput on one MI355X.

Workload (default): helmholtz_128 -- 128x128 elements at GLL order 8 on the
doubly periodic [0, 2pi]^2 (1 048 576 unknowns), lambda = 100, right-hand side
a seeded random field with an isotropic k^-3 spectrum.  One step is one
Jacobi-PCG solve of (lambda M + K) x = f to a true relative residual of 1e-10,
with f and x resident in HBM when the timed region starts.

value = unknowns / (seconds per solve) in GDOF/s.  The JSON line also carries
the roofline of the dominant kernel (the matrix-free operator apply, timed live
with HIP events on the stream it runs on) and a CPU baseline: the CPU oracle's
assembled-CSR Jacobi-PCG (the reference's algorithm) timed for a bounded number
of iterations on this host and scaled to a full solve.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="helmholtz_128")
    p.add_argument("--precision", default="fp64", choices=["fp64", "fp32_ir", "bf16_ir"])
    p.add_argument("--detj", default="physical", choices=["physical", "literal"])
    p.add_argument("--cpu-iters", type=int, default=12, help="CG iterations in the CPU baseline sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--apply-reps", type=int, default=50)
    return p.parse_args()


def apply_bytes(m):
    # algorithmic bytes of one operator apply: x read, y written (8 B each per
    # unknown) and the int32 numbering map read once per field node
    return 16.0 * m.n + 4.0 * m.nV


def apply_flops(m):
    # per element: 4 contractions of (N+1)^3 multiply-adds, ~30 flops of metric
    # per quadrature point, mass term; evaluated once per element (the
    # owner-computes halo elements are not counted)
    n1 = m.N + 1
    return m.nElsTotal * (8.0 * n1 ** 3 + 32.0 * n1 * n1)


def cpu_baseline(cfg, f, iters_gpu, sample_iters):
    """Oracle CSR Jacobi-PCG (reference algorithm), single thread, bounded sample."""
    import numpy as np

    from oracle import oracle as O
    om = O.OracleMesh(cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], cfg["periodic"], cfg["bc"])
    t0 = time.perf_counter()
    oh = O.Helm(om, cfg["c_mass"], cfg["c_lap"], cfg["detj_mode"])
    t_asm = time.perf_counter() - t0
    # warm up one iteration, then time the bounded sample
    oh.pcg(f, rtol=1e-30, maxit=1)
    t0 = time.perf_counter()
    _, info = oh.pcg(f, rtol=1e-30, maxit=sample_iters)
    t_it = (time.perf_counter() - t0) / max(info["iters"], 1)
    t_solve = t_it * iters_gpu
    return {
        "value": om.n / t_solve / 1e9,
        "unit": "GDOF/s",
        "cores": 1,
        "kind": "port",
        "sample": "%d Jacobi-PCG iterations of the assembled PETSc-pattern CSR (%d nnz) on the same "
                  "system, single thread, scaled to the GPU's %d-iteration solve; CSR assembly %.1f s "
                  "not counted" % (info["iters"], oh.L.orc_helm_nnz(oh.h), iters_gpu, t_asm),
        "seconds_per_iteration": t_it,
    }


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import se_amd
    from se_amd import workloads as W

    if a.workload.startswith("helmholtz"):
        parts = a.workload.split("_")
        ne = int(parts[1]) if len(parts) > 1 else 128
        N = int(parts[2][1:]) if len(parts) > 2 else 8
        cfg = W.helmholtz_128(nEls=ne, N=N, detj_mode=1 if a.detj == "physical" else 0)
    else:
        raise SystemExit("unknown workload %s" % a.workload)
    prec = {"fp64": se_amd.PREC_FP64, "fp32_ir": se_amd.PREC_FP32_IR, "bf16_ir": se_amd.PREC_BF16_IR}[a.precision]

    ctx = se_amd.Context(0)
    m = se_amd.Mesh(ctx, cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], cfg["periodic"], cfg["bc"])
    h = se_amd.Helm(m, cfg["c_mass"], cfg["c_lap"], cfg["detj_mode"])
    abc, _, _ = m.gll()
    rhs_field = W.make_rhs(cfg, abc)

    dev = torch.device("cuda:0")
    d_rhs = torch.from_numpy(rhs_field).to(dev)
    d_f = torch.empty(m.n, dtype=torch.float64, device=dev)
    d_x = torch.empty(m.n, dtype=torch.float64, device=dev)
    d_y = torch.empty(m.n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    se_amd.check(se_amd.lib().se_helm_rhs_dev(h.h, d_rhs.data_ptr(), d_f.data_ptr()), "rhs")
    ctx.synchronize()
    stream = torch.cuda.ExternalStream(ctx.stream)

    def solve():
        return h.pcg_dev(d_f.data_ptr(), d_x.data_ptr(), rtol=1e-10, precision=prec, check_every=32)

    for _ in range(a.warmup):
        info = solve()
    torch.cuda.synchronize()
    ctx.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    infos = []
    t_wall = time.perf_counter()
    ev0.record(stream)
    for _ in range(a.steps):
        infos.append(solve())
    ev1.record(stream)
    ev1.synchronize()
    t_wall = time.perf_counter() - t_wall
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / a.steps
    info = infos[-1]
    ok = all(i["status"] == 0 and i["true_rel_residual"] <= 1e-10 for i in infos)

    # dominant kernel: the operator apply, timed on the stream it is launched on
    for _ in range(5):
        h.apply_dev(d_x.data_ptr(), d_y.data_ptr())
    ea, eb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ea.record(stream)
    for _ in range(a.apply_reps):
        h.apply_dev(d_x.data_ptr(), d_y.data_ptr())
    eb.record(stream)
    eb.synchronize()
    t_apply = ea.elapsed_time(eb) / a.apply_reps * 1e-3
    bytes_apply = apply_bytes(m)
    achieved = bytes_apply / t_apply / 1e9
    flops_apply = apply_flops(m)

    iters = info["iters"]
    value = m.n / (ms * 1e-3) / 1e9
    out = {
        "metric": "Helmholtz CG GDOF/s at 1e-10 rel. residual vs FP64; p50/p99 latency under load",
        "value": value,
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32_ir": "f32+f64", "bf16_ir": "bf16+f64"}[a.precision],
        "data": "synthetic",
        "config": {
            "workload": cfg["workload"], "nEls": list(cfg["nEls"]), "N": cfg["N"], "lambda": cfg["c_mass"],
            "n": m.n, "precision": a.precision, "detj": a.detj, "rtol": 1e-10, "iters": iters,
            "outer_iters": info["outer_iters"], "true_rel_residual": info["true_rel_residual"],
            "all_converged": ok, "n_iter_per_s": m.n * iters / (ms * 1e-3),
            "wall_s_timed": t_wall,
        },
        "roofline": {
            "kernel": "k_helm_tile (operator apply)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "apply_us": t_apply * 1e6,
            "bytes_per_apply": bytes_apply,
            "fp64_tflops": flops_apply / t_apply / 1e12,
            "fp64_frac": flops_apply / t_apply / 1e12 / FP64_PEAK_TFLOPS,
        },
        "cpu_baseline": None,
    }
    if not a.no_cpu_baseline:
        f_host = d_f.cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(cfg, f_host, iters, a.cpu_iters)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
