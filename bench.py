#!/usr/bin/env python3
"""Throughput bench of the compact3d hot path on one MI355X.

Workload: one time step of the reference's per-step path -- GETUP (RHS1-3,
ADI GETUH1-3, RHSDP + FFT/Thomas GETDP, UPCALC) followed by DIVCHECK and CFL
of the new field (MAIN, compact3d.f:130-138) -- on the channel
256 x 193 x 256 (2pi x 2 x pi, tanh stretching 1.8), Re_tau = 395 (bulk
Re_b(2 delta) = 13750, RE = 0.75 Re_b = 10312.5 in the reference's units),
dt = 7.5e-4 (= 5e-4 in bulk-velocity units), seeded perturbed-Poiseuille
initial field resident in HBM.

Metric: Mpts.steps/s = (N1M*N2M*N3M) * steps / wall-seconds / 1e6 over the
timed steps (state resident, no host transfers).

Prints one JSON line (see the keys below).  `roofline` is for the kernel with
the largest share of the step, from HIP events recorded by the library around
each launch on its own stream; `cpu_baseline` times the FP64 oracle
(oracle/liboracle.so, OpenMP) on a few steps of the same grid.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mpts·steps/s mixed vs FP64 on 1 GPU @ div≤1e-9; p50/p99 served latency"

# algorithmic (compulsory) HBM bytes per cell and launch of each kernel of the
# FP64 step: 8 B per field element read or written once
ALG_BYTES = {
    # compulsory bytes per pressure cell of each kernel (every distinct field
    # read once, written once); adi_y is the ADI y-sweep with the RHS y terms
    # fused in (per component: state, VV1/VV3 or P, RUH, HH, HH2 in; HH, UH out)
    "rhs_x": 88, "rhs_z": 112, "rhs_zw": 40, "rhs_x2": 24, "adi_y": 168, "adi_z": 64,
    "rhsdp_x": 24, "rhsdp_z": 24, "fft_x": 16, "fft_z": 16, "pois_y": 16, "fft_xinv": 16,
    "flow_red": 40, "upd_x": 48, "upd_z": 64,
}
# adi_x runs twice per step: (u,v) read+write 32 B, then w with U added 24 B
ALG_BYTES_ADI_X_PER_STEP = 56


def workload(grid):
    n1m, n2, n3m = grid
    return dict(n1=n1m + 1, n2=n2, n3=n3m + 1, re=10312.5, alx=2 * math.pi, aly=2.0, alz=math.pi,
                dt=7.5e-4, files=0.0, wmodel=0.0, cks=1.8)


def cpu_baseline(cfg, target_s, seed):
    """FP64 oracle on the host cores: a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from compact3d_amd.initfield import perturbed_poiseuille
    from pyoracle import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    Oracle.set_threads(threads)
    o = Oracle(**cfg)
    m = o.mesh()
    U, P = perturbed_poiseuille(cfg["n1"], cfg["n2"], cfg["n3"], m["Y"], m["DX1"], m["DX3"], m["ALY"], seed=seed)
    HH = o.zeros_u()
    S = o.zeros_u(0)
    pg = pg3 = 0.0
    steps = 0
    t0 = time.perf_counter()
    while True:
        pg, pg3 = o.getup(U, P, HH, S, pg, pg3, steps + 1)
        o.divcheck(U)
        o.cfl(U)
        steps += 1
        el = time.perf_counter() - t0
        if el >= target_s or steps >= 50:
            break
    pts = (cfg["n1"] - 1) * (cfg["n2"] - 1) * (cfg["n3"] - 1)
    return dict(value=pts * steps / el / 1e6, unit="Mpts·steps/s", cores=threads, kind="port",
                sample=f"{steps} steps (GETUP+DIVCHECK+CFL) of the same {cfg['n1']-1}x{cfg['n2']}x{cfg['n3']-1} "
                       f"FP64 workload by the C oracle (bit-exact restatement of the reference's "
                       f"subroutines), OpenMP {threads} threads, {el:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", default="256x193x256", help="N1M x N2 x N3M")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-timing", action="store_true", help="no per-kernel events (for profilers)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU decomposition is a design "
                                   "only (DESIGN.md, 'Sharding').  Use --gpus 1."}))
        sys.exit(2)

    import numpy as np
    import torch

    from compact3d_amd import c3d
    from compact3d_amd.initfield import perturbed_poiseuille

    grid = tuple(int(x) for x in args.grid.lower().split("x"))
    cfg = workload(grid)
    pts = grid[0] * (grid[1] - 1) * grid[2]
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    torch.cuda.init()
    ch = c3d.Channel(**cfg)
    m = ch.mesh()  # the grid the library built, for the initial field
    U, P = perturbed_poiseuille(cfg["n1"], cfg["n2"], cfg["n3"], m["Y"], m["DX1"], m["DX3"], m["ALY"], seed=args.seed)
    ch.upload(U, P)
    del U, P
    for n in range(args.warmup):
        ch.step(n + 1, checks=False)
    torch.cuda.synchronize()
    nlaunch = 40 * args.steps + 64
    if not args.no_timing:
        ch.timing(True, nlaunch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for n in range(args.steps):
        ch._ck(ch.L.c3d_step(ch.h, args.warmup + n + 1, None, None), "step")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kt = ch.kernel_times() if not args.no_timing else {}
    ch.timing(False)
    dt_s = (t1 - t0) / args.steps
    value = pts / dt_s / 1e6
    divmax = ch.divcheck()
    cflm = ch.cfl()
    st_presg = ch.scalar(c3d.S_PRESG)
    finite = math.isfinite(divmax) and math.isfinite(cflm)

    roof = None
    kernels = {}
    if kt:
        for name, (ms, cnt) in kt.items():
            kernels[name] = dict(avg_us=1000.0 * ms / cnt, launches=cnt, share=ms / (1000.0 * dt_s * args.steps))
        dom = max((k for k in kt if k in ALG_BYTES), key=lambda k: kt[k][0])
        avg_s = kt[dom][0] / kt[dom][1] / 1000.0
        alg = ALG_BYTES[dom] * pts
        achieved = alg / avg_s / 1e9
        traffic = None
        tp = os.path.join(ROOT, "profiles", "traffic.json")
        if os.path.exists(tp):
            try:
                tj = json.load(open(tp))
                rec = tj.get(f"{dom}@{args.grid}")
                if rec:
                    traffic = rec["bytes_per_launch"]
            except Exception:
                traffic = None
        roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, kernel=dom,
                    alg_bytes_per_launch=alg, avg_launch_us=round(avg_s * 1e6, 2))
    step_frac = 112.0 * pts / dt_s / (HBM_PEAK_GBS * 1e9)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, args.cpu_seconds, args.seed)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpts·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt_s * 1000.0, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded divergence-free perturbed Poiseuille field (seed %d)" % args.seed,
        "config": {"workload": "channel DNS step (GETUP+DIVCHECK+CFL), %s, Re_tau=395 (RE=10312.5), "
                               "dt=7.5e-4, FP64 path" % args.grid,
                   "grid": args.grid, "re": cfg["re"], "dt": cfg["dt"], "cks": cfg["cks"],
                   "precision": "fp64", "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "step_alg_frac": round(step_frac, 4),
        "divmax_after": divmax,
        "cfl_after": cflm * cfg["dt"],
        "presg_after": st_presg,
        "finite": finite,
        "kernels": kernels,
    }
    print(json.dumps(out))
    ch.close()


if __name__ == "__main__":
    main()
