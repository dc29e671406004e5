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
the largest share of the step (found in a short all-kernel timing pass after
the warm-up): HIP events recorded by the library on its own stream around
each of that kernel's launches inside the timed steps give its average
duration.  `cpu_baseline` times the FP64 oracle (oracle/liboracle.so,
OpenMP) on a few steps of the same grid.
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
    # fused in (per component: state, VV1/VV3 or P, RUH, HH, HH2 in; HH, UH out);
    # the spectral kernels move (NX/2+1)/NX ~ 1/2 a complex per cell, pois_y
    # also its two per-mode factor tables, stored for kz <= NZ/2 (1/4 of a
    # double pair per cell)
    "rhs_x": 88, "rhs_z": 144, "rhs_x2": 24, "adi_y": 168, "adi_z": 64,
    "rhsdp_x": 24, "rhsdp_z": 24, "fft_x": 16, "fft_z": 16, "pois_y": 20, "fft_xinv": 16,
    "upd_x": 48, "upd_z": 64,
}
# adi_x runs twice per step: (u,v) read+write 32 B, then w with U added 24 B
ALG_BYTES_ADI_X_PER_STEP = 56
# mixed mode: FP32 intermediates and state shadows, FP64 state, RDP, DP, DIV
# (state read as FP32 shadows by the line kernels; the update reads and
# writes the FP64 state and writes the shadows).  The x and z ADI sweeps run
# z (u, v: DELTA in and out), then x (u, v, w: DELTA in and out; u's also
# writes RDP = D1x(DELTA_u)/DT in FP64), then z for w, which also reads RDP,
# DELTA_v and the DIV of the state (carried in FP32) and writes the final
# RDP: RHSDP has no kernel of its own
ALG_BYTES_MIXED = {
    "rhs_x": 44, "rhs_z": 72, "rhs_x2": 12, "adi_y": 84, "adi_z": 16,
    "adi_xr": 32, "adi_zr": 32, "fft_x": 16, "fft_z": 16, "pois_y": 20, "fft_xinv": 16,
    # the update: U and V on x-lines (DP, DELTA_u, U, DELTA_v, V in; U, V,
    # their shadows, D1x U out), then W and P on z-lines (DP, DELTA_w, W, P,
    # D1x U, U, V in; W, P, their shadows, the FP32 DIV out) with the y and z
    # parts of DIV, DIVMAX and CFL
    "upd_x": 64, "upd_z": 80,
}


def workload(grid):
    n1m, n2, n3m = grid
    return dict(n1=n1m + 1, n2=n2, n3=n3m + 1, re=10312.5, alx=2 * math.pi, aly=2.0, alz=math.pi,
                dt=7.5e-4, files=0.0, wmodel=0.0, cks=1.8)


def oracle_run(cfg, seed, target_s, max_steps):
    """the FP64 oracle (OpenMP on the host cores) stepping the seeded field of
    `cfg` (GETUP + DIVCHECK + CFL per step, as MAIN) until max_steps or
    target_s seconds; returns (steps, seconds, threads, U, P)"""
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
        if el >= target_s or steps >= max_steps:
            break
    return steps, el, threads, U, P


def cpu_baseline(cfg, target_s, seed):
    """FP64 oracle on the host cores: a bounded sample of the same workload.
    Returns the baseline record and the oracle's field after its steps."""
    steps, el, threads, U, P = oracle_run(cfg, seed, target_s, 50)
    pts = (cfg["n1"] - 1) * (cfg["n2"] - 1) * (cfg["n3"] - 1)
    rec = dict(value=pts * steps / el / 1e6, unit="Mpts·steps/s", cores=threads, kind="port",
               sample=f"{steps} steps (GETUP+DIVCHECK+CFL) of the same {cfg['n1']-1}x{cfg['n2']}x{cfg['n3']-1} "
                      f"FP64 workload by the C oracle (bit-exact restatement of the reference's "
                      f"subroutines), OpenMP {threads} threads, {el:.1f} s")
    return rec, steps, U, P


def calibration(reps=20):
    """the oracle against the reference's own subroutines (oracle/_ref, built
    from compact3d.f) on the reference's PARAM.H grid 32x33x32, one thread
    each (the reference library is built without OpenMP): seconds per pass
    of BCOND, RHS1-3, GETUH1-3, RHSDP, UPCALC, DIVCHECK and CFL (GETDP needs
    FFTW, which the reference build lacks; both sides take the oracle's DP),
    and their ratio.  None when the reference library was not built."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from pyoracle import REF_SO, Oracle, RefLib
    if not os.path.exists(REF_SO):
        return None
    from cases import inputs
    cfg = dict(re=4200.0, alx=2 * math.pi, aly=2.0, alz=math.pi, dt=1e-3, files=0.0, wmodel=0.0)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    Oracle.set_threads(1)
    try:
        ref = RefLib(**cfg)
        o = Oracle(ref.n1, ref.n2, ref.n3, cks=2.5, **cfg)
        U0, P0, HH2 = inputs(o, seed=7)
        S = o.zeros_u(0)
        DP = o.zeros_p()
        RDP0 = o.zeros_p()
        UH0 = o.zeros_u()
        RUH0 = [o.zeros_u(0) for _ in range(3)]
        HH0 = o.zeros_u()
        for c in (1, 2, 3):
            o.rhs(c, U0, P0, HH0, HH2, RUH0[c - 1], 0.003, S)
            o.getuh(c, U0, UH0, RUH0[c - 1])
        o.rhsdp(RDP0, UH0)
        o.getdp(DP, RDP0)

        def one(lib, is_ref):
            U, P = U0.copy(order="F"), P0.copy(order="F")
            HH, UH, RDP = o.zeros_u(), o.zeros_u(), o.zeros_p()
            RUH = [o.zeros_u(0) for _ in range(3)]
            t0 = time.perf_counter()
            for _ in range(reps):
                lib.bcond(U, S)
                for c in (1, 2, 3):
                    lib.rhs(c, U, P, HH, HH2, RUH[c - 1], 0.003, S)
                for c in (1, 2, 3):
                    if is_ref:
                        lib.getuh(c, U, UH, RUH[c - 1], S)
                    else:
                        lib.getuh(c, U, UH, RUH[c - 1])
                if is_ref:
                    lib.rhsdp(RDP, UH, U)
                else:
                    lib.rhsdp(RDP, UH)
                lib.upcalc(U, P, UH, DP, 0.003, -0.001)
                lib.divcheck(U)
                lib.cfl(U)
            return (time.perf_counter() - t0) / reps

        one(ref, True)
        one(o, False)
        t_ref, t_orc = one(ref, True), one(o, False)
    finally:
        Oracle.set_threads(threads)
    return dict(grid="32x33x32", threads=1, ref_s_per_pass=round(t_ref, 6), oracle_s_per_pass=round(t_orc, 6),
                ref_over_oracle=round(t_ref / t_orc, 3),
                phases="BCOND RHS1-3 GETUH1-3 RHSDP UPCALC DIVCHECK CFL (GETDP excluded: FFTW absent)")


def field_errors(ref_U, ref_P, got_U, got_P, n1, n2, n3):
    """per-field max and L2 errors relative to the reference field, and the
    mean-velocity (bulk u) error"""
    import numpy as np
    out = {}
    for name, comp, j0 in (("u", 0, 1), ("v", 1, 2), ("w", 2, 1)):
        a_ = ref_U[1:n1, j0:n2, 1:n3, comp]
        b_ = got_U[1:n1, j0:n2, 1:n3, comp]
        out[name] = dict(max=float(np.abs(a_ - b_).max() / np.abs(a_).max()),
                         l2=float(np.linalg.norm(a_ - b_) / np.linalg.norm(a_)))
    a_, b_ = ref_P[: n1 - 1, : n2 - 1, : n3 - 1], got_P[: n1 - 1, : n2 - 1, : n3 - 1]
    out["p"] = dict(max=float(np.abs(a_ - b_).max() / np.abs(a_).max()),
                    l2=float(np.linalg.norm(a_ - b_) / np.linalg.norm(a_)))
    ua = ref_U[1:n1, 1:n2, 1:n3, 0].mean(axis=(0, 2))
    ub = got_U[1:n1, 1:n2, 1:n3, 0].mean(axis=(0, 2))
    out["mean_u_profile"] = float(np.abs(ua - ub).max() / np.abs(ua).max())
    return out


def accuracy_leg(c3d, cfg, U, P, nsteps):
    """nsteps of the mixed and the FP64 path from the same field, DIVMAX read
    back after every step: the largest DIVMAX of each and the mixed path's
    per-field errors against the FP64 path at the end"""
    chs = [c3d.Channel(**cfg, precision=p) for p in (c3d.FP64, c3d.MIXED)]
    for ch in chs:
        ch.upload(U, P)
    dmax = [0.0, 0.0]
    t0 = time.perf_counter()
    for n in range(1, nsteps + 1):
        for q, ch in enumerate(chs):
            d, _ = ch.step(n)
            dmax[q] = max(dmax[q], d) if d == d else float("nan")
    el = time.perf_counter() - t0
    s64, smx = chs[0].download(), chs[1].download()
    errs = field_errors(s64["U"], s64["P"], smx["U"], smx["P"], cfg["n1"], cfg["n2"], cfg["n3"])
    for ch in chs:
        ch.close()
    return dict(steps=nsteps, divmax_max_fp64=dmax[0], divmax_max_mixed=dmax[1],
                mixed_divmax_le_1e9_every_step=bool(dmax[1] <= 1e-9), mixed_vs_fp64=errs,
                mean_u_le_1e5=bool(errs["mean_u_profile"] <= 1e-5), seconds=round(el, 2))


def configs_leg(c3d, args):
    """The two smaller benchmark channels on the GPU (FP64 and mixed) and
    the oracle on the host cores: 64x65x64 (Re_tau 180, tanh 1.5, dt 1e-3,
    200 steps; the oracle runs all of them) and 192x129x192 (500 steps; the
    oracle a bounded sample).  The GPU steps run as c3d_step calls and again
    as captured step graphs replayed one step per call (graph_*), both with
    DIVMAX and CFL read back every step."""
    from compact3d_amd.initfield import perturbed_poiseuille
    out = {}
    for name, grid, steps, cpu_s, seed in (("c0", (64, 65, 64), 200, 60.0, 0), ("c1", (192, 129, 192), 500, 10.0, 1)):
        n1m, n2, n3m = grid
        cfg = dict(n1=n1m + 1, n2=n2, n3=n3m + 1, re=4200.0, alx=2 * math.pi, aly=2.0, alz=math.pi, dt=1e-3,
                   files=0.0, wmodel=0.0, cks=1.5)
        pts = n1m * (n2 - 1) * n3m
        rec = {"grid": "%dx%dx%d" % grid, "steps": steps}
        probe = c3d.Channel(**cfg)
        m = probe.mesh()
        probe.close()
        U, P = perturbed_poiseuille(cfg["n1"], cfg["n2"], cfg["n3"], m["Y"], m["DX1"], m["DX3"], m["ALY"], seed=seed)
        for pname, prec in (("fp64", c3d.FP64), ("mixed", c3d.MIXED)):
            ch = c3d.Channel(**cfg, precision=prec)
            ch.upload(U, P)
            for n in range(1, 4):
                ch.step(n)
            t0 = time.perf_counter()
            dm = 0.0
            for n in range(4, 4 + steps):
                d, _ = ch.step(n)
                dm = max(dm, d)
            el = time.perf_counter() - t0
            # the same steps replayed as the captured step graph, one step
            # per call with its DIVMAX and CFL read back (c3d_advance(1) +
            # c3d_sync): one launch per step instead of the step's kernels
            ch.advance(2, use_graph=True)
            ch.sync()
            t0 = time.perf_counter()
            dmg = 0.0
            for _ in range(steps):
                ch.advance(1, use_graph=True)
                d, _ = ch.sync()
                dmg = max(dmg, d)
            elg = time.perf_counter() - t0
            ch.close()
            rec[pname] = dict(mpts_steps_s=round(pts * steps / el / 1e6, 1), ms_per_step=round(1000 * el / steps, 4),
                              divmax_max=dm, graph_mpts_steps_s=round(pts * steps / elg / 1e6, 1),
                              graph_ms_per_step=round(1000 * elg / steps, 4), graph_divmax_max=dmg)
        if not args.no_cpu:
            st, el, thr, _, _ = oracle_run(cfg, seed, cpu_s, steps)
            rec["cpu_oracle"] = dict(mpts_steps_s=round(pts * st / el / 1e6, 2), steps=st, seconds=round(el, 1),
                                     cores=thr)
        out[name] = rec
    return out


def operator_leg(c3d):
    """isolated operator benches on a seeded uniform[-1,1] FP64 field of
    192x129x192: the compact operators D2, D1 and I (both staggerings) along
    x and z (one periodic tridiagonal solve per line, c3d_apply_op's
    stand-alone kernel), the Poisson solve (RHSDP + GETDP) and DIVCHECK/CFL,
    kernel times from HIP events; GB/s counts one read and one write of the
    field per operator"""
    import numpy as np
    cfg = dict(n1=193, n2=129, n3=193, re=4200.0, alx=2 * math.pi, aly=2.0, alz=math.pi, dt=1e-3,
               files=0.0, wmodel=0.0, cks=1.5)
    n1, n2, n3 = cfg["n1"], cfg["n2"], cfg["n3"]
    pts = (n1 - 1) * (n2 - 1) * (n3 - 1)
    rng = np.random.default_rng(1)
    f = np.zeros((n1, n2, n3), order="F")
    f[: n1 - 1, : n2 - 1, : n3 - 1] = rng.uniform(-1, 1, (n1 - 1, n2 - 1, n3 - 1))
    ch = c3d.Channel(**cfg)
    U = np.zeros((n1 + 1, n2 + 1, n3 + 1, 3), order="F")
    U[1:n1, 0:n2 + 1, 1:n3, :] = rng.uniform(-1, 1, (n1 - 1, n2 + 1, n3 - 1, 3))
    ch.upload(U, f)
    names = {0: "D2", 1: "D1 centre->face", 2: "D1 face->centre", 3: "I centre->face", 4: "I face->centre"}
    out = {}
    for direction, dname in ((1, "x"), (3, "z")):
        for op in range(5):
            ch.apply_op(op, direction, f)
            ch.timing(True, 64)
            for _ in range(5):
                ch.apply_op(op, direction, f)
            kt = ch.kernel_times()
            ch.timing(False)
            ms, cnt = kt["apply_op"]
            us = 1000.0 * ms / cnt
            out["%s %s" % (names[op], dname)] = dict(avg_us=round(us, 2), gbs=round(16.0 * pts / us / 1e3, 1))
    ch.getuh()
    ch.timing(True, 256)
    for _ in range(5):
        ch.dpcalc()
        ch.divcheck()
        ch.cfl()
    kt = ch.kernel_times()
    ch.timing(False)
    ch.close()
    for k, (ms, cnt) in kt.items():
        out[k] = dict(avg_us=round(1000.0 * ms / cnt, 2), launches=cnt)
    return dict(grid="192x129x192", field="uniform[-1,1] FP64, seed 1", kernels=out)



def run_leg(c3d, cfg, U, P, precision, args, pts, table):
    """warm-up, a short pass with every kernel timed (the per-kernel table
    and the dominant kernel), then args.steps timed steps (GETUP + DIVCHECK +
    CFL read back every step, as MAIN's loop does) of one precision, with
    HIP events around the dominant kernel's launches only"""
    import torch
    ch = c3d.Channel(**cfg, precision=precision)
    ch.upload(U, P)
    n = 0
    for _ in range(args.warmup):
        n += 1
        ch.step(n)
    torch.cuda.synchronize()
    kt_all = {}
    if not args.no_timing:
        nprof = max(5, args.steps // 5)
        ch.timing(True, 48 * nprof + 64)
        for _ in range(nprof):
            n += 1
            ch.step(n)
        kt_all = ch.kernel_times()
        ch.timing(False)
    dom = max((k for k in kt_all if k in table), key=lambda k: kt_all[k][0]) if kt_all else None
    if dom is not None:
        ch.timing_select(dom)
        ch.timing(True, 4 * args.steps + 64)
    torch.cuda.synchronize()
    dmax = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        n += 1
        d, _ = ch.step(n)
        dmax = max(dmax, d) if d == d else float("nan")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kt_dom = ch.kernel_times() if dom is not None else {}
    ch.timing(False)
    ch.timing_select(None)
    dt_s = (t1 - t0) / args.steps
    tot_ms = sum(ms for ms, _ in kt_all.values())
    kernels = {}
    for name, (ms, cnt) in kt_all.items():
        kernels[name] = dict(avg_us=round(1000.0 * ms / cnt, 2), launches=cnt, share=round(ms / tot_ms, 4))
    return ch, dict(value=pts / dt_s / 1e6, dt_s=dt_s, divmax_max=dmax, kt=kt_dom, kernels=kernels,
                    steps_run=n)


def roofline(kt, table, pts, grid, precision):
    if not kt:
        return None
    dom = max((k for k in kt if k in table), key=lambda k: kt[k][0])
    avg_s = kt[dom][0] / kt[dom][1] / 1000.0
    alg = table[dom] * pts
    achieved = alg / avg_s / 1e9
    traffic = None
    tp = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tp):
        try:
            rec = json.load(open(tp)).get(f"{dom}@{grid}@{precision}")
            if rec:
                traffic = rec["bytes_per_launch"]
        except Exception:
            traffic = None
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, kernel=dom,
                alg_bytes_per_launch=alg, avg_launch_us=round(avg_s * 1e6, 2))


def served_leg(args):
    """The serving metric: a seeded Poisson stream of channel requests served
    by the continuous-batching loop (compact3d_amd.serve), and the same
    requests all present at t=0 (the batched throughput it is compared to).
    100-2000 steps per request; the offered load stays above what the GPU
    completes, so the latencies include queueing."""
    import copy
    from compact3d_amd import c3d
    from compact3d_amd.serve import FieldCache, Server, poisson_requests, stats
    grids = [(64, 49, 64), (96, 65, 96), (128, 65, 128), (192, 129, 192), (256, 129, 256)]
    reqs = poisson_requests(args.serve_requests, 20.0, grids, [180, 395, 1000], (100, 2000), seed=4)
    if not args.no_autotune:  # tile shapes of every configuration in the stream, before any timing
        seen = set()
        for r in reqs:
            key = (r.grid, r.precision, r.les)
            if key not in seen:
                seen.add(key)
                c3d.autotune(steps=3, **r.config(), precision=r.precision)
    fields = FieldCache()
    for r in reqs:  # seeded initial fields, generated before any timing
        ch = c3d.Channel(**r.config(), precision=r.precision)
        fields.get(r, ch.mesh())
        ch.close()
    budget = 240 << 30
    out = {}
    for name, stream in (("batched", [copy.copy(r) for r in reqs]), ("served", reqs)):
        if name == "batched":
            for r in stream:
                r.t_arrive = 0.0
        srv = Server(budget, max_live=args.serve_slots, chunk=10, use_graph=True, fields=fields,
                     policy=args.serve_policy, max_active=args.serve_active)
        done = srv.run(stream)
        wall = max(r.t_done for r in done) - min(r.t_arrive for r in done)
        st = stats(done, wall)
        st["peak_resident_gb"] = srv.peak_bytes / 1e9
        st["divmax_max"] = max(r.result["divmax"] for r in done)
        st["failed"] = len(srv.failed)
        out[name] = st
    out["served_over_batched"] = out["served"]["throughput_mpts_steps_s"] / out["batched"]["throughput_mpts_steps_s"]
    out["stream"] = ("%d requests, Poisson arrivals 20/s (seed 4), grids 64x49x64..256x129x256, Re_tau "
                     "{180, 395; 1000 wall-modelled LES}, all mixed precision, 100-2000 steps, dt 5e-4, "
                     "%d slots, chunks of up to 10 steps, captured graphs on each instance's own stream, "
                     "policy %s" % (len(reqs), args.serve_slots, args.serve_policy))
    out["policy"] = args.serve_policy
    if args.serve_policy == "srpt":
        out["max_active"] = args.serve_active
    return out


def ensemble_leg(args):
    """Batched uneven instances: wall-modelled Smagorinsky LES channels
    (Cs 0.1 with van Driest damping, log-law wall stress by Newton iteration
    at both walls, WMODEL=6) with grids drawn by seed 3 from
    {48x33x48, 64x49x64, 96x65x96, 128x65x128} and Re_tau from
    {1000, 2000, 5200}, mixed precision, advanced args.ens_steps steps by one
    captured-graph launch per step for all of them (c3d_ensemble), then the
    same instances on their own streams (c3d_advance) for comparison."""
    import numpy as np
    from compact3d_amd import c3d
    from compact3d_amd.serve import FieldCache, Request, re_from_re_tau
    grids = [(48, 33, 48), (64, 49, 64), (96, 65, 96), (128, 65, 128)]
    re_taus = [1000, 2000, 5200]
    rng = np.random.default_rng(3)
    reqs = []
    for q in range(args.ens_members):
        g = grids[int(rng.integers(len(grids)))]
        rt = re_taus[int(rng.integers(len(re_taus)))]
        reqs.append(Request(rid=q, grid=g, re=re_from_re_tau(rt), steps=args.ens_steps, seed=q % 3,
                            precision=c3d.MIXED, les=True, dt=5e-4))
    tuned = {}
    if not args.no_autotune:  # one shape set per distinct configuration, before any channel exists
        for r in reqs:
            key = "%dx%dx%d" % r.grid
            if key not in tuned:
                t = c3d.autotune(steps=4, **r.config(), precision=r.precision)
                tuned[key] = dict(ms_rule=round(t["ms_rule"], 4), ms_tuned=round(t["ms_tuned"], 4))
    fields = FieldCache()
    chans = []
    for r in reqs:
        ch = c3d.Channel(**r.config(), precision=r.precision)
        U, P = fields.get(r, ch.mesh())
        ch.upload(U, P)
        chans.append(ch)
    pts = sum(r.pts for r in reqs)
    n = args.ens_steps
    ens = c3d.Ensemble(chans)
    ens.advance(2)  # capture + warm-up
    ens.sync()
    t0 = time.perf_counter()
    ens.advance(n)
    dm, cf = ens.sync()
    t1 = time.perf_counter()
    ens.close()
    thr_ens = pts * n / (t1 - t0) / 1e6
    for ch in chans:
        ch.advance(2, True)
    for ch in chans:
        ch.sync()
    t2 = time.perf_counter()
    for ch in chans:
        ch.advance(n, True)
    dm2 = [ch.sync()[0] for ch in chans]
    t3 = time.perf_counter()
    thr_str = pts * n / (t3 - t2) / 1e6
    for ch in chans:
        ch.close()
    counts = {}
    for r in reqs:
        key = "%dx%dx%d" % r.grid
        counts[key] = counts.get(key, 0) + 1
    return dict(members=len(reqs), grids=counts, steps=n, precision="mixed", model="Smagorinsky + log-law wall model at both walls (WMODEL=6)",
                autotune=tuned or None,
                pts_per_step=pts, ensemble_ms_per_step=round(1000.0 * (t1 - t0) / n, 4),
                ensemble_mpts_steps_s=round(thr_ens, 1), streams_mpts_steps_s=round(thr_str, 1),
                divmax_max=float(max(max(dm), max(dm2))), finite=bool(np.all(np.isfinite(dm)) and np.all(np.isfinite(cf))))


def les_leg(c3d, args, given=None):
    """the wall-modelled LES channel (Smagorinsky SGSCS, the log law at both
    walls, WMODEL=6) of the served stream's Re_tau 1000 requests on the
    headline grid 256x193x256 (dt 5e-4), FP64 and mixed: ms per step and
    Mpts.steps/s over args.les_steps steps of c3d_step (DIVMAX and CFL read
    back every step) and the per-kernel table of a short event-timed pass.
    Tile shapes from `given` (the --shapes file's les_fp64 / les_mixed sets)
    or the autotuner."""
    import torch
    from compact3d_amd.initfield import perturbed_poiseuille
    from compact3d_amd.serve import re_from_re_tau
    cfg = dict(n1=257, n2=193, n3=257, re=re_from_re_tau(1000), alx=2 * math.pi, aly=2.0, alz=math.pi,
               dt=5e-4, files=2.0, wmodel=6.0, cks=1.8)
    pts = 256 * 192 * 256
    out = dict(grid="256x193x256", re_tau=1000, model="Smagorinsky + log-law wall model at both walls (WMODEL=6)")
    for name, prec in (("fp64", c3d.FP64), ("mixed", c3d.MIXED)):
        key = "les_" + name
        if given and key in given:
            c3d.autotune_set(given[key], **cfg, precision=prec)
            shapes = {k: given[key].get(k, c3d.SHAPE_DEFAULTS.get(k)) for k in c3d.SHAPE_KEYS}
        elif not args.no_autotune:
            t = c3d.autotune(steps=3, **cfg, precision=prec)
            shapes = {k: t[k] for k in c3d.SHAPE_KEYS}
        else:
            shapes = None
        ch = c3d.Channel(**cfg, precision=prec)
        m = ch.mesh()
        U, P = perturbed_poiseuille(cfg["n1"], cfg["n2"], cfg["n3"], m["Y"], m["DX1"], m["DX3"], m["ALY"], seed=2)
        ch.upload(U, P)
        del U, P
        n = 0
        for _ in range(3):
            n += 1
            ch.step(n)
        kernels = {}
        if not args.no_timing:
            ch.timing(True, 4096)
            for _ in range(5):
                n += 1
                ch.step(n)
            for k, (ms, cnt) in ch.kernel_times().items():
                kernels[k] = dict(avg_us=round(1000.0 * ms / cnt, 2), launches=cnt)
            ch.timing(False)
        torch.cuda.synchronize()
        dm = 0.0
        t0 = time.perf_counter()
        for _ in range(args.les_steps):
            n += 1
            d, _ = ch.step(n)
            dm = max(dm, d) if d == d else float("nan")
        torch.cuda.synchronize()
        dt_s = (time.perf_counter() - t0) / args.les_steps
        ch.close()
        out[name] = dict(mpts_steps_s=round(pts / dt_s / 1e6, 1), ms_per_step=round(1000.0 * dt_s, 4),
                         divmax_max=dm, kernels=kernels, shapes=shapes)
    out["mixed_over_fp64"] = round(out["mixed"]["mpts_steps_s"] / out["fp64"]["mpts_steps_s"], 3)
    return out


# kernel base names of the step -> bench keys
KERNEL_KEYS = {"k_rhsy_adiy": "adi_y", "k_rhs_x": "rhs_x", "k_rhs_x2": "rhs_x2", "k_rhs_z": "rhs_z",
               "k_adi_x": "adi_x", "k_adi_z": "adi_z", "k_rhsdp_x": "rhsdp_x",
               "k_rhsdp_z": "rhsdp_z", "k_fft_x": "fft_x", "k_fft_z": "fft_z", "k_pois_y": "pois_y",
               "k_fft_xinv": "fft_xinv", "k_rfft_x": "fft_x", "k_rfft_z": "fft_z", "k_rfft_xinv": "fft_xinv",
               "k_upd_x": "upd_x", "k_upd_xu": "upd_x", "k_upd_v": "upd_v", "k_upd_z": "upd_z",
               "k_adi_xr": "adi_xr", "k_adi_zr": "adi_zr"}
# adi_x runs twice per FP64 step; its per-launch algorithmic bytes are the
# mean (the mixed step has no adi_x: its x-sweeps are adi_xr)
ADI_X_ALG = {"fp64": ALG_BYTES_ADI_X_PER_STEP / 2.0}


def summarize_counters(rd_csv, wr_csv, grid, out_json):
    """HBM bytes per launch of every step kernel from two rocprofv3 --pmc
    runs of `bench.py --copy-calibration --no-autotune --no-cpu --no-serve
    --no-ensemble` (reads: TCC_EA0_RDREQ_{32,64,128}B request counts; writes:
    WRITE_SIZE in KiB), scaled by what the same counters read for the plain
    256 MiB copy that --copy-calibration runs first.  Dispatches before the
    first FP32 kernel belong to the FP64 leg, the rest to the mixed leg."""
    import csv
    import re

    def load(path):
        rows = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                d = int(r["Dispatch_Id"])
                rec = rows.setdefault(d, {"name": r["Kernel_Name"]})
                rec[r["Counter_Name"]] = rec.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        return rows

    rd, wr = load(rd_csv), load(wr_csv)

    def rbytes(r):
        return sum(float(r.get("TCC_EA0_RDREQ_%dB_sum" % b, r.get("TCC_EA0_RDREQ_%dB" % b, 0.0))) * b
                   for b in (32, 64, 128))

    def wbytes(r):
        return float(r.get("WRITE_SIZE", 0.0)) * 1024.0

    copy_bytes = 256 * 1024 * 1024
    cr = [d for d in sorted(rd) if "copyBuffer" in rd[d]["name"] or "copy" in rd[d]["name"].lower()]
    cw = [d for d in sorted(wr) if "copyBuffer" in wr[d]["name"] or "copy" in wr[d]["name"].lower()]
    rcopy = max(rbytes(rd[d]) for d in cr)
    wcopy = max(wbytes(wr[d]) for d in cw)
    rf, wf = copy_bytes / rcopy, copy_bytes / wcopy

    def legs(rows):
        first32 = min((d for d in rows if "<float" in rows[d]["name"]), default=None)
        out = {}
        for d, r in rows.items():
            m = re.search(r"(k_\w+)", r["name"])
            if not m or m.group(1) not in KERNEL_KEYS:
                continue
            leg = "fp64" if (first32 is None or d < first32) else "mixed"
            out.setdefault((KERNEL_KEYS[m.group(1)], leg), []).append(r)
        return out

    lr, lw = legs(rd), legs(wr)
    n1m, n2, n3m = grid
    pts = n1m * (n2 - 1) * n3m
    res = {"_calibration": {"copy_bytes": copy_bytes, "read_bytes_of_copy": rcopy, "write_counter_bytes_of_copy": wcopy,
                            "read_factor": round(rf, 4), "write_factor": round(wf, 4)}}
    gname = "%dx%dx%d" % grid
    for key in sorted(lr):
        if key not in lw:
            continue
        k, leg = key
        r = rf * sum(rbytes(x) for x in lr[key]) / len(lr[key])
        w = wf * sum(wbytes(x) for x in lw[key]) / len(lw[key])
        table = ALG_BYTES if leg == "fp64" else ALG_BYTES_MIXED
        alg = (ADI_X_ALG[leg] if k == "adi_x" else table[k]) * pts
        res["%s@%s@%s" % (k, gname, leg)] = {"bytes_per_launch": int(r + w), "fetch": int(r), "write": int(w),
                                              "alg_bytes": int(alg), "traffic_over_alg": round((r + w) / alg, 3)}
    with open(out_json, "w") as f:
        json.dump(res, f, indent=1)
    return res


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
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--serve-policy", default="srpt", choices=["srpt", "fifo"],
                    help="serving loop policy: shortest remaining work first, or arrival order")
    ap.add_argument("--serve-active", type=int, default=6,
                    help="srpt: instances with a chunk queued at a time (6: the lowest p50 and mean "
                         "latency of 4..12 at full throughput on the 24-request stream)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched LES ensemble leg")
    ap.add_argument("--no-autotune", action="store_true", help="keep the rule-derived tile shapes")
    ap.add_argument("--shapes", metavar="JSON",
                    help="tile shapes of the main legs from a file ({'fp64': {...}, 'mixed': {...}}, the 'autotune' "
                         "field of an earlier run) instead of autotuning them; for profiling the same kernels")
    ap.add_argument("--acc-steps", type=int, default=1000,
                    help="steps of the mixed-vs-FP64 accuracy leg (0 skips it)")
    ap.add_argument("--no-configs", action="store_true", help="skip the 64x65x64 and 192x129x192 legs")
    ap.add_argument("--no-operators", action="store_true", help="skip the isolated operator benches")
    ap.add_argument("--les-steps", type=int, default=20,
                    help="steps of the 256x193x256 wall-modelled LES leg per precision (0 skips it)")
    ap.add_argument("--ens-members", type=int, default=48)
    ap.add_argument("--ens-steps", type=int, default=200)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain device copy of 256 MiB three times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    ap.add_argument("--summarize-counters", nargs=2, metavar=("READ_CSV", "WRITE_CSV"),
                    help="write profiles/traffic.json from two rocprofv3 --pmc counter CSVs and exit")
    args = ap.parse_args()
    if args.summarize_counters:
        grid = tuple(int(x) for x in args.grid.lower().split("x"))
        res = summarize_counters(args.summarize_counters[0], args.summarize_counters[1], grid,
                                 os.path.join(ROOT, "profiles", "traffic.json"))
        for k, v in res.items():
            if not k.startswith("_"):
                print("%-28s %8.1f MB  (alg %8.1f MB, %.3f)" % (k, v["bytes_per_launch"] / 1e6, v["alg_bytes"] / 1e6,
                                                              v["traffic_over_alg"]))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU; the multi-GPU decomposition is a design "
                                   "only.  Use --gpus 1."}))
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
    if args.copy_calibration:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    tune = {}
    given = None
    if args.shapes:
        with open(args.shapes) as f:
            given = json.load(f)
        for name, prec in (("fp64", c3d.FP64), ("mixed", c3d.MIXED)):
            c3d.autotune_set(given[name], **cfg, precision=prec)
            tune[name] = {k: given[name].get(k, c3d.SHAPE_DEFAULTS.get(k)) for k in c3d.SHAPE_KEYS}
            tune[name]["from"] = args.shapes
    elif not args.no_autotune:
        for name, prec in (("fp64", c3d.FP64), ("mixed", c3d.MIXED)):
            t = c3d.autotune(steps=4, **cfg, precision=prec)
            tune[name] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in t.items()}
    probe = c3d.Channel(**cfg)
    m = probe.mesh()  # the grid the library built, for the initial field
    probe.close()
    U, P = perturbed_poiseuille(cfg["n1"], cfg["n2"], cfg["n3"], m["Y"], m["DX1"], m["DX3"], m["ALY"], seed=args.seed)
    ch64, r64 = run_leg(c3d, cfg, U, P, c3d.FP64, args, pts, ALG_BYTES)
    chmx, rmx = run_leg(c3d, cfg, U, P, c3d.MIXED, args, pts, ALG_BYTES_MIXED)
    del U, P
    # both legs ran the same number of steps from the same field: mixed vs FP64
    s64, smx = ch64.download(), chmx.download()
    n1, n2, n3 = cfg["n1"], cfg["n2"], cfg["n3"]
    errs = {}
    for name, comp in (("u", 0), ("v", 1), ("w", 2)):
        a_ = s64["U"][1:n1, 1:n2, 1:n3, comp]
        b_ = smx["U"][1:n1, 1:n2, 1:n3, comp]
        errs[name] = dict(max=float(np.abs(a_ - b_).max() / np.abs(a_).max()),
                          l2=float(np.linalg.norm(a_ - b_) / np.linalg.norm(a_)))
    a_, b_ = s64["P"][: n1 - 1, : n2 - 1, : n3 - 1], smx["P"][: n1 - 1, : n2 - 1, : n3 - 1]
    errs["p"] = dict(max=float(np.abs(a_ - b_).max() / np.abs(a_).max()),
                     l2=float(np.linalg.norm(a_ - b_) / np.linalg.norm(a_)))
    ub64 = float(np.mean(s64["U"][1:n1, 1:n2 - 1, 1:n3, 0]))
    ubmx = float(np.mean(smx["U"][1:n1, 1:n2 - 1, 1:n3, 0]))
    errs["mean_u"] = abs(ubmx - ub64) / abs(ub64)
    del s64, smx
    refines = chmx.scalar(c3d.S_REFINES)
    cflm = chmx.cfl()
    st_presg = chmx.scalar(c3d.S_PRESG)
    ch64.close()
    chmx.close()
    value = rmx["value"]
    dt_s = rmx["dt_s"]
    finite = math.isfinite(rmx["divmax_max"]) and math.isfinite(cflm)
    roof = roofline(rmx["kt"], ALG_BYTES_MIXED, pts, args.grid, "mixed")
    roof64 = roofline(r64["kt"], ALG_BYTES, pts, args.grid, "fp64")
    step_frac = 112.0 * pts / r64["dt_s"] / (HBM_PEAK_GBS * 1e9)

    les = les_leg(c3d, args, given) if args.les_steps > 0 else None
    served = None if args.no_serve else served_leg(args)
    ensemble = None if args.no_ensemble else ensemble_leg(args)
    accuracy = None
    if args.acc_steps > 0:
        U, P = perturbed_poiseuille(cfg["n1"], cfg["n2"], cfg["n3"], m["Y"], m["DX1"], m["DX3"], m["ALY"],
                                    seed=args.seed)
        accuracy = accuracy_leg(c3d, cfg, U, P, args.acc_steps)
        del U, P
    configs = None if args.no_configs else configs_leg(c3d, args)
    operators = None if args.no_operators else operator_leg(c3d)

    cpu = None
    vs_oracle = None
    if not args.no_cpu:
        cpu, ksteps, Uo, Po = cpu_baseline(cfg, args.cpu_seconds, args.seed)
        cpu["calibration"] = calibration()
        # the GPU paths after the oracle's steps, from the same field
        U, P = perturbed_poiseuille(cfg["n1"], cfg["n2"], cfg["n3"], m["Y"], m["DX1"], m["DX3"], m["ALY"],
                                    seed=args.seed)
        vs_oracle = {"steps": ksteps}
        for pname, prec in (("fp64", c3d.FP64), ("mixed", c3d.MIXED)):
            ch = c3d.Channel(**cfg, precision=prec)
            ch.upload(U, P)
            for n in range(1, ksteps + 1):
                ch.step(n)
            st = ch.download()
            ch.close()
            vs_oracle[pname] = field_errors(Uo, Po, st["U"], st["P"], n1, n2, n3)
        del U, P, Uo, Po

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
        "dtype": "f32+f64 mixed",
        "data": "synthetic: seeded divergence-free perturbed Poiseuille field (seed %d)" % args.seed,
        "config": {"workload": "channel DNS step (GETUP+DIVCHECK+CFL, DIVMAX and CFL read back every "
                               "step), %s, Re_tau=395 (RE=10312.5), dt=7.5e-4; value = mixed path, "
                               "fp64 = FP64 path on the same field" % args.grid,
                   "grid": args.grid, "re": cfg["re"], "dt": cfg["dt"], "cks": cfg["cks"],
                   "precision": "mixed", "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp64": {"value": round(r64["value"], 2), "ms_per_step": round(r64["dt_s"] * 1000.0, 4),
                 "divmax_max": r64["divmax_max"], "roofline": roof64, "kernels": r64["kernels"],
                 "step_alg_frac": round(step_frac, 4)},
        "mixed_over_fp64": round(value / r64["value"], 3),
        "mixed_divmax_max": rmx["divmax_max"],
        "mixed_refinements": refines,
        "mixed_vs_fp64_err": errs,
        "vs_oracle": vs_oracle,
        "accuracy": accuracy,
        "configs": configs,
        "operators": operators,
        "autotune": tune or None,
        "les": les,
        "served": served,
        "ensemble": ensemble,
        "cfl_after": cflm * cfg["dt"],
        "presg_after": st_presg,
        "finite": finite,
        "kernels": rmx["kernels"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
