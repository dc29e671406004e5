"""The headline configuration: the channel 256 x 193 x 256 (2pi x 2 x pi,
tanh stretching 1.8) at Re_tau = 395 (RE = 10312.5 in the reference's
units), dt = 7.5e-4 (5e-4 in bulk-velocity units).

* one FP64 GETUP phase by phase against the oracle (<= 1e-12), with
  DIVCHECK/CFL on non-solenoidal fields, CHKMF, WALLSS and PRESG;
* 1000 steps in mixed precision against 1000 FP64 steps from the same
  seeded field: DIVMAX <= 1e-9 after every mixed step, mean-velocity
  profile within 1e-5, per-field max/L2 errors reported."""
import time

import numpy as np
import pytest

from cases import config, interior_p, interior_u

pytestmark = pytest.mark.gpu

C2 = config(257, 193, 257, re=10312.5, dt=7.5e-4, cks=1.8)


@pytest.mark.timeout(900)
def test_c2_phases_match_oracle():
    from phases import check_phases
    print("config 256x193x256 FP64 phases vs oracle", flush=True)
    check_phases(C2, seed=2)


@pytest.mark.timeout(900)
def test_c2_mixed_vs_fp64_1000_steps():
    from compact3d_amd import c3d
    from compact3d_amd.initfield import perturbed_poiseuille
    n1, n2, n3 = C2["n1"], C2["n2"], C2["n3"]
    chs = [c3d.Channel(**C2, precision=p) for p in (c3d.FP64, c3d.MIXED)]
    m = chs[0].mesh()
    U, P = perturbed_poiseuille(n1, n2, n3, m["Y"], m["DX1"], m["DX3"], m["ALY"], seed=2)
    for ch in chs:
        ch.upload(U, P)
    nsteps = 1000
    dmax = [0.0, 0.0]
    t0 = time.perf_counter()
    for n in range(1, nsteps + 1):
        for q, ch in enumerate(chs):
            d, c = ch.step(n)
            dmax[q] = max(dmax[q], d)
            assert d <= 1e-9, (q, n, d)
        if n % 100 == 0:
            print(f"step {n}: DIVMAX max FP64 {dmax[0]:.2e} mixed {dmax[1]:.2e} "
                  f"({time.perf_counter() - t0:.1f} s)", flush=True)
    sa, sb = chs[0].download(), chs[1].download()
    errs = {}
    for name, a, b in (("u", interior_u(sa["U"], n1, n2, n3, 0), interior_u(sb["U"], n1, n2, n3, 0)),
                       ("v", interior_u(sa["U"], n1, n2, n3, 1, j0=2), interior_u(sb["U"], n1, n2, n3, 1, j0=2)),
                       ("w", interior_u(sa["U"], n1, n2, n3, 2), interior_u(sb["U"], n1, n2, n3, 2)),
                       ("P", interior_p(sa["P"], n1, n2, n3), interior_p(sb["P"], n1, n2, n3))):
        d = a - b
        errs[name] = (float(np.max(np.abs(d)) / np.max(np.abs(a))),
                      float(np.sqrt(np.sum(d * d) / np.sum(a * a))))
    va, vb = chs[0].stats()[0], chs[1].stats()[0]  # mean u profile VM(J,1)
    emean = float(np.max(np.abs(va - vb)) / np.max(np.abs(va)))
    print("1000 steps mixed vs FP64: max/L2 relative",
          {k: "%.2e / %.2e" % v for k, v in errs.items()}, "mean u %.2e" % emean,
          "DIVMAX max FP64 %.2e mixed %.2e" % tuple(dmax), flush=True)
    assert emean <= 1e-5
    assert max(v[0] for k, v in errs.items() if k != "P") <= 1e-4
    for ch in chs:
        ch.close()
