"""Shared test cases: channel configurations in the reference's units and
seeded inputs.  Units follow the reference (centreline velocity 1, half
height 1, nu = 1/RE): a bulk Reynolds number Re_b(2 delta) maps to
RE = 0.75 * Re_b (compact3d.f:627, :656)."""
from __future__ import annotations

import math

import numpy as np

from compact3d_amd.initfield import perturbed_poiseuille

PI = math.pi

# grid of the reference's PARAM.H (M1=M2=M3=33): 32 x 33 x 32
REF_GRID = (33, 33, 33)


def config(n1, n2, n3, re=4200.0, alx=2 * PI, aly=2.0, alz=PI, dt=1e-3, files=0.0, wmodel=0.0, cks=2.5):
    return dict(n1=n1, n2=n2, n3=n3, re=re, alx=alx, aly=aly, alz=alz, dt=dt, files=files,
                wmodel=wmodel, cks=cks)


def inputs(orc, seed=0, p_amp=1e-2, hh_amp=1.0):
    """Perturbed Poiseuille U (divergence-free), random P and random previous
    HH (to exercise the AB2 term) for an Oracle instance."""
    m = orc.mesh()
    n1, n2, n3 = orc.n1, orc.n2, orc.n3
    U, P = perturbed_poiseuille(n1, n2, n3, m["Y"], m["DX1"], m["DX3"], m["ALY"], seed=seed)
    rng = np.random.default_rng(1000 + seed)
    P[: n1 - 1, : n2 - 1, : n3 - 1] = rng.uniform(-1, 1, (n1 - 1, n2 - 1, n3 - 1)) * p_amp
    HH = np.zeros_like(U)
    HH[1:n1, 1:n2, 1:n3, :] = rng.uniform(-1, 1, (n1 - 1, n2 - 1, n3 - 1, 3)) * hh_amp
    HH[1:n1, 1, 1:n3, 1] = 0.0
    return U, P, HH


def relerr(ref, got):
    ref = np.asarray(ref)
    got = np.asarray(got)
    scale = np.max(np.abs(ref))
    if scale == 0:
        return float(np.max(np.abs(got)))
    return float(np.max(np.abs(ref - got)) / scale)


def interior_u(a, n1, n2, n3, comp=None, j0=1):
    """interior slice of a U-like array: I=1..N1-1, J=j0..N2-1, K=1..N3-1"""
    sl = (slice(1, n1), slice(j0, n2), slice(1, n3))
    return a[sl] if comp is None else a[sl + (comp,)]


def interior_p(a, n1, n2, n3):
    return a[: n1 - 1, : n2 - 1, : n3 - 1]


# values that exercise the F20.10 edge cases of WRITEUP's records: signed
# zero, rounding at the 10th decimal, NaN, infinities and overflow to
# asterisks (|x| >= 1e9, or <= -1e8)
F20_SPECIAL = (-0.0, -1e-12, 5e-11, 4.9999e-11, 2.5e-10, 1.00000000005, 0.125, float("nan"), float("inf"),
               -float("inf"), 2e9, 999999999.9999, 1e9, -99999999.99, -1e8, -123456789.5)


def writeup_inputs(orc, seed=7):
    """U, P for the WRITEUP cases: the seeded channel field of inputs(), with
    F20_SPECIAL planted in P(1:16,1,1), and a random bottom-wall TAUW(I,K,1)
    (I = 1..N1-1, K = 1..N3-1) as an array [K-1, I-1]"""
    U, P, _ = inputs(orc, seed=seed)
    for q, v in enumerate(F20_SPECIAL):
        P[q, 0, 0] = v
    rng = np.random.default_rng(2000 + seed)
    tauw = rng.standard_normal((orc.n3 - 1, orc.n1 - 1)) * 1e-3 + 2.5e-3
    return U, P, tauw
