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


IWM_KAPPA, IWM_B = 0.4, 5.0


def iwm_equil(u, dz, re):
    """the log-law friction velocity u/ut = ln(ut dz re)/0.4 + 5 (bisection
    on ln ut; independent of the models' Newton iterations)"""
    u = np.asarray(u, dtype=np.float64)
    lo, hi = np.full(u.shape, -30.0), np.full(u.shape, 5.0)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        ut = np.exp(mid)
        f = u / ut - np.log(ut * dz * re) / IWM_KAPPA - IWM_B
        lo = np.where(f > 0, mid, lo)
        hi = np.where(f > 0, hi, mid)
    return np.exp(0.5 * (lo + hi))


def iwm_integrals(utx, uty, ux, uy, dz, re):
    """the wall-layer integrals of u and w (the residual terms of the
    integral wall model's balance) for friction velocities utx, uty and the
    velocity (ux, uy) at the height dz"""
    k = IWM_KAPPA
    us = (utx ** 4 + uty ** 4) ** 0.25
    delv = 1.0 / re / us
    deli = np.minimum(11.0 / re / us, dz)
    vel = np.sqrt(ux * ux + uy * uy)
    lg = np.log(deli / dz)
    out = []
    for ut, u in ((utx, ux), (uty, uy)):
        a = (u / us + u / vel / k * lg - deli / delv * ut / us) / (1 - deli / dz)
        c = u / us - a
        out.append(0.5 * ut * deli ** 2 / delv + us * dz * (0.5 * a * (1 - deli ** 2 / dz ** 2) + c * (1 - deli / dz)
                                                          - u / vel / k * (1 - deli / dz + deli / dz * lg)))
    return out


def iwm_inputs(n1, n3, dy, re, dt, seed=0):
    """A state of the integral wall model (WMODEL = 2) consistent with the
    log law, so that its Newton iteration converges with small linear
    corrections: the velocity of the third cell from the wall (a log-law
    speed with a seeded variation, spanwise component positive), friction
    velocities near the equilibrium value along the velocity's direction, the wall-layer integrals that
    go with them, and a time-filter weight of 0.2.  Returns U
    ((0:N1,0:N2,0:N3,3) sized from dy, only J = 3 set), P (first cell set),
    the 43 state planes [plane, K-1, I-1] and wm_dt four steps into a
    five-step cycle."""
    from pyoracle import IWM_NPL, IWM_PLANES
    n1m, n3m, n2 = n1 - 1, n3 - 1, len(dy) - 1
    rng = np.random.default_rng(seed)
    dz = dy[1] + dy[2] + dy[3] / 2.0
    ut0 = 0.045
    u0 = ut0 * (np.log(dz * ut0 * re) / IWM_KAPPA + IWM_B)
    x = np.arange(n1m)[None, :] * 2 * PI / n1m
    z = np.arange(n3m)[:, None] * 2 * PI / n3m
    ux = u0 * (1 + 0.08 * np.sin(x) * np.cos(2 * z) + 0.02 * rng.standard_normal((n3m, n1m)))
    uy = u0 * (0.15 + 0.05 * np.cos(x + z) + 0.01 * rng.standard_normal((n3m, n1m)))
    vel = np.sqrt(ux * ux + uy * uy)
    ut = iwm_equil(vel, dz, re) * (1 + 0.01 * rng.standard_normal((n3m, n1m)))
    utx, uty = ut * ux / vel, ut * uy / vel
    il, iv = iwm_integrals(utx, uty, ux, uy, dz, re)
    pl = np.zeros((IWM_NPL, n3m, n1m))
    ix = {name: q for q, name in enumerate(IWM_PLANES)}
    for name, v in (("utx", utx), ("uty", uty), ("tauwx", utx * utx), ("tauwy", uty * uty),
                    ("flt_u", ux), ("flt_w", uy), ("flt_um", ux), ("flt_wm", uy),
                    ("inte_lu", il), ("inte_lv", iv), ("inte_luu", ux * ux * dz), ("inte_lvv", uy * uy * dz),
                    ("inte_luv", ux * uy * dz), ("intem_lu", il), ("intem_lv", iv), ("flt_us", ut),
                    ("tr", np.full_like(ux, 0.2)), ("dz", np.full_like(ux, dz)), ("delv", 1.0 / re / ut),
                    ("deli", np.minimum(11.0 / re / ut, dz))):
        pl[ix[name]] = v
    U = np.zeros((n1 + 1, n2 + 1, n3 + 1, 3), order="F")
    U[1:n1, 3, 1:n3, 0] = (ux * (1 + 0.01 * rng.standard_normal(ux.shape))).T
    U[1:n1, 3, 1:n3, 2] = (uy * (1 + 0.01 * rng.standard_normal(uy.shape))).T
    P = np.zeros((n1, n2, n3), order="F")
    P[:n1m, 0, :n3m] = 1e-3 * rng.standard_normal((n1m, n3m))
    return U, P, pl, 4 * dt
