"""Seeded initial fields for the channel: laminar Poiseuille profile plus a
divergence-free perturbation.

The reference's INIUP (compact3d.f:503-723) draws unseeded uniform noise that
is not divergence-free; runs meant to be reproduced instead inject a field
through READUP (compact3d.f:729-749).  This module builds such a field from a
seed:

* laminar u at cell centres, REALU = YH*(ALY-YH)*(ALY-1) + YH*(2ALY-YH)*(2-ALY)
  (the reference's profile, compact3d.f:627; centreline 1 for ALY=2);
* a perturbation equal to the discrete curl of a vector potential A whose
  components are iid N(0,1) on the grid edges, damped by (1-y^2)^2 with
  y = 2Y/ALY - 1.  The curl uses the same operators as the divergence the
  solver measures (DIVCHECK, compact3d.f:9066-9128): fourth-order compact
  face-to-centre derivatives D1 in x and z and (f_{j+1}-f_j)/DY_j in y, so
  D1x D1z = D1z D1x makes the discrete divergence vanish to rounding;
* A1 and A3 vanish on the walls, so v = 0 there and the perturbation carries
  no net flow rate.

Arrays are returned in the reference's Fortran layouts:
U (N1+1, N2+1, N3+1, 3), P (N1, N2, N3).
"""
from __future__ import annotations

import numpy as np


def _d1_face_to_centre(f: np.ndarray, dx: float, axis: int) -> np.ndarray:
    """x_i + 22 x_i... : solve (1,22,1) x = 24*dx*(f_{i+1} - f_i) periodically
    along `axis` (exactly, as a circulant, in Fourier space)."""
    n = f.shape[axis]
    th = 2.0 * np.pi * np.arange(n) / n
    sym = 24.0 * dx * (np.exp(1j * th) - 1.0) / (22.0 + 2.0 * np.cos(th))
    shape = [1] * f.ndim
    shape[axis] = n
    return np.real(np.fft.ifft(np.fft.fft(f, axis=axis) * sym.reshape(shape), axis=axis))


def laminar_profile(Y: np.ndarray, aly: float) -> np.ndarray:
    """REALU at cell centres J=1..N2M (compact3d.f:626-627); Y is Y(0:N2)."""
    yh = 0.5 * (Y[1:-1] + Y[2:])
    return yh * (aly - yh) * (aly - 1.0) + yh * (2 * aly - yh) * (2.0 - aly)


def perturbed_poiseuille(n1: int, n2: int, n3: int, Y: np.ndarray, dx1: float, dx3: float,
                         aly: float, seed: int, amp: float = 0.1):
    """Return (U, P) in Fortran layout.  `amp` is the perturbation rms as a
    fraction of the bulk velocity of the laminar profile."""
    n1m, n2m, n3m = n1 - 1, n2 - 1, n3 - 1
    Y = np.asarray(Y, dtype=np.float64)
    DY = np.zeros(n2 + 1)
    DY[1:n2] = Y[2:n2 + 1] - Y[1:n2]
    rng = np.random.default_rng(seed)
    # edge potentials on index grids (i 0..n1m-1, j 0..n2-1 for faces J=1..N2, k)
    yf = 2.0 * Y[1:n2 + 1] / aly - 1.0                 # faces J=1..N2
    yc = 2.0 * (0.5 * (Y[1:n2] + Y[2:n2 + 1])) / aly - 1.0   # centres J=1..N2M
    gf = (1.0 - yf ** 2) ** 2
    gc = (1.0 - yc ** 2) ** 2
    gf[0] = 0.0
    gf[-1] = 0.0
    A1 = rng.standard_normal((n1m, n2, n3m)) * gf[None, :, None]   # (ci, fj, fk)
    A2 = rng.standard_normal((n1m, n2m, n3m)) * gc[None, :, None]  # (fi, cj, fk)
    A3 = rng.standard_normal((n1m, n2, n3m)) * gf[None, :, None]   # (fi, fj, ck)
    dy = DY[1:n2][None, :, None]
    u = (A3[:, 1:, :] - A3[:, :-1, :]) / dy - _d1_face_to_centre(A2, dx3, 2)
    v = _d1_face_to_centre(A1, dx3, 2) - _d1_face_to_centre(A3, dx1, 0)   # faces J=1..N2
    w = _d1_face_to_centre(A2, dx1, 0) - (A1[:, 1:, :] - A1[:, :-1, :]) / dy
    lam = laminar_profile(Y, aly)
    ub = float(np.sum(lam * DY[1:n2]) / aly)
    rms = np.sqrt((np.mean(u ** 2) + np.mean(v[:, 1:-1, :] ** 2) + np.mean(w ** 2)) / 3.0)
    s = amp * ub / rms if rms > 0 else 0.0
    U = np.zeros((n1 + 1, n2 + 1, n3 + 1, 3), order="F")
    U[1:n1, 1:n2, 1:n3, 0] = lam[None, :, None] + s * u
    U[1:n1, 1:n2 + 1, 1:n3, 1] = s * v
    U[1:n1, 1:n2, 1:n3, 2] = s * w
    U[1:n1, 1, 1:n3, 1] = 0.0
    U[1:n1, n2, 1:n3, 1] = 0.0
    U[1:n1, n2, 1:n3, 0] = U[1:n1, n2m, 1:n3, 0] * (2.0 - aly)
    U[1:n1, n2, 1:n3, 2] = U[1:n1, n2m, 1:n3, 2] * (2.0 - aly)
    P = np.zeros((n1, n2, n3), order="F")
    return U, P


def uniform_field(n1: int, n2: int, n3: int, seed: int):
    """iid uniform[-1,1] FP64 fields of the solver's shape (isolated operator
    benches): U interior, P interior; ghosts zero."""
    rng = np.random.default_rng(seed)
    U = np.zeros((n1 + 1, n2 + 1, n3 + 1, 3), order="F")
    U[1:n1, 1:n2, 1:n3, :] = rng.uniform(-1.0, 1.0, (n1 - 1, n2 - 1, n3 - 1, 3))
    U[1:n1, n2, 1:n3, 1] = 0.0
    U[1:n1, 1, 1:n3, 1] = 0.0
    P = np.zeros((n1, n2, n3), order="F")
    P[:n1 - 1, :n2 - 1, :n3 - 1] = rng.uniform(-1.0, 1.0, (n1 - 1, n2 - 1, n3 - 1))
    return U, P
