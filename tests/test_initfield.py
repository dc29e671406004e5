"""The seeded initial field: divergence-free in the solver's own DIVCHECK
metric, laminar flow rate, requested perturbation amplitude, reproducible."""
import numpy as np

from cases import config


def test_perturbed_poiseuille_properties():
    from compact3d_amd.initfield import laminar_profile, perturbed_poiseuille
    from pyoracle import Oracle
    o = Oracle(**config(49, 33, 25, cks=1.5))
    m = o.mesh()
    U, P = perturbed_poiseuille(49, 33, 25, m["Y"], m["DX1"], m["DX3"], m["ALY"], seed=11, amp=0.1)
    assert o.divcheck(U) <= 1e-12
    lam = laminar_profile(m["Y"], m["ALY"])
    dy = m["DY"][1:33]
    flow = np.einsum("ijk,j->", U[1:49, 1:33, 1:25, 0], dy) / (48 * 24)
    assert abs(flow - np.sum(lam * dy)) <= 1e-12
    pert = U[1:49, 1:33, 1:25, 0] - lam[None, :, None]
    ub = np.sum(lam * dy) / m["ALY"]
    rms = np.sqrt((np.mean(pert ** 2) + np.mean(U[1:49, 2:32, 1:25, 1] ** 2) + np.mean(U[1:49, 1:33, 1:25, 2] ** 2)) / 3)
    assert 0.05 * ub < rms < 0.15 * ub
    assert np.all(U[:, 1, :, 1] == 0) and np.all(U[:, 33, :, 1] == 0)
    U2, _ = perturbed_poiseuille(49, 33, 25, m["Y"], m["DX1"], m["DX3"], m["ALY"], seed=11, amp=0.1)
    assert np.array_equal(U, U2)
