"""The Fortran drop-in: the project's ISO_C_BINDING module and wrappers named
GETUP/DIVCHECK/CFL/CHKMF/WALLSS/WRITEUP, driven by a MAIN-like loop reading a
READUP CONTINUE file, give the same trajectory and the same WRITEUP file as
the Python binding of the same library; GETUP's host sync puts the stepped
field into the caller's arrays."""
import os
import subprocess

import numpy as np
import pytest

from cases import config, inputs

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "compact3d_amd", "fortran")


@pytest.mark.parametrize("files,wmodel", [(0.0, 0.0), (2.0, 2.0)])
def test_fortran_driver_matches_python(tmp_path, files, wmodel):
    """DNS, and Smagorinsky LES with the integral wall model (its compute step
    NTIME = 5 falls inside the 6 steps: GETUP's NTIME crosses the binding)"""
    from compact3d_amd import c3d
    from compact3d_amd.restart import write_continue
    from pyoracle import Oracle
    exe = os.path.join(FDIR, "c3d_driver")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", FDIR], check=True)
    n1, n2, n3 = 33, 33, 33
    cfg = config(n1, n2, n3, files=files, wmodel=wmodel)
    orc = Oracle(**cfg)
    U, P, _ = inputs(orc, seed=5)
    cont = tmp_path / "CONTINUE"
    write_continue(str(cont), U, P, 0.0, 0.0)
    nst = 6
    args = [exe, str(n1), str(n2), str(n3), repr(cfg["re"]), repr(cfg["alx"]), repr(cfg["aly"]),
            repr(cfg["alz"]), repr(cfg["dt"]), repr(files), repr(wmodel), repr(cfg["cks"]), str(nst), "1", str(cont),
            "0.5", "3"]
    (tmp_path / "RESULT").mkdir()
    out = subprocess.run(args, check=True, capture_output=True, text=True, timeout=120, cwd=str(tmp_path)).stdout
    print(out)
    lines = [ln for ln in out.splitlines() if ln.startswith("STEP=")]
    assert len(lines) == nst
    presg_f = float(lines[-1].split("DP=")[1].split()[0])
    ch = c3d.Channel(**cfg, cflmax=0.5)
    ch.upload(U, P)
    # GETUP's host sync (every 3 steps): the caller's U holds the stepped field
    syncs = {int(ln.split("=")[1].split()[0]): float(ln.split("U(5,5,5,1)=")[1])
             for ln in out.splitlines() if ln.startswith("HOSTSYNC=")}
    assert sorted(syncs) == [3, 6]
    for n in range(nst):
        ch.getup(n + 1)
        if n + 1 in syncs:
            u = ch.download()["U"][5, 5, 5, 0]
            assert abs(syncs[n + 1] - u) <= 1e-14 * abs(u), (n + 1, syncs[n + 1], u)
    presg_p = ch.scalar(c3d.S_PRESG)
    u555 = ch.download()["U"][5, 5, 5, 0]
    u_f = float([ln for ln in out.splitlines() if ln.startswith("U(5,5,5,1)=")][0].split("=")[1])
    print("PRESG fortran", presg_f, "python", presg_p, "U", u_f, u555)
    assert abs(presg_f - presg_p) <= 1e-13 * abs(presg_p)
    # the driver prints 16 significant digits
    assert abs(u_f - u555) <= 1e-14 * abs(u555)
    # the driver's final WRITEUP (NTIME = NTST + 1) equals the Python binding's
    from compact3d_amd.restart import result_names
    fname, sname = result_names(nst + 1)
    ch.write_flow(str(tmp_path / "py_flow.plt"), nst + 1)
    assert (tmp_path / "RESULT" / fname).read_bytes() == (tmp_path / "py_flow.plt").read_bytes()
    assert (tmp_path / "RESULT" / sname).exists()
    ch.close()
