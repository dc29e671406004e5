"""On-disk state formats.

* READUP's CONTINUE file (compact3d.f:741-746): Fortran unformatted
  sequential records  PRESG,PRESG3 / U(I,J,K,1:3) with K fastest, J=0..N2,
  I=1..N1M / P(I,J,K) with K fastest, J=1..N2M, I=1..N1M.  (The reference's
  own CWRITE writes a formatted file that READUP cannot read, :854-872; these
  functions are the pair READUP expects.)
* WRITEUP's Tecplot output (compact3d.f:757-835): RESULT/FLOW-<100000+NTIME>.plt
  (cell-centre U, V, W, P) and RESULT/STRAIN-<100000+NTIME>.plt (the
  bottom-wall shear stress fluctuation over UTAU**2), text as the
  reference's Fortran runtime writes it (write_result).
* The build's own checkpoint: an .npz with U, P, HH (the AB2 history that
  CONTINUE omits), SGSVIS, PRESG, PRESG3, the grid and model switches, and
  the state the wall and SGS models carry from step to step: WSM(1:3) and
  TAOW (SGSCS's friction velocity and WALLMODEL_WERNER's viscous branch read
  the previous step's), TAUW/TAUB, the integral wall model's 43 planes with
  its NTIME, wm_dt, CFLMAX and initial means (WMODEL = 2), and in mixed
  precision the FP64 divergence the next step starts from.  A run resumed
  from it continues bit for bit.
"""
from __future__ import annotations

import numpy as np

_MARK = np.dtype("<i4")


def _rec(f, arr):
    b = np.ascontiguousarray(arr, dtype="<f8").tobytes()
    np.array([len(b)], _MARK).tofile(f)
    f.write(b)
    np.array([len(b)], _MARK).tofile(f)


def _read_rec(f):
    n = int(np.fromfile(f, _MARK, 1)[0])
    data = np.frombuffer(f.read(n), dtype="<f8")
    m = int(np.fromfile(f, _MARK, 1)[0])
    if m != n:
        raise ValueError("corrupt record markers")
    return data


def write_continue(path, U, P, presg, presg3):
    """U (N1+1, N2+1, N3+1, 3) and P (N1, N2, N3) in the reference layout."""
    n1, n2, n3 = P.shape
    with open(path, "wb") as f:
        _rec(f, np.array([presg, presg3]))
        # (((U(I,J,K,1),U(I,J,K,2),U(I,J,K,3),K=1,N3M),J=0,N2),I=1,N1M)
        _rec(f, np.transpose(U[1:n1, 0:n2 + 1, 1:n3, :], (0, 1, 2, 3)))
        # (((P(I,J,K),K=1,N3M),J=1,N2M),I=1,N1M)
        _rec(f, P[: n1 - 1, : n2 - 1, : n3 - 1])


def read_continue(path, n1, n2, n3):
    U = np.zeros((n1 + 1, n2 + 1, n3 + 1, 3), order="F")
    P = np.zeros((n1, n2, n3), order="F")
    with open(path, "rb") as f:
        presg, presg3 = _read_rec(f)
        U[1:n1, 0:n2 + 1, 1:n3, :] = _read_rec(f).reshape(n1 - 1, n2 + 1, n3 - 1, 3)
        P[: n1 - 1, : n2 - 1, : n3 - 1] = _read_rec(f).reshape(n1 - 1, n2 - 1, n3 - 1)
    return U, P, float(presg), float(presg3)


def _carried_scalars(channel):
    from . import c3d
    out = [c3d.S_WSM1, c3d.S_WSM2, c3d.S_WSM3, c3d.S_TAOW]
    if channel.wmodel == 2.0:
        out += [c3d.S_NTIME, c3d.S_WMDT, c3d.S_CFLMAX, c3d.S_UINIT, c3d.S_USINIT, c3d.S_PBAR]
    return out


def save_checkpoint(path, channel):
    """the resident state of `channel` (a c3d.Channel) to an .npz"""
    from . import c3d
    st = channel.download()
    keys = _carried_scalars(channel)
    z = dict(U=st["U"], P=st["P"], HH=st["HH"], SGSVIS=st["SGSVIS"], PRESG=st["PRESG"], PRESG3=st["PRESG3"],
             grid=np.array([channel.n1, channel.n2, channel.n3]),
             model=np.array([channel.files, channel.wmodel, channel.precision], dtype=np.float64),
             scalar_ids=np.array(keys), scalars=np.array([channel.scalar(k) for k in keys]),
             TAUW=channel.field(c3d.F_TAUW), TAUB=channel.field(c3d.F_TAUB))
    if channel.wmodel == 2.0:
        z["IWM"] = channel.field(c3d.F_IWM)
    if channel.precision == c3d.MIXED:
        z["DIV"] = channel.field(c3d.F_DIV)
    np.savez(path, **z)


def load_checkpoint(path, channel):
    """restores a save_checkpoint file into `channel`, which must have the
    same grid, FILES, WMODEL and precision"""
    from . import c3d
    z = np.load(path, allow_pickle=False)
    if tuple(z["grid"]) != (channel.n1, channel.n2, channel.n3):
        raise ValueError("checkpoint grid does not match the channel")
    if tuple(z["model"]) != (channel.files, channel.wmodel, float(channel.precision)):
        raise ValueError("checkpoint FILES, WMODEL or precision does not match the channel")
    channel.upload(np.asfortranarray(z["U"]), np.asfortranarray(z["P"]), HH=np.asfortranarray(z["HH"]),
                   SGSVIS=np.asfortranarray(z["SGSVIS"]), presg=float(z["PRESG"]), presg3=float(z["PRESG3"]))
    for k, v in zip(z["scalar_ids"], z["scalars"]):
        channel.set_scalar(int(k), float(v))
    channel.set_field(c3d.F_TAUW, z["TAUW"])
    channel.set_field(c3d.F_TAUB, z["TAUB"])
    if "IWM" in z:
        channel.set_field(c3d.F_IWM, z["IWM"])
    if "DIV" in z:
        channel.set_field(c3d.F_DIV, z["DIV"])


def result_names(ntime):
    """WRITEUP's file names: FILENAME = 100000+NTIME written '(I6)' (:782)"""
    v = 100000 + int(ntime)
    tag = "%6d" % v
    if len(tag) > 6:
        tag = "*" * 6
    return "FLOW-%s.plt" % tag, "STRAIN-%s.plt" % tag


def write_result(channel, ntime, utau, root="."):
    """WRITEUP(U,P,PRESG,PRESG3,NTIME) of the resident field: writes
    root/RESULT/FLOW-*.plt and root/RESULT/STRAIN-*.plt.  utau is WRITEUP's
    UTAU: SQRT(WSMO(1)) without a wall model, else SQRT(TAOWO) (:785-789).
    Returns the two paths."""
    import os
    d = os.path.join(root, "RESULT")
    os.makedirs(d, exist_ok=True)
    flow, strain = (os.path.join(d, n) for n in result_names(ntime))
    channel.write_flow(flow, ntime)
    channel.write_strain(strain, ntime, utau)
    return flow, strain

