"""The build's checkpoint (compact3d_amd/restart.py): k steps, save, load
into a fresh handle, and the remaining steps equal an uninterrupted run bit
for bit, in FP64 and mixed precision, for DNS and for LES with the models
that carry state between steps (WSM and TAOW read back by SGSCS and
WALLMODEL_WERNER, the integral wall model's planes and NTIME)."""
import numpy as np
import pytest

from cases import config, inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("files,wmodel", [(0.0, 0.0), (2.0, 0.0), (2.0, 2.0), (2.0, 4.0)])
def test_resume_equals_continuous(tmp_path, precision, files, wmodel):
    from compact3d_amd import c3d
    from compact3d_amd.restart import load_checkpoint, save_checkpoint
    from pyoracle import Oracle
    cfg = config(33, 33, 33, re=5200.0 if files else 4200.0, dt=5e-4, files=files, wmodel=wmodel, cks=1.5)
    U, P, _ = inputs(Oracle(**cfg), seed=13)
    kw = dict(cfg, precision=precision, cflmax=0.5)
    k, n = 6, 12  # NTIME 5 and 10 are the integral wall model's compute steps
    print("resume case precision", precision, "files", files, "wmodel", wmodel, flush=True)
    a = c3d.Channel(**kw)
    a.upload(U, P)
    if files:
        a.set_scalar(c3d.S_WSM1, 1e-3)
        a.set_scalar(c3d.S_WSM3, 1e-4)
    for nt in range(1, k + 1):
        a.step(nt)
    path = str(tmp_path / "ck.npz")
    save_checkpoint(path, a)
    b = c3d.Channel(**kw)
    load_checkpoint(path, b)
    for nt in range(k + 1, n + 1):
        da, ca = a.step(nt)
        db, cb = b.step(nt)
        assert da == db and ca == cb, nt
    sa, sb = a.download(), b.download()
    for key in ("U", "P", "HH", "SGSVIS"):
        assert np.array_equal(sa[key], sb[key]), key
    assert sa["PRESG"] == sb["PRESG"] and sa["PRESG3"] == sb["PRESG3"]
    if wmodel == 2.0:
        assert np.array_equal(a.field(c3d.F_IWM), b.field(c3d.F_IWM))
    a.close()
    b.close()


def test_checkpoint_refuses_other_models(tmp_path):
    from compact3d_amd import c3d
    from compact3d_amd.restart import load_checkpoint, save_checkpoint
    cfg = config(33, 33, 33)
    a = c3d.Channel(**cfg)
    a.upload(np.zeros((34, 34, 34, 3), order="F"), np.zeros((33, 33, 33), order="F"))
    path = str(tmp_path / "ck.npz")
    save_checkpoint(path, a)
    b = c3d.Channel(**dict(cfg, files=2.0, wmodel=1.0))
    with pytest.raises(ValueError):
        load_checkpoint(path, b)
    a.close()
    b.close()
