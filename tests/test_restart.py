"""CONTINUE (READUP) files round-trip, and the reference's own reader
(READUP's record structure, compact3d.f:741-746) is what write_continue emits:
checked by reading the file back with the record layout spelled out."""
import numpy as np

from compact3d_amd.restart import read_continue, write_continue


def test_continue_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    n1, n2, n3 = 9, 7, 5
    U = np.zeros((n1 + 1, n2 + 1, n3 + 1, 3), order="F")
    U[1:n1, 0:n2 + 1, 1:n3, :] = rng.standard_normal((n1 - 1, n2 + 1, n3 - 1, 3))
    P = np.zeros((n1, n2, n3), order="F")
    P[: n1 - 1, : n2 - 1, : n3 - 1] = rng.standard_normal((n1 - 1, n2 - 1, n3 - 1))
    f = tmp_path / "CONTINUE"
    write_continue(str(f), U, P, 0.25, -0.5)
    U2, P2, a, b = read_continue(str(f), n1, n2, n3)
    assert np.array_equal(U, U2) and np.array_equal(P, P2) and (a, b) == (0.25, -0.5)
    # record 2 lists U(I,J,K,1),U(I,J,K,2),U(I,J,K,3) with K innermost, then J, then I
    raw = f.read_bytes()
    n = int(np.frombuffer(raw[:4], "<i4")[0])
    off = 4 + n + 4
    rec2 = np.frombuffer(raw[off + 4: off + 4 + (n1 - 1) * (n2 + 1) * (n3 - 1) * 3 * 8], "<f8")
    assert rec2[0] == U[1, 0, 1, 0] and rec2[1] == U[1, 0, 1, 1] and rec2[2] == U[1, 0, 1, 2]
    assert rec2[3] == U[1, 0, 2, 0]
    assert rec2[3 * (n3 - 1)] == U[1, 1, 1, 0]
