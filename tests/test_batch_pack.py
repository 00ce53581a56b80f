"""Host side of the batched solver: the packing of instances into the C
ABI's concatenated layout (include/sdpa_batch.h), no device calls."""
import numpy as np
import pytest

from sdpa_amd import problems as PR

B = pytest.importorskip("sdpa_amd.batch")


def dense_of(d):
    """F_k as dense symmetric matrices from the records (1-based, either triangle)."""
    n = int(d.block_struct[0])
    F = np.zeros((d.m + 1, n, n))
    for k, i, j, v in zip(d.k, d.i, d.j, d.v):
        F[k, i - 1, j - 1] += v
        if i != j:
            F[k, j - 1, i - 1] += v
    return F


def test_pack_layout():
    insts = [PR.maxcut(12, 0.3, 1, trace_row=True), PR.lovasz(9, 0.4, 2), PR.example1()]
    P = B.Packed(insts)
    assert P.count == 3
    assert list(P.n) == [12, 9, 2] and list(P.m) == [13, insts[1].m, 3]
    for t, d in enumerate(insts):
        n, m = int(P.n[t]), int(P.m[t])
        F = dense_of(d)
        Cm = P.C[P.coff[t]: P.coff[t] + n * n].reshape(n, n, order="F")
        np.testing.assert_array_equal(Cm, -F[0])  # internal C = -F0, both triangles
        np.testing.assert_array_equal(P.b[P.boff[t]: P.boff[t] + m], d.c)
        ptr = P.cptr[P.pptr[t]: P.pptr[t] + m + 1]
        assert np.all(np.diff(ptr) >= 0)
        for i in range(m):
            A = np.zeros((n, n))
            for e in range(ptr[i], ptr[i + 1]):
                r, c = P.er[e], P.ec[e]
                assert 0 <= r <= c < n
                A[r, c] += P.ev[e]
                if r != c:
                    A[c, r] += P.ev[e]
            np.testing.assert_array_equal(A, F[i + 1])


def test_pack_merges_duplicates_and_lower_entries():
    recs = [(0, 1, 1, 1, 1.0), (1, 1, 2, 1, 3.0), (1, 1, 1, 2, 1.0), (2, 1, 2, 2, 1.0)]
    K, L, I, J, V = PR._records(recs)
    d = PR._make(2, [2], [1.0, 1.0], K, L, I, J, V, "dup")
    P = B.Packed([d])
    assert list(P.cptr) == [0, 1, 2]
    assert (P.er[0], P.ec[0], P.ev[0]) == (0, 1, 4.0)


def test_pack_rejects_out_of_range():
    with pytest.raises(ValueError):
        B.Packed([PR.maxcut(65, 0.1, 1)])
    with pytest.raises(ValueError):
        B.Packed([PR.multi_block(2, 5, 8, 3, 4, seed=1)])


def test_config_struct_layout():
    """sdpa_batch_config: four ints then a double (24 bytes, the C layout)."""
    import ctypes
    from sdpa_amd import batch as B
    assert ctypes.sizeof(B.Config) == 24
    assert B.Config.mu_switch.offset == 16
    c = B.config(threads=128, heavy=32, order=1, precision=1, mu_switch=1e-3)
    assert c.as_dict() == {"threads": 128, "heavy": 32, "order": 1, "precision": 1,
                           "mu_switch": 1e-3}
