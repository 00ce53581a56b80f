"""Seeded synthetic SDP instances in SDPA's .dat-s convention.

Every generator returns an :class:`SdpData` holding the problem exactly as a
``.dat-s`` file would: ``m``, ``blockStruct`` (LP blocks negative), the cost
vector ``c`` and 1-based records ``(k, l, i, j, v)`` where ``k = 0`` is F0.
SDPA solves ``min c.x  s.t.  sum_i F_i x_i - F_0 >= 0`` and its dual
``max F_0.Y  s.t.  F_i.Y = c_i, Y >= 0`` (manual section 4); internally it
stores ``C = -F_0`` and ``A_i = F_i`` (sdpa_io.cpp:510-529).

Workloads (the sizes are the ones the benchmarks and tests quote):

* ``maxcut``      -- MaxCut relaxation of G(n, p): F0 = L/4, F_i = e_i e_i^T, c = 1.
* ``lovasz``      -- Lovasz theta of G(n, p): F0 = J, F_1 = I (c_1 = 1),
                     one F_e = e_a e_b^T + e_b e_a^T per edge (c_e = 0).
* ``dense_random``-- one dense block, A_i symmetric N(0,1), feasible by construction.
* ``multi_block`` -- many small sparse SDP blocks plus an LP block, feasible by construction.
* ``maxcut_batch``-- independent small MaxCut instances (optional trace row).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class SdpData:
    m: int
    block_struct: np.ndarray  # int32, LP blocks negative
    c: np.ndarray             # float64, length m
    k: np.ndarray             # int32 records, 0 = F0
    l: np.ndarray             # int32, 1-based block
    i: np.ndarray             # int32, 1-based row
    j: np.ndarray             # int32, 1-based column
    v: np.ndarray             # float64
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.k.shape[0])

    @property
    def n_block(self) -> int:
        return int(self.block_struct.shape[0])

    def packed_size(self) -> int:
        """Length of the packed block space (SDP blocks n*n, then LP entries)."""
        t = 0
        for s in self.block_struct:
            s = int(s)
            t += s * s if s >= 2 else abs(s)
        return t


def _records(recs):
    if not recs:
        z = np.zeros(0, np.int32)
        return z, z, z, z, np.zeros(0)
    a = np.asarray(recs, dtype=np.float64)
    return (a[:, 0].astype(np.int32), a[:, 1].astype(np.int32), a[:, 2].astype(np.int32),
            a[:, 3].astype(np.int32), a[:, 4].copy())


def _make(m, bs, c, K, L, I, J, V, name, **meta) -> SdpData:
    return SdpData(m=int(m), block_struct=np.asarray(bs, np.int32), c=np.asarray(c, np.float64),
                   k=np.asarray(K, np.int32), l=np.asarray(L, np.int32), i=np.asarray(I, np.int32),
                   j=np.asarray(J, np.int32), v=np.asarray(V, np.float64), name=name, meta=meta)


def erdos_renyi(n: int, p: float, rng: np.random.Generator) -> np.ndarray:
    """Upper-triangle edge list (a < b, 0-based) of G(n, p)."""
    iu, ju = np.triu_indices(n, 1)
    keep = rng.random(iu.shape[0]) < p
    return np.stack([iu[keep], ju[keep]], axis=1)


def maxcut(n: int = 100, p: float = 0.1, seed: int = 1, trace_row: bool = False) -> SdpData:
    rng = np.random.default_rng(seed)
    edges = erdos_renyi(n, p, rng)
    deg = np.zeros(n)
    np.add.at(deg, edges[:, 0], 1.0)
    np.add.at(deg, edges[:, 1], 1.0)
    # F0 = L/4 (upper triangle): diagonal deg/4, off-diagonal -1/4 on edges
    K = [np.zeros(n, np.int32), np.zeros(len(edges), np.int32)]
    I = [np.arange(1, n + 1), edges[:, 0] + 1]
    J = [np.arange(1, n + 1), edges[:, 1] + 1]
    V = [deg / 4.0, -0.25 * np.ones(len(edges))]
    keep = V[0] != 0.0
    K[0], I[0], J[0], V[0] = K[0][keep], I[0][keep], J[0][keep], V[0][keep]
    K.append(np.arange(1, n + 1)); I.append(np.arange(1, n + 1)); J.append(np.arange(1, n + 1))
    V.append(np.ones(n))
    m = n
    c = list(np.ones(n))
    if trace_row:
        m = n + 1
        K.append(np.full(n, n + 1)); I.append(np.arange(1, n + 1)); J.append(np.arange(1, n + 1))
        V.append(np.ones(n))
        c.append(float(n))
    K = np.concatenate(K); I = np.concatenate(I); J = np.concatenate(J); V = np.concatenate(V)
    return _make(m, [n], c, K, np.ones_like(K), I, J, V, f"maxcut_n{n}_p{p}_s{seed}",
                 edges=len(edges))


def lovasz(n: int = 200, p: float = 0.3, seed: int = 2) -> SdpData:
    rng = np.random.default_rng(seed)
    edges = erdos_renyi(n, p, rng)
    ne = len(edges)
    m = ne + 1
    iu, ju = np.triu_indices(n)
    K = [np.zeros(iu.shape[0], np.int32), np.ones(n, np.int32), np.arange(2, ne + 2)]
    I = [iu + 1, np.arange(1, n + 1), edges[:, 0] + 1]
    J = [ju + 1, np.arange(1, n + 1), edges[:, 1] + 1]
    V = [np.ones(iu.shape[0]), np.ones(n), np.ones(ne)]
    c = np.zeros(m)
    c[0] = 1.0
    K = np.concatenate(K); I = np.concatenate(I); J = np.concatenate(J); V = np.concatenate(V)
    return _make(m, [n], c, K, np.ones_like(K), I, J, V, f"lovasz_n{n}_p{p}_s{seed}", edges=ne)


def _rand_orth(n, rng):
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))


def dense_random(n: int = 500, m: int = 2000, seed: int = 3, as_arrays: bool = False):
    """One dense n x n block, A_i = sym(N(0,1)); X0 = Q D Q^T, D ~ U[1,10],
    b = A(X0); C = sum y0_i A_i + S0 with y0 ~ N(0,1) and S0 built like X0.
    File form: F_i = A_i, c = b, F0 = -C."""
    rng = np.random.default_rng(seed)
    iu, ju = np.triu_indices(n)
    nu = iu.shape[0]
    Aup = rng.standard_normal((m, nu))  # upper-triangle entries of every A_i
    Q = _rand_orth(n, rng)
    X0 = (Q * rng.uniform(1, 10, n)) @ Q.T
    Q2 = _rand_orth(n, rng)
    S0 = (Q2 * rng.uniform(1, 10, n)) @ Q2.T
    y0 = rng.standard_normal(m)
    w = np.where(iu == ju, 1.0, 2.0)
    b = Aup @ (w * X0[iu, ju])
    Cup = Aup.T @ y0 + S0[iu, ju]
    if as_arrays:
        return dict(n=n, m=m, Aup=Aup, b=b, Cup=Cup, X0=X0, S0=S0, y0=y0, iu=iu, ju=ju)
    K = np.concatenate([np.zeros(nu, np.int32), np.repeat(np.arange(1, m + 1, dtype=np.int32), nu)])
    I = np.concatenate([iu + 1, np.tile(iu + 1, m)])
    J = np.concatenate([ju + 1, np.tile(ju + 1, m)])
    V = np.concatenate([-Cup, Aup.reshape(-1)])
    return _make(m, [n], b, K, np.ones_like(K), I, J, V, f"dense_n{n}_m{m}_s{seed}")


def multi_block(n_sdp_blocks: int = 200, nmin: int = 5, nmax: int = 50, lp: int = 1000,
                m: int = 5000, density: float = 1e-3, seed: int = 5, lp_per_con: int = 1) -> SdpData:
    """Block-diagonal SDP: sparse symmetric N(0,1) data at the given density
    per block (at least one entry per constraint), an LP block of size lp,
    made feasible like dense_random (X0, S0 positive definite per block)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, n_sdp_blocks)
    bs = list(sizes) + ([-lp] if lp > 0 else [])
    recs = {}  # (k, l) -> list of (i, j, v)
    for k in range(1, m + 1):
        placed = False
        while not placed:
            for bl, n in enumerate(sizes):
                cnt = rng.binomial(n * (n + 1) // 2, density)
                for _ in range(cnt):
                    a, b2 = sorted(rng.integers(0, n, 2))
                    recs.setdefault((k, bl + 1), {})[(a, b2)] = rng.standard_normal()
                    placed = True
            if lp > 0:
                for t in rng.choice(lp, lp_per_con, replace=False):
                    recs.setdefault((k, len(sizes) + 1), {})[(t, t)] = rng.standard_normal()
                    placed = True
    # feasible point
    X0 = [(_q := _rand_orth(n, rng)) * rng.uniform(1, 10, n) @ _q.T for n in sizes]
    S0 = [(_q := _rand_orth(n, rng)) * rng.uniform(1, 10, n) @ _q.T for n in sizes]
    x0lp = rng.uniform(1, 10, lp)
    s0lp = rng.uniform(1, 10, lp)
    y0 = rng.standard_normal(m)
    b = np.zeros(m)
    C = [S.copy() for S in S0]
    Clp = s0lp.copy()
    for (k, l), ent in recs.items():
        for (a, b2), v in ent.items():
            if l <= len(sizes):
                w = 1.0 if a == b2 else 2.0
                b[k - 1] += v * w * X0[l - 1][a, b2]
                C[l - 1][a, b2] += y0[k - 1] * v
                if a != b2:
                    C[l - 1][b2, a] += y0[k - 1] * v
            else:
                b[k - 1] += v * x0lp[a]
                Clp[a] += y0[k - 1] * v
    K, L, I, J, V = [], [], [], [], []
    for bl, n in enumerate(sizes):
        iu, ju = np.triu_indices(n)
        vals = -C[bl][iu, ju]
        K.append(np.zeros(iu.shape[0], np.int32)); L.append(np.full(iu.shape[0], bl + 1))
        I.append(iu + 1); J.append(ju + 1); V.append(vals)
    if lp > 0:
        K.append(np.zeros(lp, np.int32)); L.append(np.full(lp, len(sizes) + 1))
        I.append(np.arange(1, lp + 1)); J.append(np.arange(1, lp + 1)); V.append(-Clp)
    for (k, l), ent in sorted(recs.items()):
        for (a, b2), v in ent.items():
            K.append([k]); L.append([l]); I.append([a + 1]); J.append([b2 + 1]); V.append([v])
    K = np.concatenate(K); L = np.concatenate(L); I = np.concatenate(I); J = np.concatenate(J)
    V = np.concatenate(V)
    return _make(m, bs, b, K, L, I, J, V, f"multiblock_{n_sdp_blocks}_lp{lp}_m{m}_s{seed}")


def block_separated(n_blocks: int = 2, n: int = 12, per_block: int = 60, nnz: int = 3,
                    seed: int = 14, couple: int = 0) -> SdpData:
    """SDP blocks whose constraints never share a block: constraint k has nnz
    N(0,1) entries in block k // per_block only, so the aggregate Schur
    pattern is block diagonal (the shape for which SDPA's ordering_bMat
    leaves the dense path once m > 100).  couple > 0 appends that many
    constraints with one diagonal entry in every block (an arrow pattern).
    Feasible like dense_random."""
    rng = np.random.default_rng(seed)
    m = n_blocks * per_block + couple
    X0 = [(_q := _rand_orth(n, rng)) * rng.uniform(1, 10, n) @ _q.T for _ in range(n_blocks)]
    S0 = [(_q := _rand_orth(n, rng)) * rng.uniform(1, 10, n) @ _q.T for _ in range(n_blocks)]
    y0 = rng.standard_normal(m)
    b = np.zeros(m)
    C = [S.copy() for S in S0]
    K, L, I, J, V = [], [], [], [], []
    for k in range(n_blocks * per_block):
        bl = k // per_block
        ent = {}
        while len(ent) < nnz:
            a, c = sorted(rng.integers(0, n, 2))
            ent[(a, c)] = rng.standard_normal()
        for (a, c), v in sorted(ent.items()):
            w = 1.0 if a == c else 2.0
            b[k] += v * w * X0[bl][a, c]
            C[bl][a, c] += y0[k] * v
            if a != c:
                C[bl][c, a] += y0[k] * v
            K.append(k + 1); L.append(bl + 1); I.append(a + 1); J.append(c + 1); V.append(v)
    for t in range(couple):
        k = n_blocks * per_block + t
        for bl in range(n_blocks):
            a = int(rng.integers(0, n))
            v = 1.0 + rng.uniform()
            b[k] += v * X0[bl][a, a]
            C[bl][a, a] += y0[k] * v
            K.append(k + 1); L.append(bl + 1); I.append(a + 1); J.append(a + 1); V.append(v)
    for bl in range(n_blocks):
        iu, ju = np.triu_indices(n)
        for a, c in zip(iu, ju):
            K.append(0); L.append(bl + 1); I.append(a + 1); J.append(c + 1); V.append(-C[bl][a, c])
    return _make(m, [n] * n_blocks, b, K, L, I, J, V,
                 f"block_separated_{n_blocks}x{n}_m{m}_s{seed}" + (f"_c{couple}" if couple else ""))


def maxcut_batch(count: int = 4096, nmin: int = 8, nmax: int = 64, p: float = 0.2, seed: int = 4):
    """Independent MaxCut instances; every fourth one carries a trace row."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        out.append(maxcut(n, p, seed=int(rng.integers(0, 2**31 - 1)), trace_row=(t % 4 == 0)))
    return out


def example1() -> SdpData:
    """LP64/example1.dat-s (manual section 4.1): m=3, one 2x2 block."""
    recs = [(0, 1, 1, 1, -11), (0, 1, 2, 2, 23), (1, 1, 1, 1, 10), (1, 1, 1, 2, 4),
            (2, 1, 2, 2, -8), (3, 1, 1, 2, -8), (3, 1, 2, 2, -2)]
    K, L, I, J, V = _records(recs)
    return _make(3, [2], [48, -8, 20], K, L, I, J, V, "example1")


def write_dats(d: SdpData, path: str) -> None:
    """Write an SdpData in the .dat-s text format read by sdpa_io.cpp:33-66, 473-606."""
    with open(path, "w") as f:
        f.write(f'"{d.name}"\n{d.m} = mDIM\n{d.n_block} = nBLOCK\n')
        f.write(" ".join(str(int(s)) for s in d.block_struct) + " = bLOCKsTRUCT\n")
        f.write(" ".join(repr(float(x)) for x in d.c) + "\n")
        for r in range(d.nnz):
            f.write(f"{d.k[r]} {d.l[r]} {d.i[r]} {d.j[r]} {float(d.v[r])!r}\n")
