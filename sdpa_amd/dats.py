"""Readers for SDPA's input formats (.dat dense, .dat-s sparse) in Python.

The token rules follow IO::read / IO::setElement (sdpa_io.cpp:33-66, 473-606):
leading lines that start with '*' or '"' are comments; then m is read from
the first other line; every later number is found by skipping characters
that cannot start a number (the ``%*[^0-9+-]`` scan).  Dense input lists
F0..Fm block by block, SDP blocks as full n x n matrices (only i <= j and
non-zero entries are kept) and LP blocks as their diagonal.
"""
from __future__ import annotations

import re

import numpy as np

from .problems import SdpData

_NUM = re.compile(rb"[+-]?(?:\d+\.?\d*|\.\d+)(?:[eEdD][+-]?\d+)?")


def _numbers(buf: bytes):
    return [float(t.replace(b"d", b"e").replace(b"D", b"e")) for t in _NUM.findall(buf)]


def read(path: str, sparse: bool | None = None) -> SdpData:
    if sparse is None:
        sparse = path.endswith("-s")
    raw = open(path, "rb").read()
    lines = raw.split(b"\n")
    pos = 0
    while pos < len(lines) and (lines[pos][:1] in (b"*", b'"')):
        pos += 1
    m = int(_numbers(lines[pos])[0])
    toks = _numbers(b"\n".join(lines[pos + 1:]))
    t = 0
    nblock = int(toks[t]); t += 1
    bs = [int(x) for x in toks[t:t + nblock]]; t += nblock
    c = np.array(toks[t:t + m]); t += m
    K, L, I, J, V = [], [], [], [], []
    if sparse:
        rest = toks[t:]
        nrec = len(rest) // 5
        arr = np.array(rest[: nrec * 5]).reshape(nrec, 5)
        K, L, I, J, V = (arr[:, 0].astype(np.int32), arr[:, 1].astype(np.int32),
                         arr[:, 2].astype(np.int32), arr[:, 3].astype(np.int32), arr[:, 4])
    else:
        for k in range(m + 1):
            for bl, s in enumerate(bs):
                if s > 0 and s != 1:
                    vals = np.array(toks[t:t + s * s]).reshape(s, s); t += s * s
                    for a in range(s):
                        for b in range(a, s):
                            if vals[a, b] != 0.0:
                                K.append(k); L.append(bl + 1); I.append(a + 1); J.append(b + 1)
                                V.append(vals[a, b])
                else:
                    size = abs(s)
                    vals = toks[t:t + size]; t += size
                    for a in range(size):
                        if vals[a] != 0.0:
                            K.append(k); L.append(bl + 1); I.append(a + 1); J.append(a + 1)
                            V.append(vals[a])
    return SdpData(m=m, block_struct=np.array(bs, np.int32), c=c, k=np.asarray(K, np.int32),
                   l=np.asarray(L, np.int32), i=np.asarray(I, np.int32), j=np.asarray(J, np.int32),
                   v=np.asarray(V, np.float64), name=path.rsplit("/", 1)[-1])
