"""Global edge ("line") numbering of a tetrahedral mesh.

Restates the rule of MKLINE_V2 and MKN6 (src/common/m_line_type.f90:121-213,219-266), which
write lineinfo.dat for the solver (tmtgem/mesh/mkline.f90:19-30):
  * an edge runs from its lower to its higher global node id;
  * edges are numbered by ascending start node n1; the edges of one start node are numbered
    in discovery order: elements containing n1 in ascending element id, and inside each
    element the local node positions k = 1..4;
  * n6line(t, j) = +-id of local edge j = kl(j) = (1,2),(2,3),(1,3),(1,4),(2,4),(3,4)
    (src/solver/m_fem_edge_util.f90:141-147), negative when n4(t,kl(j,1)) > n4(t,kl(j,2)).
The result is vectorised with numpy sorts instead of per-node linear searches.
"""
from __future__ import annotations

import numpy as np

# Local edge -> local node pairs (0-based), m_fem_edge_util.f90:141-147.
KL = np.array([[0, 1], [1, 2], [0, 2], [0, 3], [1, 3], [2, 3]], dtype=np.int64)


def mkline_mkn6(nodeg: int, n4: np.ndarray):
    """Return (line (nline,2) int32 1-based, n6line (ntet,6) int32 signed 1-based)."""
    n4 = np.asarray(n4, dtype=np.int64)
    ntet = n4.shape[0]
    if ntet == 0:
        return np.zeros((0, 2), np.int32), np.zeros((0, 6), np.int32)
    if n4.min() < 1 or n4.max() > nodeg:
        raise ValueError("node id out of range")
    a_idx, b_idx = np.nonzero(~np.eye(4, dtype=bool))  # 12 ordered local pairs
    na = n4[:, a_idx]          # (ntet, 12)
    nb = n4[:, b_idx]
    keep = na < nb             # the 6 pairs with n1 < n2
    e = np.broadcast_to(np.arange(ntet)[:, None], na.shape)[keep]
    bpos = np.broadcast_to(b_idx[None, :], na.shape)[keep]
    n1 = na[keep]
    n2 = nb[keep]
    # first discovery of each (n1, n2): smallest element id (b is then unique)
    o = np.lexsort((e, n2, n1))
    n1, n2, e, bpos = n1[o], n2[o], e[o], bpos[o]
    first = np.ones(n1.size, bool)
    first[1:] = (n1[1:] != n1[:-1]) | (n2[1:] != n2[:-1])
    n1, n2, e, bpos = n1[first], n2[first], e[first], bpos[first]
    o = np.lexsort((bpos, e, n1))
    line = np.stack([n1[o], n2[o]], axis=1)
    nline = line.shape[0]
    # n6line: look the (min, max) pair up among the numbered edges
    key_line = line[:, 0] * (nodeg + 1) + line[:, 1]
    srt = np.argsort(key_line)
    s_keys = key_line[srt]
    p = n4[:, KL[:, 0]]
    q = n4[:, KL[:, 1]]
    lo = np.minimum(p, q)
    hi = np.maximum(p, q)
    sgn = np.where(p > q, -1, 1)
    k = lo * (nodeg + 1) + hi
    pos = np.searchsorted(s_keys, k)
    if np.any(pos >= nline) or np.any(s_keys[np.minimum(pos, nline - 1)] != k):
        raise ValueError("edge lookup failed")
    ids = srt[pos] + 1
    n6line = (sgn * ids).astype(np.int32)
    return line.astype(np.int32), n6line
