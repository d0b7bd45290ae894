"""Pure-Python restatement of the reference's accessibility-grid logic.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
path's accessibility lookups and of pm_block_pockets.  Small grids only
(plain Python loops).

Follows src/grid.jl:
  arg_nearest_neighbor   :778-780   1 + round(xf (n - 1)), ties to even
  apply_pbc_to_index     :783-793   0 -> n, n + 1 -> 1
  accessible             :813-836   any of the 27 points around the nearest
  segment_grid           :383-456   flood fill in scan order (i outermost)
  connections            :485-530   face pairs (c, a, b faces) + reverses
  classify               :533-629   simple cycles with nonzero net translation
  block                  :632-646, 704-764

The simple cycles are enumerated by a plain depth-first search over simple
paths (every cycle found once, from its smallest vertex), not by Johnson's
blocking algorithm, so it checks the library's search rather than repeats it.
"""
from __future__ import annotations

from collections import deque

import numpy as np


def arg_nearest_neighbor(n_pts, xf):
    return [1 + int(np.rint(float(x) * float(int(n) - 1))) for x, n in zip(xf, n_pts)]


def apply_pbc_to_index(idx, n_pts):
    idx = list(idx)
    for k in range(3):
        if idx[k] == 0:
            idx[k] = n_pts[k]
        if idx[k] == n_pts[k] + 1:
            idx[k] = 1
    return idx


def accessible(data, xf, repfactors=None):
    """data: (n1, n2, n3) booleans; xf fractional (of the replicated box when
    repfactors is given)."""
    data = np.asarray(data)
    n_pts = data.shape
    if repfactors is not None:
        xf = [(float(x) * float(r)) % 1.0 for x, r in zip(xf, repfactors)]
    nn = arg_nearest_neighbor(n_pts, xf)
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                a, b, c = apply_pbc_to_index((nn[0] + i, nn[1] + j, nn[2] + k), n_pts)
                if data[a - 1, b - 1, c - 1]:
                    return True
    return False


def segment_grid(energy, tol):
    """Flood-fill labels (n1, n2, n3): 1.. segments, -1 not occupiable."""
    e = np.asarray(energy, dtype=np.float64)
    n1, n2, n3 = e.shape
    seg = np.zeros(e.shape, dtype=np.int64)
    n_seg = 0
    for i in range(n1):
        for j in range(n2):
            for k in range(n3):
                if seg[i, j, k] != 0:
                    continue
                if e[i, j, k] > tol:
                    seg[i, j, k] = -1
                    continue
                n_seg += 1
                queue = deque([(i, j, k)])
                while queue:
                    p = queue.popleft()
                    seg[p] = n_seg
                    for di in (-1, 0, 1):
                        for dj in (-1, 0, 1):
                            for dk in (-1, 0, 1):
                                if di == 0 and dj == 0 and dk == 0:
                                    continue
                                q = (p[0] + di, p[1] + dj, p[2] + dk)
                                if min(q) < 0 or q[0] >= n1 or q[1] >= n2 or q[2] >= n3:
                                    continue
                                if seg[q] != 0:
                                    continue
                                if e[q] < tol:
                                    seg[q] = n_seg
                                    queue.append(q)
                                else:
                                    seg[q] = -1
    return seg, n_seg


def connections(seg):
    """Ordered list of (src, dst, (da, db, dc)) with 1-based segments."""
    n1, n2, n3 = seg.shape
    out = []

    def note(s1, s2, d):
        if s1 == -1 or s2 == -1:
            return
        c = (int(s1), int(s2), d)
        if c in out:
            return
        out.append(c)
        out.append((int(s2), int(s1), (-d[0], -d[1], -d[2])))

    for i in range(n1):
        for j in range(n2):
            note(seg[i, j, n3 - 1], seg[i, j, 0], (0, 0, 1))
    for j in range(n2):
        for k in range(n3):
            note(seg[n1 - 1, j, k], seg[0, j, k], (1, 0, 0))
    for i in range(n1):
        for k in range(n3):
            note(seg[i, n2 - 1, k], seg[i, 0, k], (0, 1, 0))
    return out


def expanded_graph(n_seg, conns):
    """Vertices 0..n_seg-1 segments, then one per connection; adjacency and
    the translation carried by each vertex."""
    nv = n_seg + len(conns)
    adj = [[] for _ in range(nv)]
    trans = [(0, 0, 0)] * n_seg
    for c, (s, d, t) in enumerate(conns):
        v = n_seg + c
        adj[s - 1].append(v)
        adj[v].append(d - 1)
        trans.append(t)
    return adj, trans


def simple_cycles(adj):
    """All simple cycles (vertex lists, smallest vertex first), by DFS over
    simple paths from each start through larger vertices."""
    cycles = []
    for s in range(len(adj)):
        stack = [(s, iter(adj[s]))]
        on_path = {s}
        path = [s]
        while stack:
            v, it = stack[-1]
            w = next(it, None)
            if w is None:
                stack.pop()
                on_path.discard(v)
                path.pop()
                continue
            if w == s:
                cycles.append(list(path))
            elif w > s and w not in on_path:
                on_path.add(w)
                path.append(w)
                stack.append((w, iter(adj[w])))
    return cycles


def classify(n_seg, conns, cycles=None):
    """0/1 per segment: on a simple cycle with a nonzero net translation."""
    adj, trans = expanded_graph(n_seg, conns)
    if cycles is None:
        cycles = simple_cycles(adj)
    cls = [0] * n_seg
    for cyc in cycles:
        t = [0, 0, 0]
        for v in cyc:
            for a in range(3):
                t[a] += trans[v][a]
        if t != [0, 0, 0]:
            for v in cyc:
                if v < n_seg:
                    cls[v] = 1
    return cls, len(cycles)


def block(energy, tol, block_pockets=True):
    """(accessible bool grid, labels or None, n_segments, n_blocked, n_cycles)."""
    e = np.asarray(energy, dtype=np.float64)
    if not block_pockets:
        return e < tol, None, 0, 0, 0
    seg, n_seg = segment_grid(e, tol)
    conns = connections(seg)
    cls, n_cycles = classify(n_seg, conns)
    acc = np.zeros(e.shape, dtype=bool)
    for s in range(1, n_seg + 1):
        if cls[s - 1]:
            acc |= seg == s
    return acc, seg, n_seg, n_seg - sum(cls), n_cycles
