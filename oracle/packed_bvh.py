"""Test infrastructure (the oracle): the exact nearest-first closest-point traversal of the
engine's packed wide BVH, restated on the CPU to count the BVH nodes and primitives one
closest-point query reads.

The reference answers each distance query with FCPW's Aggregate::findClosestPoint
(utils/fcpw_geometric_queries.h:1069-1077; FCPW itself is absent).  The algorithmic cost of
that query on the engine's scene layout is the work of an exact nearest-first traversal from
the root: visit a node, prune children whose box lies beyond the best distance found so far,
descend into the nearest, stack the others farthest-first, test the primitives of every leaf
reached.  This module walks the packed arrays that zg_scene_copy_bvh returns (node layout:
lo[DIM][W], hi[DIM][W], child[W] as int32 bits, padded to 64-byte multiples; a child < 0 is a
leaf ~((first << 4) | (count - 1)); primitives in leaf order, triangles a.xyz b.xyz c.xyz id
0 0) with those rules, so that the GPU's count of the same traversal (count_visits 2 of
zg_wos_solve) can be checked against it.  Only tests use it.
"""
from __future__ import annotations

import numpy as np

FLT_MAX = float(np.finfo(np.float32).max)


def node_floats(dim: int, width: int) -> int:
    raw = 2 * dim * width + width
    return ((raw + 15) // 16) * 16


def _closest_on_triangle2(a, b, c, p):
    """Squared distance from p to triangle abc (Ericson, Real-Time Collision Detection 5.1.5)."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        q = a
    else:
        bp = p - b
        d3, d4 = ab @ bp, ac @ bp
        if d3 >= 0 and d4 <= d3:
            q = b
        else:
            vc = d1 * d4 - d3 * d2
            if vc <= 0 and d1 >= 0 and d3 <= 0:
                q = a + ab * (d1 / (d1 - d3))
            else:
                cp = p - c
                d5, d6 = ab @ cp, ac @ cp
                if d6 >= 0 and d5 <= d6:
                    q = c
                else:
                    vb = d5 * d2 - d1 * d6
                    if vb <= 0 and d2 >= 0 and d6 <= 0:
                        q = a + ac * (d2 / (d2 - d6))
                    else:
                        va = d3 * d6 - d5 * d4
                        if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
                            q = b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)))
                        else:
                            den = 1.0 / (va + vb + vc)
                            q = a + ab * (vb * den) + ac * (vc * den)
    d = p - q
    return float(d @ d)


def _closest_on_segment2(a, b, p):
    ab = b - a
    t = float(np.clip(((p - a) @ ab) / max(ab @ ab, 1e-300), 0.0, 1.0))
    d = p - (a + t * ab)
    return float(d @ d)


def _sort_network(k, v):
    """The same compare-exchange network as the device traversal (ties keep their order)."""
    w = len(k)
    if w == 4:
        pairs = [(0, 1), (2, 3), (0, 2), (1, 3), (1, 2)]
    else:
        pairs = [(0, 1), (2, 3), (4, 5), (6, 7), (0, 2), (1, 3), (4, 6), (5, 7), (1, 2), (5, 6),
                 (0, 4), (1, 5), (2, 6), (3, 7), (2, 4), (3, 5), (1, 2), (3, 4), (5, 6)]
    for i, j in pairs:
        if k[j] < k[i]:
            k[i], k[j] = k[j], k[i]
            v[i], v[j] = v[j], v[i]


class PackedBvh:
    def __init__(self, nodes: np.ndarray, prims: np.ndarray, dim: int, width: int):
        self.dim, self.width = dim, width
        nf = node_floats(dim, width)
        self.nodes = np.asarray(nodes, np.float32).reshape(-1, nf)
        self.child = self.nodes.view(np.int32)[:, 2 * dim * width: 2 * dim * width + width]
        self.lo = self.nodes[:, : dim * width].reshape(-1, dim, width).astype(np.float64)
        self.hi = self.nodes[:, dim * width: 2 * dim * width].reshape(-1, dim, width).astype(np.float64)
        pf = 12 if dim == 3 else 4
        P = np.asarray(prims, np.float32).reshape(-1, pf).astype(np.float64)
        if dim == 3:
            self.tri = P[:, :9].reshape(-1, 3, 3)
        else:
            self.seg = P[:, :4].reshape(-1, 2, 2)

    def _prim2(self, i, x):
        if self.dim == 3:
            t = self.tri[i]
            return _closest_on_triangle2(t[0], t[1], t[2], x)
        s = self.seg[i]
        return _closest_on_segment2(s[0], s[1], x)

    def closest(self, x):
        """(squared distance, nodes visited, primitives tested) of the exact query at x."""
        x = np.asarray(x, np.float64)
        W = self.width
        best2 = FLT_MAX
        nodes = prims = 0
        stack = []
        item, item_d2, has = 0, 0.0, len(self.nodes) > 0
        while True:
            if not has:
                if not stack:
                    break
                item, item_d2 = stack.pop()
                if item_d2 > best2:
                    continue
            if item >= 0:
                d = np.maximum(np.maximum(self.lo[item] - x[:, None], x[:, None] - self.hi[item]), 0.0)
                d2 = list((d * d).sum(0))
                ch = list(self.child[item])
                nodes += 1
                for k in range(W):
                    if d2[k] > best2:
                        d2[k] = FLT_MAX
                _sort_network(d2, ch)
                has = d2[0] < FLT_MAX
                item, item_d2 = int(ch[0]), d2[0]
                for k in range(W - 1, 0, -1):
                    if d2[k] < FLT_MAX:
                        stack.append((int(ch[k]), d2[k]))
                if not has or item >= 0:
                    continue          # nothing left here, or the nearest child is an inner node
            # a leaf: its primitives one by one while the leaf box stays within the bound
            code = ~item
            first, cnt = code >> 4, (code & 15) + 1
            while True:
                dd = self._prim2(first, x)
                prims += 1
                if dd < best2:
                    best2 = dd
                if cnt > 1 and item_d2 <= best2:
                    first, cnt = first + 1, cnt - 1
                    continue
                break
            has = False
        return best2, nodes, prims
