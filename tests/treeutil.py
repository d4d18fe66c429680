"""Canonical form of a BatchLayout quadtree, for comparing trees built different ways.

The reference's tree (BarnesHut.h:135-220) keeps single-child chains: a cell whose
points all fall in one quadrant gets one internal child with the same point set.  A
chain shares one centroid (an unweighted mean over one child is that child's centroid),
and the traversal (BarnesHut.h:69-86) accepts somewhere in the chain exactly when the
lowest (smallest-diameter) node of the chain is accepted.  The canonical form therefore
replaces every chain by its lowest node and lists, in pre-order, each node as
(diameter, cx, cy) followed by its non-empty quadrants as ("L", vertex) or a subtree.
"""
from __future__ import annotations


def _canon(children, diam, cx, cy, v):
    # collapse chains: exactly one non-empty quadrant and it is internal
    while True:
        nonempty = [(q, c) for q, c in enumerate(children[v]) if c is not None]
        if len(nonempty) == 1 and nonempty[0][1][0] == "N":
            v = nonempty[0][1][1]
            continue
        break
    out = [("node", float(diam[v]).hex(), float(cx[v]).hex(), float(cy[v]).hex())]
    for q, c in enumerate(children[v]):
        if c is None:
            continue
        if c[0] == "L":
            out.append((q, "L", int(c[1])))
        else:
            out.append((q, _canon(children, diam, cx, cy, c[1])))
    return tuple(out)


def canonical_from_reference(d) -> tuple:
    """From a bh_*.npz fixture: pre-order nodes (id, parent, quad) plus leaves (parent, quad, point)."""
    m = len(d["node_id"])
    children = [[None] * 4 for _ in range(m)]
    for v in range(m):
        p, q = int(d["node_parent"][v]), int(d["node_quad"][v])
        if p >= 0:
            children[p][q] = ("N", v)
    for p, q, pt in d["leaf"]:
        children[int(p)][int(q)] = ("L", int(pt))
    return _canon(children, d["node_diam"], d["node_cx"], d["node_cy"], 0)


def canonical_from_child_array(child, diam, cx, cy, root) -> tuple:
    """From child[m,4] arrays (>=0 node, -1 empty, <=-2 leaf vertex -c-2)."""
    m = len(diam)
    children = [[None] * 4 for _ in range(m)]
    for v in range(m):
        for q in range(4):
            c = int(child[v][q])
            if c >= 0:
                children[v][q] = ("N", c)
            elif c <= -2:
                children[v][q] = ("L", -c - 2)
    return _canon(children, diam, cx, cy, root)
