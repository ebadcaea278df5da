"""Minimal CGNS/Python tree helpers.

A node is ``[name, value, children, type]`` exactly as in Cassiopee's ``Converter.Internal``
(the tree format the FastS API works on, ``Fast/FastS/FastS/PyTree.py``); Cassiopee itself is
not available, so the handful of accessors the solver API needs are provided here with the
same names and behaviour.
"""
import numpy as np


def createNode(name, ntype, value=None, children=None, parent=None):
    if isinstance(value, (int, np.integer)):
        value = np.array([value], dtype=np.int32)
    elif isinstance(value, (float, np.floating)):
        value = np.array([value], dtype=np.float64)
    elif isinstance(value, str):
        value = np.array(list(value), dtype="c")
    node = [name, value, [] if children is None else children, ntype]
    if parent is not None:
        parent[2].append(node)
    return node


def createChild(parent, name, ntype, value=None, children=None):
    return createNode(name, ntype, value, children, parent)


def createUniqueChild(parent, name, ntype, value=None, children=None):
    n = getNodeFromName1(parent, name)
    if n is None:
        return createChild(parent, name, ntype, value, children)
    new = createNode(name, ntype, value)
    n[1] = new[1]
    n[3] = ntype
    if children is not None:
        n[2] = children
    return n


def getValue(node):
    v = node[1]
    if isinstance(v, np.ndarray) and v.dtype.kind in ("S", "c"):
        return v.tobytes().decode()
    if isinstance(v, np.ndarray) and v.size == 1 and v.dtype.kind in "if":
        return v.ravel()[0].item()
    return v


def setValue(node, value):
    node[1] = createNode("", "", value)[1]


def getNodeFromName1(node, name):
    for c in node[2]:
        if c[0] == name:
            return c
    return None


def getNodeFromName2(node, name):
    for c in node[2]:
        if c[0] == name:
            return c
        for cc in c[2]:
            if cc[0] == name:
                return cc
    return None


def getNodeFromName(node, name):
    if node[0] == name:
        return node
    for c in node[2]:
        r = getNodeFromName(c, name)
        if r is not None:
            return r
    return None


def getNodesFromType1(node, ntype):
    return [c for c in node[2] if c[3] == ntype]


def getNodesFromType2(node, ntype):
    out = []
    for c in node[2]:
        if c[3] == ntype:
            out.append(c)
        for cc in c[2]:
            if cc[3] == ntype:
                out.append(cc)
    return out


def getNodesFromType(node, ntype):
    out = [node] if node[3] == ntype else []
    for c in node[2]:
        out.extend(getNodesFromType(c, ntype))
    return out


def getNodeFromType1(node, ntype):
    for c in node[2]:
        if c[3] == ntype:
            return c
    return None


def getBases(t):
    return getNodesFromType1(t, "CGNSBase_t")


def getZones(t):
    if t[3] == "Zone_t":
        return [t]
    if t[3] == "CGNSBase_t":
        return getNodesFromType1(t, "Zone_t")
    out = []
    for b in getBases(t):
        out.extend(getNodesFromType1(b, "Zone_t"))
    return out


def newCGNSTree():
    t = createNode("CGNSTree", "CGNSTree_t")
    createChild(t, "CGNSLibraryVersion", "CGNSLibraryVersion_t", 3.1)
    return t


def newCGNSBase(name="Base", cellDim=3, physDim=3, parent=None):
    return createNode(name, "CGNSBase_t", np.array([cellDim, physDim], dtype=np.int32), parent=parent)


def newZone(name, ni, nj, nk, parent=None):
    """Structured zone of ni x nj x nk vertices (ghost vertices included)."""
    dims = np.array([[ni, ni - 1, 0], [nj, nj - 1, 0], [nk, nk - 1, 0]], dtype=np.int32)
    z = createNode(name, "Zone_t", dims, parent=parent)
    createChild(z, "ZoneType", "ZoneType_t", "Structured")
    return z


def newBC(name, bctype, prange, parent):
    bc = createChild(parent, name, "BC_t", bctype)
    createChild(bc, "PointRange", "IndexRange_t", np.asarray(prange, dtype=np.int32).reshape(3, 2))
    return bc
