"""Minimal CGNS/Python tree helpers.

A node is ``[name, value, children, type]`` exactly as in Cassiopee's ``Converter.Internal``
(the tree format the FastS API works on, ``Fast/FastS/FastS/PyTree.py``); Cassiopee itself is
not available, so the handful of accessors the solver API needs are provided here with the
same names and behaviour.
"""
import numpy as np

# The solver API (fast_amd.PyTree) keeps the flow fields on the GPU between calls and refreshes
# the tree's arrays when code outside the API next reads the tree through these accessors: it
# installs _access_hook, and wraps its own tree walks in quiet() so that they do not trigger it.
_access_hook = None
_quiet_depth = 0


def _touch():
    if _access_hook is not None and _quiet_depth == 0:
        _access_hook()


class quiet:
    """Context manager: accessor calls inside it do not trigger the access hook."""

    def __enter__(self):
        global _quiet_depth
        _quiet_depth += 1
        return self

    def __exit__(self, *exc):
        global _quiet_depth
        _quiet_depth -= 1
        return False


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
    _touch()
    v = node[1]
    if isinstance(v, np.ndarray) and v.dtype.kind in ("S", "c"):
        return v.tobytes().decode()
    if isinstance(v, np.ndarray) and v.size == 1 and v.dtype.kind in "if":
        return v.ravel()[0].item()
    return v


def setValue(node, value):
    node[1] = createNode("", "", value)[1]


def getNodeFromName1(node, name):
    _touch()
    for c in node[2]:
        if c[0] == name:
            return c
    return None


def getNodeFromName2(node, name):
    _touch()
    for c in node[2]:
        if c[0] == name:
            return c
        for cc in c[2]:
            if cc[0] == name:
                return cc
    return None


def getNodeFromName(node, name):
    _touch()
    return _node_from_name(node, name)


def _node_from_name(node, name):
    if node[0] == name:
        return node
    for c in node[2]:
        r = _node_from_name(c, name)
        if r is not None:
            return r
    return None


def getNodesFromType1(node, ntype):
    _touch()
    return [c for c in node[2] if c[3] == ntype]


def getNodesFromType2(node, ntype):
    _touch()
    out = []
    for c in node[2]:
        if c[3] == ntype:
            out.append(c)
        for cc in c[2]:
            if cc[3] == ntype:
                out.append(cc)
    return out


def getNodesFromType(node, ntype):
    _touch()
    return _nodes_from_type(node, ntype)


def _nodes_from_type(node, ntype):
    out = [node] if node[3] == ntype else []
    for c in node[2]:
        out.extend(_nodes_from_type(c, ntype))
    return out


def getNodeFromType(node, ntype):
    _touch()
    r = _nodes_from_type(node, ntype)
    return r[0] if r else None


getNodeByName1 = getNodeFromName1


def getNodeFromPath(t, path):
    """The node at path "Base/Zone/..." below t (a leading "/" or the root's name is allowed)."""
    _touch()
    parts = [q for q in path.split("/") if q]
    if parts and parts[0] == t[0]:
        parts = parts[1:]
    node = t
    for q in parts:
        node = next((c for c in node[2] if c[0] == q), None)
        if node is None:
            return None
    return node


def getZonePaths(t):
    _touch()
    return ["%s/%s" % (b[0], z[0]) for b in getNodesFromType1(t, "CGNSBase_t")
            for z in getNodesFromType1(b, "Zone_t")]


def getPathsFromType(t, ntype):
    """Paths (from below the root) of the nodes of type ntype."""
    _touch()
    out = []

    def walk(node, prefix):
        for c in node[2]:
            p = c[0] if not prefix else prefix + "/" + c[0]
            if c[3] == ntype:
                out.append(p)
            walk(c, p)
    walk(t, "")
    return out


def getNodeFromType1(node, ntype):
    _touch()
    for c in node[2]:
        if c[3] == ntype:
            return c
    return None


def getBases(t):
    _touch()
    return getNodesFromType1(t, "CGNSBase_t")


def getZones(t):
    _touch()
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
