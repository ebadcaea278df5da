"""Zone classification and metric sizing (host side).

``classify_zone`` restates ``nature_geom_dom`` (``Fast/FastC/FastC/Metric/nature_geom_dom.for``
:6-560), called by ``allocate_metric`` (``Fast/FastS/FastS/Metric/allocate_metric.cpp:158``):
it decides whether a zone is 3D curvilinear (type 0), 3D with a homogeneous constant-step k
direction (type 1), 3D Cartesian with constant steps (type 2) or 2D (type 3), and sizes the
metric arrays accordingly.  The tests are tolerance based (``souszero = 2e-8``).
The metric values themselves are computed on the GPU (``fm_zone_metric``).
"""
import numpy as np
from . import params as P

SOUSZERO = 2.0e-8


def _vertex_arrays(p, x, y, z):
    nv = (int(p[P.NIJK_XYZ]), int(p[P.NIJK_XYZ + 1]), int(p[P.NIJK_XYZ + 2]))
    return (np.asarray(x).reshape(nv, order="F"), np.asarray(y).reshape(nv, order="F"),
            np.asarray(z).reshape(nv, order="F"))


def degenerate_cells(p, x, y, z):
    """degener(ind) of nature_geom_dom.for:52-160: 1 where all four edges of a cell in one
    direction have zero length (cells collapsed by a boundary condition)."""
    X, Y, Z = _vertex_arrays(p, x, y, z)
    def L(a0, a1):
        return np.sqrt((X[a0] - X[a1]) ** 2 + (Y[a0] - Y[a1]) ** 2 + (Z[a0] - Z[a1]) ** 2)
    s = slice(0, -1); e = slice(1, None)
    # direction i: edges (j,k) (j+1,k) (j+1,k+1) (j,k+1)
    di = ((L((s, s, s), (e, s, s)) <= SOUSZERO) & (L((s, e, s), (e, e, s)) <= SOUSZERO)
          & (L((s, e, e), (e, e, e)) <= SOUSZERO) & (L((s, s, e), (e, s, e)) <= SOUSZERO))
    dj = ((L((s, s, s), (s, e, s)) <= SOUSZERO) & (L((e, s, s), (e, e, s)) <= SOUSZERO)
          & (L((e, s, e), (e, e, e)) <= SOUSZERO) & (L((s, s, e), (s, e, e)) <= SOUSZERO))
    dk = ((L((s, s, s), (s, s, e)) <= SOUSZERO) & (L((e, s, s), (e, s, e)) <= SOUSZERO)
          & (L((e, e, s), (e, e, e)) <= SOUSZERO) & (L((s, e, s), (s, e, e)) <= SOUSZERO))
    return di | dj | dk


def classify_zone(p, x, y, z, lale=0):
    """Fill ITYPZONE, NIJK_MTR(5), NDIMDX_MTR, NEQ_IJ, NEQ_K in ``p`` (in place); return type."""
    nx = [int(v) for v in p[P.NIJK_XYZ:P.NIJK_XYZ + 5]]
    if degenerate_cells(p, x, y, z).any():
        raise ValueError("zone has degenerate (collapsed) cells: not supported")
    if nx[2] == 2:
        typ, neq_ij, neq_k = 3, 2, 0
        mtr = (1, nx[0], 0, nx[3], nx[4]); ndm = nx[0] * nx[1]
    else:
        X, Y, Z = _vertex_arrays(p, x, y, z)
        zcart = np.all(np.abs(Z - Z[0:1, 0:1, :]) < SOUSZERO)
        ycart = np.all(np.abs(Y - Y[0:1, :, 0:1]) < SOUSZERO)
        xcart = np.all(np.abs(X - X[:, 0:1, 0:1]) < SOUSZERO)
        dz = np.diff(Z, axis=2); dy = np.diff(Y, axis=1); dx = np.diff(X, axis=0)
        zcte = np.all(np.abs(dz - dz[0, 0, 0]) < SOUSZERO)
        ycte = np.all(np.abs(dy - dy[0, 0, 0]) < SOUSZERO)
        xcte = np.all(np.abs(dx - dx[0, 0, 0]) < SOUSZERO)
        if xcart and xcte and ycart and ycte and zcart and zcte and lale <= 1:
            typ, neq_ij, neq_k = 2, 1, 1
            mtr = (0, 0, 0, 0, 0); ndm = 1
        else:
            homo = False
            if zcart and zcte:
                xflag = np.all(np.abs(X - X[:, :, 0:1]) < SOUSZERO)
                yflag = np.all(np.abs(Y - Y[:, :, 0:1]) < SOUSZERO)
                homo = bool(xflag and yflag and lale <= 1)
            if homo:
                typ, neq_ij, neq_k = 1, 2, 1
                mtr = (1, nx[0], 0, nx[3], 0); ndm = nx[0] * nx[1]
            else:
                typ, neq_ij, neq_k = 0, 3, 3
                mtr = (1, nx[0], nx[0] * nx[1], nx[3], nx[4]); ndm = nx[0] * nx[1] * nx[2]
    p[P.NIJK_MTR:P.NIJK_MTR + 5] = mtr
    p[P.NDIMDX_MTR] = ndm
    p[P.ITYPZONE] = typ
    p[P.NEQ_IJ] = neq_ij
    p[P.NEQ_K] = neq_k
    return typ


def metric_sizes(p):
    """(size of ti, tj, tk, vol) in float64 elements."""
    ndm = int(p[P.NDIMDX_MTR])
    return ndm * int(p[P.NEQ_IJ]), ndm * int(p[P.NEQ_IJ]), ndm * int(p[P.NEQ_K]), ndm
