"""FastS solver API over CGNS/Python trees, running on the MI355X.

Drop-in for the FastS compute path of ``Fast/FastS/FastS/PyTree.py``:

* ``warmup(t, tc=None, ...) -> (t, tc, metrics)``       PyTree.py:341-545
* ``_compute(t, metrics, nitrun, tc=None, ..., NIT=1)``  PyTree.py:42-236
* ``allocate_metric(t)`` / ``_init_metric(t, metrics, hook)``  PyTree.py:267-335
* ``display_temporal_criteria(t, metrics, nitrun, format=None, ...)``  PyTree.py:2031-2076
* ``_applyBC(t, metrics, hook1, nstep, var)``           PyTree.py:1155
* ``_fillGhostcells(zones, tc, metrics, ...)``           transfers + ``_applyBC``; zones joined by
  CGNS GridConnectivity1to1 nodes (injection raccords, FastC/Com/setInterpTransfersFast.cpp)
* ``_computeEnstrophy(t, metrics)``                      PyTree.py:1878
* ``createStressNodes(t, BC, windows)`` / ``_computeStress(t, teff, metrics, xyz_ref)``
  PyTree.py:2277-2566 (wall efforts, bceffort.for)
* ``_createConvergenceHistory(t, nrecs)``               PyTree.py:2152
* numerics setters ``_setNum2Base`` / ``_setNum2Zones``  (FastC/PyTree.py:115-137)

The tree holds the fields exactly as the reference's does (FlowSolution#Centers Density,
VelocityX/Y/Z, Temperature and their _M1/_P1 time levels, ghost cells included).  The state
lives on the GPU between calls: ``_compute`` leaves the tree arrays as they were and marks them
stale; the next time code outside this API reads the tree through ``fast_amd.Internal`` (or
calls ``_syncTree``, ``saveTree``) the three levels are copied back into the tree arrays in
place.  Once the tree has been handed out that way it is the source of truth again: the next
``_compute`` (or BC fill, stress, enstrophy) first copies the tree's levels to the GPU, so edits
made to the tree arrays take effect.  A loop of ``_compute`` calls that does not read the tree
therefore moves no field over PCIe.  Arrays fetched before a ``_compute`` must be fetched again
(or ``_syncTree`` called) to see the new state.  Like the reference's module-global
``FastC.HOOK``, one solver state lives per process.
"""
import functools
import os
import sys

import numpy as np

from . import Internal
from . import params as P
from .zone import HostZone

VARS = ["Density", "VelocityX", "VelocityY", "VelocityZ", "Temperature"]
HOOK = {}


def _api(f):
    """Solver API functions walk the tree without triggering the tree refresh."""
    @functools.wraps(f)
    def wrapper(*args, **kw):
        with Internal.quiet():
            return f(*args, **kw)
    return wrapper


def _pull_levels():
    """Copy the three time levels of every zone from the GPU into the tree arrays (in place)."""
    from . import _lib
    ctx = HOOK["ctx"]
    with Internal.quiet():
        for (z, hz), zid in zip(HOOK["zones"], HOOK["ids"]):
            for sfx, f, a in (("", _lib.FIELD_RO, hz.ro), ("_M1", _lib.FIELD_RO_M1, hz.ro_m1),
                              ("_P1", _lib.FIELD_RO_P1, hz.ro_p1)):
                ctx.get(zid, f, a)
                _level_to_tree(z, hz, sfx, a)
    HOOK["stale"] = False
    HOOK["exposed"] = True


def _on_tree_access():
    if HOOK.get("stale"):
        _pull_levels()


def _device_current():
    """Before GPU work: if the tree has been handed out since the last GPU update, its arrays
    are the state (they may have been edited): copy them to the GPU."""
    from . import _lib
    if not HOOK.get("exposed"):
        return
    ctx = HOOK["ctx"]
    for (z, hz), zid in zip(HOOK["zones"], HOOK["ids"]):
        for sfx, f, a in (("", _lib.FIELD_RO, hz.ro), ("_M1", _lib.FIELD_RO_M1, hz.ro_m1),
                          ("_P1", _lib.FIELD_RO_P1, hz.ro_p1)):
            _tree_to_level(z, hz, sfx, a)
            ctx.set(zid, f, a)
    HOOK["exposed"] = False


def _gpu_updated():
    """After GPU work that changed the state: the tree arrays are stale until next read."""
    HOOK["stale"] = True
    HOOK["exposed"] = False
    Internal._access_hook = _on_tree_access


def _syncTree(t=None):
    """Copy the current state of every zone (levels N, N-1, N+1) into the tree arrays now."""
    if HOOK.get("stale"):
        _pull_levels()


syncTree = _syncTree
_MODELS = {"NSLaminar": "lamin", "Euler": "euler"}
_STATE_KEYS = ("gamma", "cv", "ro", "p", "u", "T", "mach", "reynolds", "prandtl", "mus", "ts",
               "cs", "vx", "vy", "vz")


# ---------------------------------------------------------------------------------------
# numerics (FastC._setNum2Base / _setNum2Zones)
def _set_define(node, num):
    d = Internal.createUniqueChild(node, ".Solver#define", "UserDefinedData_t")
    for k, v in num.items():
        Internal.createUniqueChild(d, k, "DataArray_t", v)


@_api
def _setNum2Base(t, numb):
    for b in Internal.getBases(t):
        _set_define(b, numb)


@_api
def _setNum2Zones(t, numz):
    for z in Internal.getZones(t):
        _set_define(z, numz)


def _get_define(node):
    out = {}
    d = Internal.getNodeFromName1(node, ".Solver#define")
    if d is not None:
        for c in d[2]:
            out[c[0]] = Internal.getValue(c)
    return out


def _numerics(base, zone):
    n = _get_define(base)
    n.update(_get_define(zone))
    n.setdefault("temporal_scheme", "implicit")
    if n.get("scheme") in ("roe_kap", "roe_min", "roe_nul"):
        n["slope"] = {"roe_kap": "o3", "roe_min": "minmod", "roe_nul": "o1"}[n["scheme"]]
    keys = set(P.Numerics.DEFAULTS)
    return P.Numerics(**{k: v for k, v in n.items() if k in keys})


# ---------------------------------------------------------------------------------------
# reference state (Cassiopee's C.addState is not available: stored as DataArray_t children)
def _addState(t, state, model="NSLaminar"):
    """ReferenceState and FlowEquationSet/GoverningEquations of every base (the nodes
    FastC._buildOwnData reads, Fast/FastC/FastC/PyTree.py:1452-1490)."""
    if model not in _MODELS:
        raise NotImplementedError("governing equations %r are not supported" % model)
    for b in Internal.getBases(t):
        rs = Internal.createUniqueChild(b, "ReferenceState", "ReferenceState_t")
        for k in _STATE_KEYS:
            Internal.createUniqueChild(rs, k, "DataArray_t", float(state[k]))
        fe = Internal.createUniqueChild(b, "FlowEquationSet", "FlowEquationSet_t")
        Internal.createUniqueChild(fe, "GoverningEquations", "GoverningEquations_t", model)


def _get_state(base, zone):
    rs = Internal.getNodeFromName1(zone, "ReferenceState") or Internal.getNodeFromName1(base, "ReferenceState")
    if rs is None:
        raise ValueError("ReferenceState missing (use fast_amd.PyTree._addState)")
    return {k: Internal.getValue(Internal.getNodeFromName1(rs, k)) for k in _STATE_KEYS}


# ---------------------------------------------------------------------------------------
# zones <-> tree
def newZoneFromHost(hz, base, bcs=None):
    """Add a zone built from a HostZone to a base: coordinates, centre fields (N level),
    ZoneBC with PointRange windows in the ghost-inclusive vertex numbering."""
    p = hz.param_int
    nv = (int(p[P.NIJK_XYZ]), int(p[P.NIJK_XYZ + 1]), int(p[P.NIJK_XYZ + 2]))
    nc = (int(p[P.NIJK]), int(p[P.NIJK + 1]), int(p[P.NIJK + 2]))
    z = Internal.newZone(hz.name, nv[0], nv[1], nv[2], parent=base)
    g = Internal.createChild(z, "GridCoordinates", "GridCoordinates_t")
    for n, a in (("CoordinateX", hz.x), ("CoordinateY", hz.y), ("CoordinateZ", hz.z)):
        Internal.createChild(g, n, "DataArray_t", a.reshape(nv, order="F"))
    fs = Internal.createChild(z, "FlowSolution#Centers", "FlowSolution_t")
    Internal.createChild(fs, "GridLocation", "GridLocation_t", "CellCenter")
    nd = hz.nd
    for i, v in enumerate(VARS):
        Internal.createChild(fs, v, "DataArray_t", hz.ro[i * nd:(i + 1) * nd].reshape(nc, order="F").copy(order="F"))
    zbc = Internal.createChild(z, "ZoneBC", "ZoneBC_t")
    pt = int(p[P.PT_BC])
    nb = int(p[pt]) if pt > 0 else 0
    inv = {}
    for k, v in P.BC_TAGS.items():          # the first name of each type (BCWallInviscid, not its alias)
        inv.setdefault(v, k)
    for n in range(nb):
        q = int(p[pt + 1 + n])
        typ, idir = int(p[q]), int(p[q + 1])
        fen = [int(v) for v in p[q + 2:q + 8]]
        rng = _fen_to_pointrange(p, idir, fen)
        Internal.newBC("BC%d_%s" % (n, ["imin", "imax", "jmin", "jmax", "kmin", "kmax"][idir - 1]),
                       inv[typ], rng, zbc)
    return z


def _fen_to_pointrange(p, idir, fen):
    """Inverse of getRange (Fast/FastC/FastC/getRange.cpp:57-110) for full-face windows."""
    nf, nfk = int(p[P.NIJK + 3]), int(p[P.NIJK + 4])
    nv = [int(p[P.NIJK_XYZ]), int(p[P.NIJK_XYZ + 1]), int(p[P.NIJK_XYZ + 2])]
    rng = []
    for d in range(3):
        g = nf if d < 2 else nfk
        lo, hi = fen[2 * d] + g, fen[2 * d + 1] + g + 1
        rng.append([lo, hi])
    d = (idir - 1) // 2
    v = 1 if idir % 2 == 1 else nv[d]
    rng[d] = [v, v]
    return rng


def _addMatch(z, name, donor, rng_r, rng_d, transform=(1, 2, 3)):
    """A CGNS GridConnectivity1to1 node (what Cassiopee's connectMatch writes) under
    ZoneGridConnectivity: receiver window rng_r of z matches window rng_d of zone ``donor``."""
    zgc = Internal.getNodeFromName1(z, "ZoneGridConnectivity")
    if zgc is None:
        zgc = Internal.createChild(z, "ZoneGridConnectivity", "ZoneGridConnectivity_t")
    gc = Internal.createChild(zgc, name, "GridConnectivity1to1_t", donor)
    Internal.createChild(gc, "PointRange", "IndexRange_t", np.array(rng_r, dtype=np.int32))
    Internal.createChild(gc, "PointRangeDonor", "IndexRange_t", np.array(rng_d, dtype=np.int32))
    Internal.createChild(gc, "Transform", '"int[IndexDimension]"', np.array(transform, dtype=np.int32))
    return gc


def _matches(zones):
    """Injection raccords of every GridConnectivity1to1 of the zones (list of (tree zone,
    HostZone)); the donor is found by name."""
    from . import connect
    index = {z[0]: n for n, (z, hz) in enumerate(zones)}
    racs = []
    for n, (z, hz) in enumerate(zones):
        zgc = Internal.getNodeFromName1(z, "ZoneGridConnectivity")
        if zgc is None:
            continue
        for gc in Internal.getNodesFromType1(zgc, "GridConnectivity1to1_t"):
            donor = Internal.getValue(gc)
            if donor not in index:
                raise ValueError("GridConnectivity1to1 %s: unknown donor zone %r" % (gc[0], donor))
            d = index[donor]
            tr = Internal.getNodeFromName1(gc, "Transform")
            transform = (1, 2, 3) if tr is None else tuple(int(v) for v in np.asarray(tr[1]).ravel())
            rng_r = np.asarray(Internal.getNodeFromName1(gc, "PointRange")[1])
            rng_d = np.asarray(Internal.getNodeFromName1(gc, "PointRangeDonor")[1])
            rcv, dnr = connect.from_pointranges(hz.param_int, rng_r, zones[d][1].param_int, rng_d, transform)
            racs.append(connect.Raccord(n, d, rcv, dnr))
    return racs


def _zone_bcs(z, p):
    bcs = []
    zbc = Internal.getNodeFromName1(z, "ZoneBC")
    if zbc is None:
        return bcs
    for bc in Internal.getNodesFromType1(zbc, "BC_t"):
        btype = Internal.getValue(bc)
        rng = Internal.getValue(Internal.getNodeFromType1(bc, "IndexRange_t"))
        in_bc = [int(rng[0][0]), int(rng[0][1]), int(rng[1][0]), int(rng[1][1]), int(rng[2][0]), int(rng[2][1])]
        idir, fen = P.get_range(in_bc, p)
        bcs.append((btype, (idir, fen), None))
    return bcs


def _host_zone(base, z):
    dims = z[1]
    nv = [int(dims[0][0]), int(dims[1][0]), int(dims[2][0])]
    ni, nj, nk = nv[0] - 5, nv[1] - 5, nv[2] - 5
    num = _numerics(base, z)
    st = _get_state(base, z)
    g = Internal.getNodeFromName1(z, "GridCoordinates")
    x, y, zz = (np.asarray(Internal.getNodeFromName1(g, n)[1]).ravel(order="F").astype(np.float64)
                for n in ("CoordinateX", "CoordinateY", "CoordinateZ"))
    p0 = P.build_param_int(ni, nj, nk, num)
    bcs = _zone_bcs(z, p0)
    fe = Internal.getNodeFromName1(z, "FlowEquationSet") or Internal.getNodeFromName1(base, "FlowEquationSet")
    ge = Internal.getNodeFromName1(fe, "GoverningEquations") if fe is not None else None
    model = Internal.getValue(ge) if ge is not None else "NSLaminar"
    if model not in _MODELS:
        raise NotImplementedError("governing equations %r are not supported" % model)
    hz = HostZone(z[0], ni, nj, nk, num, st, x, y, zz, bcs=bcs, flow=_MODELS[model])
    return hz


def _fields(z, suffix=""):
    fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
    return [Internal.getNodeFromName1(fs, v + suffix) for v in VARS]


def _tree_to_level(z, hz, suffix, out):
    nd = hz.nd
    for i, n in enumerate(_fields(z, suffix)):
        out[i * nd:(i + 1) * nd] = np.asarray(n[1]).ravel(order="F")


def _level_to_tree(z, hz, suffix, a):
    nd = hz.nd
    for i, n in enumerate(_fields(z, suffix)):
        np.copyto(n[1], a[i * nd:(i + 1) * nd].reshape(n[1].shape, order="F"))


# ---------------------------------------------------------------------------------------
@_api
def allocate_metric(t, verbose=0):
    """Classify every zone, build its descriptor vectors, register it on the GPU and return
    the list of metric arrays (ti|tj|tk|vol per zone)."""
    from . import _lib
    ctx = HOOK.get("ctx")
    if ctx is None:
        ctx = _lib.Context(0)
        HOOK.clear()
        HOOK.update(ctx=ctx, zones=[], ids=[])
    metrics = []
    for b in Internal.getBases(t):
        for z in Internal.getNodesFromType1(b, "Zone_t"):
            hz = _host_zone(b, z)
            zid = ctx.add_zone(hz)
            HOOK["zones"].append((z, hz))
            HOOK["ids"].append(zid)
            o = Internal.createUniqueChild(z, ".Solver#ownData", "UserDefinedData_t")
            Internal.createUniqueChild(o, "Parameter_int", "DataArray_t", hz.param_int)
            Internal.createUniqueChild(o, "Parameter_real", "DataArray_t", hz.param_real)
            if verbose:
                print("typezone: %s, %s %s" % (["3D curvilinear", "3D homogeneous k", "3D Cartesian", "2D"][hz.zone_type],
                                              z[0], hz.dims))
            metrics.append([hz.metric])
    return metrics


@_api
def _init_metric(t, metrics, hook=None):
    """Face vectors and volumes (computed on the GPU when the zone was registered)."""
    from . import _lib
    ctx = HOOK["ctx"]
    for (z, hz), zid, m in zip(HOOK["zones"], HOOK["ids"], metrics):
        ctx.get(zid, _lib.FIELD_METRIC, m[0])


@_api
def warmup(t, tc=None, graph=None, infos_ale=None, Adjoint=False, tmy=None, list_graph=None,
           Padding=None, verbose=0):
    """Build the solver state: metrics, time levels _M1/_P1, ghost cells.  -> (t, tc, metrics)

    Zones are joined by the injection raccords of ``tc``, a Cassiopee connectivity tree (the
    ``ID_*`` subregions of its donor zones, PyTree.py:492-499), or, without ``tc``, by the
    GridConnectivity1to1 nodes of t; ``tc`` may also be a list of fast_amd.connect.Raccord.
    As in the reference, HOOK['param_int_tc'] / HOOK['param_real_tc'] then hold the raccords
    in their flattened form."""
    from . import _lib, connect
    racs_given = tc is not None and isinstance(tc, list) and all(hasattr(r, "rcv") for r in tc)
    if tc is not None and not racs_given and not Internal.getZones(tc):
        raise TypeError("tc must be a connectivity tree or a list of fast_amd.connect.Raccord")
    HOOK.clear()
    metrics = allocate_metric(t, verbose)
    _init_metric(t, metrics, HOOK)
    ctx = HOOK["ctx"]
    if tc is None:
        racs = _matches(HOOK["zones"])
    elif racs_given:
        racs = list(tc)
    else:
        index = {z[0]: n for n, (z, hz) in enumerate(HOOK["zones"])}
        racs = connect.from_tc(tc, index, [hz.nd for z, hz in HOOK["zones"]])
    ctx.set_transfers(racs)
    HOOK["raccords"] = racs
    HOOK["param_int_tc"], HOOK["param_real_tc"] = connect.pack_tc(racs)
    first_it = 1
    has_m1s = []
    for (z, hz), zid in zip(HOOK["zones"], HOOK["ids"]):
        fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
        _tree_to_level(z, hz, "", hz.ro)
        has_m1 = Internal.getNodeFromName1(fs, "Density_M1") is not None
        has_m1s.append(has_m1)
        if has_m1:
            _tree_to_level(z, hz, "_M1", hz.ro_m1)
        else:
            first_it = 0       # createPrimVars sets FIRST_IT = 0 when _M1 had to be created
            hz.ro_m1[:] = hz.ro
        ctx.set(zid, _lib.FIELD_RO, hz.ro)
    # ghost cells: transfers between zones, then the physical boundary conditions
    ctx.transfer(0)
    for (z, hz), zid, has_m1 in zip(HOOK["zones"], HOOK["ids"], has_m1s):
        fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
        ctx.apply_bc(zid, 0)
        ctx.get(zid, _lib.FIELD_RO, hz.ro)
        if not has_m1:
            hz.ro_m1[:] = hz.ro
        hz.ro_p1[:] = hz.ro
        ctx.set(zid, _lib.FIELD_RO_M1, hz.ro_m1)
        ctx.set(zid, _lib.FIELD_RO_P1, hz.ro_p1)
        for sfx, a in (("", hz.ro), ("_M1", hz.ro_m1), ("_P1", hz.ro_p1)):
            nd = hz.nd
            nc = z[1][:, 1]
            for i, v in enumerate(VARS):
                arr = a[i * nd:(i + 1) * nd].reshape(tuple(int(c) for c in nc), order="F").copy(order="F")
                Internal.createUniqueChild(fs, v + sfx, "DataArray_t", arr)
    if Adjoint:
        # FastC/PyTree.py:782-794 (createPrimVars with Adjoint): the dJ/dW fields, zero
        for z, hz in HOOK["zones"]:
            fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
            shape = tuple(int(c) for c in z[1][:, 1])
            for v in ADJ_VARS["cd"] + ADJ_VARS["cl"]:
                if Internal.getNodeFromName1(fs, v) is None:
                    Internal.createUniqueChild(fs, v, "DataArray_t", np.zeros(shape, order="F"))
    HOOK["first_it"] = first_it
    HOOK["nitrun_done"] = 0
    # the tree now holds the state and is in the caller's hands
    HOOK["stale"] = False
    HOOK["exposed"] = True
    Internal._access_hook = _on_tree_access
    return t, tc, metrics


@_api
def _compute(t, metrics, nitrun, tc=None, graph=None, tc2=None, graph2=None, layer="c", NIT=1,
             ucData=None, vtune=None):
    """NIT time steps (all sub-iterations) of every zone on the GPU (computePT.cpp:44-721).  The
    state stays on the GPU; the tree arrays are refreshed when the tree is next read."""
    import time
    from . import _lib
    ctx = HOOK["ctx"]
    _device_current()
    base = Internal.getBases(t)[0]
    bdef = _get_define(base)
    modulo_verif = int(bdef.get("modulo_verif", 200))
    hz0 = HOOK["zones"][0][1]
    nssiter = int(hz0.param_int[P.NSSITER])
    t0 = time.time()
    ctx.compute(nitrun, 1, nssiter, nit_c=NIT, modulo_verif=modulo_verif, first_it=HOOK["first_it"])
    if vtune is not None:
        ctx.sync()
    t1 = time.time()
    HOOK["first_it"] = 1
    HOOK["nitrun_done"] = nitrun + NIT
    _gpu_updated()
    # CFL max / min / mean of the last verification step (computePT.cpp:629-645), stored in
    # .Solver#ownData/CFL_minmaxmoy when a verification step ran
    verified = modulo_verif > 0 and any((nitrun * NIT + it) % modulo_verif == 0 for it in range(NIT))
    if verified:
        for (z, hz), zid in zip(HOOK["zones"], HOOK["ids"]):
            try:
                cfl = ctx.cfl(zid)
            except _lib.FastMIError:
                continue
            o = Internal.createUniqueChild(z, ".Solver#ownData", "UserDefinedData_t")
            Internal.createUniqueChild(o, "CFL_minmaxmoy", "DataArray_t", cfl)
    if vtune is not None:
        return t1 - t0, 0.0
    return None


@_api
def _fillGhostcells(zones, tc, metrics, timelevel_target, vars, nstep, omp_mode=0, hook1=None,
                    nitmax=1, rk=1, exploc=0, num_passage=1):
    """Transfers between zones then physical boundary conditions on the N level of every zone
    (Fast/FastS/FastS/PyTree.py:1162-1230 _fillGhostcells: setInterpTransfers then _applyBC)."""
    ctx = HOOK["ctx"]
    _device_current()
    ctx.transfer(0)
    for (z, hz), zid in zip(HOOK["zones"], HOOK["ids"]):
        ctx.apply_bc(zid, 0, nstep)
    _gpu_updated()


@_api
def _applyBC(t, metrics, hook1=None, nstep=1, var="Density"):
    """Physical boundary conditions on the N level of every zone (BC/applyBC.cpp:33-255)."""
    ctx = HOOK["ctx"]
    _device_current()
    for (z, hz), zid in zip(HOOK["zones"], HOOK["ids"]):
        ctx.apply_bc(zid, 0, nstep)
    _gpu_updated()


# ---------------------------------------------------------------------------------------
# Convergence monitor (display_ss_iteration.cpp, dpssiter.for, conv2pytree.for)
_FORMATS = {None: 1, "double": 0, "flush": -1, "ascii": -2, "store": 3}
_CUT0X = 1.e-30


def _log10c(v):
    return np.log10(np.maximum(v, _CUT0X))


def _fortran_e(v, w=15, d=8):
    """Fortran Ew.d: 0.ddddddddE+xx, right-justified in w columns."""
    if v == 0.0 or not np.isfinite(v):
        m, e = (0.0, 0) if v == 0.0 else (v, 0)
        body = ("%.*f" % (d, m)) if np.isfinite(v) else str(v)
        return (body + "E+00").rjust(w)
    e = int(np.floor(np.log10(abs(v)))) + 1
    m = abs(v) / 10.0 ** e
    if round(m, d) >= 1.0:
        m /= 10.0
        e += 1
    body = "%s0.%0*dE%s%02d" % ("-" if v < 0 else "", d, int(round(m * 10 ** d)), "-" if e < 0 else "+", abs(e))
    return body.rjust(w)


def _dpssiter(name, nitrun, rdm, lft, iverb):
    """dpssiter.for: print the log10 residuals of the recorded sub-iterations (first one always,
    every one when verbose or when Linf grew) and the difference last - first; return the 4 neq
    values conv2pytree records (log10 of the first L2 and Linf, log10 differences last - first)."""
    neq = 5
    nrec = rdm.shape[0]
    l2, li = rdm[:, :neq], rdm[:, neq:]
    iverbs = iverb
    if np.any(li[nrec - 1] > li[0]):
        iverbs = 2
    if lft < 3:
        it_reelle = list(range(1, nrec)) + [nrec]
        fmt = (lambda v: "  %14.9f" % v) if lft == 0 else (lambda v: "  %7.2f" % v)
        for i in range(nrec):
            if i == 0 or iverbs > 1:
                print(" Res_L2 (zone=%s, ssiter=%3d %3d)%s" % (name, it_reelle[i], i + 1, "".join(fmt(v) for v in _log10c(l2[i]))))
                print(" Res_Loo(zone=%s, ssiter=%3d %3d)%s" % (name, it_reelle[i], i + 1, "".join(fmt(v) for v in _log10c(li[i]))))
        d2 = _log10c(l2[nrec - 1]) - _log10c(l2[0])
        di = _log10c(li[nrec - 1]) - _log10c(li[0])
        print(" Res_L2 (zone=%s, Diff Ssi 1-%3d)%s" % (name, nrec, "".join(fmt(v) for v in d2)))
        print(" Res_Loo(zone=%s, Diff Ssi 1-%3d)%s" % (name, nrec, "".join(fmt(v) for v in di)))
        if d2.max() > 0:
            print("Warning convergence L2  Newton: dom=%s Pas de temps=%6d divergence=%12.5f" % (name, nitrun, d2.max()))
        if di.max() > 0:
            print("Warning convergence Loo Newton: dom=%s Pas de temps=%6d divergence=%12.5f" % (name, nitrun, di.max()))
    cvg = np.concatenate([_log10c(l2[0]), _log10c(li[0]), _log10c(l2[nrec - 1]) - _log10c(l2[0]),
                          _log10c(li[nrec - 1]) - _log10c(li[0])])
    return cvg


def _history_nodes(h):
    g = lambda n: Internal.getNodeFromName1(h, n)[1]
    return g("IterationNumber"), g("RSD_L2"), g("RSD_oo"), g("RSD_L2_diff"), g("RSD_oo_diff")


def _conv2pytree(z, cvg, nitrun, lft, nd):
    """conv2pytree.for: append one record (iteration, log10 L2 / Linf of the first recorded
    sub-iteration and their drop over the step) to the zone's ZoneConvergenceHistory, whose
    value is the number of records; format "flush" appends the records to residuals.bin
    (Fortran unformatted) and empties the history, "ascii" appends them to residuals.dat.
    A full history records nothing more (the reference writes past the end of its arrays)."""
    h = Internal.getNodeFromName1(z, "ZoneConvergenceHistory")
    if h is None:
        return
    neq = 5
    itnum, rl2, roo, rl2d, rood = _history_nodes(h)
    nrec = itnum.size
    last = int(np.asarray(h[1]).ravel()[0])
    if lft < 0:
        name40 = z[0][:40].ljust(40)
        if lft == -1:
            import struct
            payload = name40.encode() + struct.pack("<ii", neq, last)
            for i in range(last):
                payload += struct.pack("<i", int(itnum[i]))
                for j in range(neq):
                    payload += struct.pack("<dd", rl2[i * neq + j], roo[i * neq + j])
            with open("residuals.bin", "ab") as f:
                f.write(struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload)))
            last = 0
        else:
            with open("residuals.dat", "a") as f:
                if nd == 0:
                    f.write('VARIABLES="It" "Ro_L2" "RoU_L2" "RoV_L2" "RoW_L2" "RoE_L2" '
                            '"Ro_oo" "RoU_oo" "RoV_oo" "RoW_oo" "RoE_oo"\n')
                f.write('ZONE T="%s"\n' % z[0][:40])
                for i in range(last):
                    vals = []
                    for j in range(neq):
                        vals += [rl2[i * neq + j], roo[i * neq + j]]
                    f.write("%7d%s\n" % (int(itnum[i]), "".join(" " + _fortran_e(v) for v in vals)))
    if lft >= 0 and last < nrec:
        itnum[last] = nitrun
        sl = slice(last * neq, (last + 1) * neq)
        rl2[sl] = cvg[0:neq]; roo[sl] = cvg[neq:2 * neq]
        rl2d[sl] = cvg[2 * neq:3 * neq]; rood[sl] = cvg[3 * neq:4 * neq]
        last += 1
    h[1] = np.array([last], dtype=np.asarray(h[1]).dtype)


@_api
def display_temporal_criteria(t, metrics, nitrun, format=None, gmres=None, verbose="firstlast",
                              stopAtNan=True):
    """Convergence monitor of the last verified step (PyTree.py:2039-2076 ->
    display_ss_iteration.cpp:29-191): per zone, the L2 and Linf residuals of every recorded
    sub-iteration (divided by DTC), printed as dpssiter.for does (format None: short, "double":
    long), recorded in the zone's ZoneConvergenceHistory if it exists (conv2pytree.for;
    "store" records without printing, "flush" / "ascii" also write the records to
    residuals.bin / residuals.dat).  Returns None, or with ``gmres`` given
    {"L2": sum of the L2 norms over zones, equations and sub-iterations, "Loo": their maximum
    Linf}.  stopAtNan: exit(1) when a zone's density is not finite."""
    from . import _lib
    if format not in _FORMATS:
        raise ValueError("format must be one of %s" % sorted(k for k in _FORMATS if k))
    lft = _FORMATS[format]
    iverb = 0 if verbose == "firstlast" else 2
    ctx = HOOK["ctx"]
    res_l2, res_li = 0.0, -1.e15
    for nd, ((z, hz), zid) in enumerate(zip(HOOK["zones"], HOOK["ids"])):
        cvg = None
        implicit = int(hz.param_int[P.ITYPCP]) <= 1
        if lft >= 0 and implicit:
            o = Internal.getNodeFromName1(z, ".Solver#ownData")
            c = Internal.getNodeFromName1(o, "CFL_minmaxmoy") if o is not None else None
            if c is not None and lft < 3:
                cf = np.asarray(c[1]).ravel()
                print("CFL_MOY_MAX_MIN(Zone= %32.32s, nitrun=%d) %f    %f    %f " % (z[0], nitrun, cf[2], cf[0], cf[1]))
            try:
                rdm = ctx.residual_history(zid)
            except _lib.FastMIError as e:
                if "no residual" not in str(e):
                    raise
                rdm = None
            if rdm is not None and rdm.shape[0] > 0:
                res_l2 += float(rdm[:, :5].sum())
                res_li = max(res_li, float(rdm[:, 5:].max()))
                cvg = _dpssiter(z[0], nitrun, rdm, lft, iverb)
        if cvg is not None or lft < 0:
            _conv2pytree(z, cvg, nitrun, lft, nd)
    if stopAtNan:
        for (z, hz), zid in zip(HOOK["zones"], HOOK["ids"]):
            # the kinetic-energy sum is not finite when an interior density or velocity is not
            if not np.isfinite(ctx.tke(zid, 0)[0]):
                sys.exit(1)
    if gmres is None:
        return None
    return {"L2": res_l2, "Loo": res_li}


displayTemporalCriteria = display_temporal_criteria
_displayTemporalCriteria = display_temporal_criteria


@_api
def _createConvergenceHistory(t, nrecs):
    """GlobalConvergenceHistory per base and ZoneConvergenceHistory per zone, each valued with
    its record count (PyTree.py:2152-2190): IterationNumber (nrecs), RSD_L2, RSD_L2_diff,
    RSD_oo, RSD_oo_diff (5 nrecs; -inf for the global Linf)."""
    neqs = 5
    for b in Internal.getBases(t):
        g = Internal.createUniqueChild(b, "GlobalConvergenceHistory", "ConvergenceHistory_t", 0)
        g[2] = []
        Internal.createChild(g, "IterationNumber", "DataArray_t", np.zeros(nrecs, dtype=np.int32))
        for v in ("RSD_L2", "RSD_L2_diff"):
            Internal.createChild(g, v, "DataArray_t", np.zeros(nrecs * neqs))
        for v in ("RSD_oo", "RSD_oo_diff"):
            Internal.createChild(g, v, "DataArray_t", np.full(nrecs * neqs, -np.inf))
        for z in Internal.getNodesFromType1(b, "Zone_t"):
            h = Internal.createUniqueChild(z, "ZoneConvergenceHistory", "ConvergenceHistory_t", 0)
            h[2] = []
            Internal.createChild(h, "IterationNumber", "DataArray_t", np.zeros(nrecs, dtype=np.int32))
            for v in ("RSD_L2", "RSD_L2_diff", "RSD_oo", "RSD_oo_diff"):
                Internal.createChild(h, v, "DataArray_t", np.zeros(nrecs * neqs))
    return None


@_api
def _calc_global_convergence(t, nrec=None):
    """GlobalConvergenceHistory of every base from its zones (PyTree.py:3852-3910): iteration
    numbers of the first zone, Linf the maximum over the zones, L2 the average weighted by each
    zone's point count (the product of the zone's vertex counts, as the reference takes it);
    all records, or record ``nrec`` only."""
    neqs = 5
    ib, ie = (None, None) if nrec is None else (neqs * int(nrec), neqs * (int(nrec) + 1))
    for b in Internal.getBases(t):
        zones = Internal.getNodesFromType1(b, "Zone_t")
        g = Internal.getNodeFromName1(b, "GlobalConvergenceHistory")
        z0 = Internal.getNodeFromName1(zones[0], "ZoneConvergenceHistory")
        it0 = Internal.getNodeFromName1(z0, "IterationNumber")[1]
        itb = Internal.getNodeFromName1(g, "IterationNumber")[1]
        cnt = np.asarray(g[1]).ravel()
        if nrec is None:
            cnt[0] = itb.shape[0]
            itb[:] = it0[:]
        else:
            cnt[0] += 1
            itb[int(nrec)] = it0[int(nrec)]
        g[1] = cnt.reshape(np.asarray(g[1]).shape)
        ncb = 0
        for z in zones:
            h = Internal.getNodeFromName1(z, "ZoneConvergenceHistory")
            ic = Internal.getNodeFromName1(z, "isCalc")
            if ic is not None and Internal.getValue(ic) != 1:
                continue
            d = z[1]
            ncz = int(d[0][0]) * int(d[1][0]) * int(d[2][0])
            ncb += ncz
            for v in ("RSD_oo", "RSD_oo_diff"):
                gb = Internal.getNodeFromName1(g, v)[1]
                gb[ib:ie] = np.maximum(gb[ib:ie], Internal.getNodeFromName1(h, v)[1][ib:ie])
            for v in ("RSD_L2", "RSD_L2_diff"):
                gb = Internal.getNodeFromName1(g, v)[1]
                gb[ib:ie] += Internal.getNodeFromName1(h, v)[1][ib:ie] * ncz
        for v in ("RSD_L2", "RSD_L2_diff"):
            gb = Internal.getNodeFromName1(g, v)[1]
            gb[ib:ie] /= float(ncb)
    return None


_calcGlobalConvergence = _calc_global_convergence


def write_plt_format(t, path, fcvg, nd, convergence_name="ZoneConvergenceHistory"):
    """One block of the tecplot ascii history (PyTree.py:2196-2240): the records of the
    convergence node below ``path``."""
    node = Internal.getNodeFromPath(t, path + "/" + convergence_name)
    last = int(np.asarray(node[1]).ravel()[0])
    it = Internal.getNodeFromName1(node, "IterationNumber")[1]
    nrec = it.size
    a = Internal.getNodeFromName1(node, "RSD_L2")[1]
    neq = a.size // nrec
    arrs = [np.reshape(Internal.getNodeFromName1(node, v)[1], (neq, nrec), order="F")
            for v in ("RSD_L2", "RSD_oo", "RSD_L2_diff", "RSD_oo_diff")]
    if nd == 0:
        fcvg.write("VARIABLES = it RO_l2 ROU_l2 ROV_l2 ROW_l2 ROE_l2 ROoo ROUoo ROVoo ROWoo ROEoo "
                   "RO_l2_diff ROU_l2_diff ROV_l2_diff ROW_l2_diff ROE_l2_diff ROoo_diff ROUoo_diff "
                   "ROVoo_diff ROWoo_diff ROEoo_diff\n")
    fcvg.write('ZONE T="bloc %s " I=%d F=POINT\n' % (path, last))
    for l in range(last):
        vals = "".join("{0:7f} ".format(arr[k, l]) for arr in arrs for k in range(neq))
        fcvg.write("%s %s\n" % (1 + it[l], vals))
    return 1


@_api
def _extractConvergenceHistory(t, fileout, perZones=True, perBases=True):
    """The convergence histories of the tree in a tecplot ascii file (PyTree.py:2245-2270):
    one block per zone, then one per base whose GlobalConvergenceHistory has records."""
    nd = 0
    with open(fileout, "w") as f:
        if perZones:
            for path in Internal.getZonePaths(t):
                nd = write_plt_format(t, path, f, nd, "ZoneConvergenceHistory")
        if perBases:
            for path in Internal.getPathsFromType(t, "CGNSBase_t"):
                b = Internal.getNodeFromPath(t, path)
                g = Internal.getNodeFromName1(b, "GlobalConvergenceHistory")
                if g is not None and g[2]:
                    nd = write_plt_format(t, path, f, nd, "GlobalConvergenceHistory")
    return None


def extractConvergenceHistory(t, fileout, perZones=True, perBases=True):
    return _extractConvergenceHistory(t, fileout, perZones=perZones, perBases=perBases)


# ---------------------------------------------------------------------------------------
# Restart (FastC.save / load, Fast/FastC/FastC/PyTree.py:3514-3800).  The reference writes CGNS
# through Cassiopee (HDF5), which is not available: the same tree goes into a NumPy archive
# (no pickled objects) under the given file name.
@_api
def saveTree(t, fileName="restart.cgns", split="single", compress=0):
    """Save the tree for a restart: the current state is pulled from the GPU first; the _P1
    level and .Solver#ownData are left out as in the reference (FastC/PyTree.py:3755-3800), so
    a restart has N and N-1 and runs BDF2 from its first step."""
    if HOOK.get("stale"):
        _pull_levels()
    nodes, arrays = [], {}

    def walk(node, parent):
        if node[0] == ".Solver#ownData" or node[0].endswith("_P1"):
            return
        idx = len(nodes)
        v = node[1]
        kind = "none"
        if isinstance(v, np.ndarray):
            kind = "array"
            arrays["v%d" % idx] = np.asfortranarray(v) if v.ndim > 1 else v
        elif v is not None:
            kind = "array"
            arrays["v%d" % idx] = np.asarray(v)
        nodes.append((node[0], node[3], parent, kind, int(isinstance(v, np.ndarray) and v.ndim > 1 and v.flags.f_contiguous)))
        for c in node[2]:
            walk(c, idx)
    walk(t, -1)
    import json
    arrays["__index__"] = np.frombuffer(json.dumps(nodes).encode(), dtype=np.uint8)
    with open(fileName, "wb") as f:
        np.savez(f, **arrays)


def loadTree(fileName="t.cgns", split="single", graph=False, exploc=0):
    """Read a tree written by saveTree."""
    import json
    d = np.load(fileName, allow_pickle=False)
    nodes = json.loads(bytes(d["__index__"]).decode())
    built = []
    for idx, (name, ntype, parent, kind, fortran) in enumerate(nodes):
        v = None
        if kind == "array":
            v = d["v%d" % idx]
            if fortran:
                v = np.asfortranarray(v)
        node = [name, v, [], ntype]
        if parent >= 0:
            built[parent][2].append(node)
        built.append(node)
    t = built[0]
    return (t, None) if graph else t


def save(t, fileName="restart.cgns", tc=None, fileNameC="tc_restart.cgns", ts=None,
         fileNameS="tstat_restart.cgns", split="single", compress=0):
    """Save the trees (FastC.PyTree.save): t, and tc / ts when given."""
    saveTree(t, fileName=fileName, split=split, compress=compress)
    if tc is not None and not isinstance(tc, list):
        saveTree(tc, fileName=fileNameC, split=split, compress=compress)
    if ts is not None:
        saveTree(ts, fileName=fileNameS, split=split, compress=compress)


def load(fileName="t.cgns", fileNameC=None, fileNameS=None, split="single", exploc=0):
    """(t, tc, ts, graph) as FastC.PyTree.load: t from fileName, tc / ts when their names are
    given, no communication graph (one GPU)."""
    t = loadTree(fileName, split=split)
    tc = loadTree(fileNameC, split=split) if fileNameC is not None else None
    ts = loadTree(fileNameS, split=split) if fileNameS is not None else None
    return t, tc, ts, None


@_api
def _computeEnstrophy(t, metrics):
    """(enstrophy, tke) volume averages over the interior cells of all zones
    (computePT_enstrophy.cpp:206-227): sum 1/2 rho |curl u|^2 vol / sum vol and
    sum 1/2 rho |u|^2 vol / sum vol, from the N level."""
    ctx = HOOK["ctx"]
    _device_current()
    k = e = v = 0.0
    for (z, hz), zid in zip(HOOK["zones"], HOOK["ids"]):
        s = ctx.tke(zid, 0)
        k += s[0]; v += s[1]; e += s[2]
    return e / v, k / v


# ---------------------------------------------------------------------------------------
# Wall efforts (FastS.PyTree.createStressNodes / _computeStress, PyTree.py:2277-2566)
EFF_VARS = ["Density", "MomentumX", "MomentumY", "MomentumZ", "EnergyStagnationDensity",
            "gradxVelocityX", "gradyVelocityX", "gradzVelocityX", "gradxVelocityY", "gradyVelocityY",
            "gradzVelocityY", "gradxVelocityZ", "gradyVelocityZ", "gradzVelocityZ",
            "gradxTemperature", "gradyTemperature", "gradzTemperature", "CoefPressure",
            "ViscosityMolecular", "Density2", "Pressure"]
EFF_LOOP, EFF_I0, EFF_NONZ, EFF_IDIR = 23, 29, 34, 35


@_api
def createStressNodes(t, BC=None, windows=None):
    """A tree of surface zones, one per selected wall window, to receive the efforts
    (PyTree.py:2277-2530).  BC: list of BC types (e.g. ['BCWall', 'BCWallViscous']); windows:
    list of (zone name, i0, i1, j0, j1, k0, k1, 'imin'..'kmax') in the ghost-inclusive vertex
    numbering of the zone.  Each surface zone carries the face window of its NS zone in
    Parameter_int (EFF_LOOP 23-28, EFF_I0 29-33, EFF_NONZ 34, EFF_IDIR 35) and the 21 centre
    fields of the reference, zero until _computeStress."""
    teff = Internal.newCGNSTree()
    no_z = 0
    for b0 in Internal.getBases(t):
        b = Internal.newCGNSBase(b0[0], parent=teff)
        for z in Internal.getNodesFromType1(b0, "Zone_t"):
            own = Internal.getNodeFromName1(z, ".Solver#ownData")
            pnode = Internal.getNodeFromName1(own, "Parameter_int") if own is not None else None
            if pnode is None:
                raise ValueError("createStressNodes must be called after warmup (zone %s)" % z[0])
            p = np.asarray(pnode[1])
            sel = []
            if BC is not None:
                zbc = Internal.getNodeFromName1(z, "ZoneBC")
                for bc in (Internal.getNodesFromType1(zbc, "BC_t") if zbc is not None else []):
                    if Internal.getValue(bc) in BC:
                        rng = np.asarray(Internal.getNodeFromType1(bc, "IndexRange_t")[1])
                        sel.append((bc[0], rng, None))
            for w in (windows or []):
                if w[0] == z[0]:
                    sel.append(("window", np.array([[w[1], w[2]], [w[3], w[4]], [w[5], w[6]]]), w[7]))
            for inum, (name, rng, face) in enumerate(sel):
                flat = [int(rng[0][0]), int(rng[0][1]), int(rng[1][0]), int(rng[1][1]),
                        int(rng[2][0]), int(rng[2][1])]
                idir, fen = P.get_range(flat, p)
                if face is not None:
                    idir = ["imin", "imax", "jmin", "jmax", "kmin", "kmax"].index(face) + 1
                if idir == 0:
                    raise ValueError("window %s of zone %s is not on a face" % (name, z[0]))
                d = (idir - 1) // 2
                win = list(fen)
                nface = win[2 * d] if idir % 2 == 1 else win[2 * d] + 1
                win[2 * d] = win[2 * d + 1] = nface
                tang = [a for a in range(3) if a != d]
                nt = [win[2 * a + 1] - win[2 * a] + 1 for a in tang]
                zp = Internal.newZone("%s_%s%d" % (z[0], name, inum), nt[0] + 1, nt[1] + 1, 1, parent=b)
                g = Internal.getNodeFromName1(z, "GridCoordinates")
                nf, nfk = int(p[P.NIJK + 3]), int(p[P.NIJK + 4])
                sl = [slice(None)] * 3
                for a in range(3):
                    g0 = nf if a < 2 else nfk
                    if a == d:
                        sl[a] = slice(win[2 * a] + g0 - 1, win[2 * a] + g0)
                    else:
                        sl[a] = slice(win[2 * a] + g0 - 1, win[2 * a + 1] + g0 + 1)
                gc = Internal.createChild(zp, "GridCoordinates", "GridCoordinates_t")
                for cn in ("CoordinateX", "CoordinateY", "CoordinateZ"):
                    a3 = np.asarray(Internal.getNodeFromName1(g, cn)[1])[tuple(sl)]
                    Internal.createChild(gc, cn, "DataArray_t",
                                         np.ascontiguousarray(a3).reshape(nt[0] + 1, nt[1] + 1, 1, order="F"))
                fs = Internal.createChild(zp, "FlowSolution#Centers", "FlowSolution_t")
                Internal.createChild(fs, "GridLocation", "GridLocation_t", "CellCenter")
                for v in EFF_VARS:
                    Internal.createChild(fs, v, "DataArray_t", np.zeros((nt[0], nt[1], 1), order="F"))
                pe = np.array(p, dtype=np.int32, copy=True)
                pe[P.NIJK:P.NIJK + 5] = [nt[0], nt[1], 1, nf, 0]
                pe[P.NEQ] = len(EFF_VARS)
                pe[P.NDIMDX] = nt[0] * nt[1]
                pe[EFF_LOOP:EFF_LOOP + 6] = win
                n0 = win[1] - win[0] + 1
                pe[EFF_I0:EFF_I0 + 5] = [win[0], win[2], win[4], n0, n0 * (win[3] - win[2] + 1)]
                pe[EFF_NONZ] = no_z
                pe[EFF_IDIR] = idir
                o = Internal.createUniqueChild(zp, ".Solver#ownData", "UserDefinedData_t")
                Internal.createUniqueChild(o, "Parameter_int", "DataArray_t", pe)
            no_z += 1
    return teff


ADJ_VARS = {c: ["dp%sp_dp%s" % (c.upper(), v) for v in ("Density", "MomentumX", "MomentumY", "MomentumZ",
                                                         "EnergyStagDens")] for c in ("cd", "cl")}


@_api
def _compute_dpJ_dpW(t, teff, metrics, cosAoA, sinAoA, surfinv):
    """Gradients of the pressure drag and lift coefficients of the slip-wall windows of teff with
    respect to the conservative state, assigned into the fields dpCDp_dp* and dpCLp_dp* of the
    wall cells (compute_dpJ_dpW.cpp:33-261 via PyTree.py:3719-3735; the fields are created by
    warmup(..., Adjoint=True)).  AUSM+(P) o3 Euler zones."""
    ctx = HOOK["ctx"]
    _device_current()
    for ze in Internal.getZones(teff):
        pe = np.asarray(Internal.getNodeFromName1(Internal.getNodeFromName1(ze, ".Solver#ownData"),
                                                  "Parameter_int")[1])
        no = int(pe[EFF_NONZ])
        z, hz = HOOK["zones"][no]
        fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
        arrs = {}
        for c in ("cd", "cl"):
            nodes = [Internal.getNodeFromName1(fs, v) for v in ADJ_VARS[c]]
            if any(nd_ is None for nd_ in nodes):
                raise ValueError("zone %s has no %s fields: call warmup(..., Adjoint=True)" % (z[0], ADJ_VARS[c][0]))
            arrs[c] = (nodes, np.concatenate([nd_[1].ravel(order="F") for nd_ in nodes]))
        win = [int(v) for v in pe[EFF_LOOP:EFF_LOOP + 6]]
        ctx.dpj_dpw(HOOK["ids"][no], int(pe[EFF_IDIR]), win, cosAoA, sinAoA, surfinv,
                    arrs["cd"][1], arrs["cl"][1])
        for c in ("cd", "cl"):
            nodes, flat = arrs[c]
            nd = hz.nd
            for e, node in enumerate(nodes):
                node[1][...] = flat[e * nd:(e + 1) * nd].reshape(node[1].shape, order="F")
    return None


@_api
def _computeStress(t, teff, metrics, xyz_ref=(0., 0., 0.)):
    """Efforts on the windows of teff from the N level of every zone, fields stored in teff
    (compute_effort.cpp -> bceffort.for).  Returns the 11 sums of the reference: forces x, y, z
    and moments about xyz_ref (pressure minus the free-stream pressure plus viscous stress,
    scaled by 2 / (rho_inf V_inf^2)), the wetted area, and the mass, x-, y-, z-momentum flux
    sums."""
    ctx = HOOK["ctx"]
    _device_current()
    effort = np.zeros(11)
    for z in Internal.getZones(teff):
        pe = np.asarray(Internal.getNodeFromName1(Internal.getNodeFromName1(z, ".Solver#ownData"),
                                                  "Parameter_int")[1])
        zid = HOOK["ids"][int(pe[EFF_NONZ])]
        win = [int(v) for v in pe[EFF_LOOP:EFF_LOOP + 6]]
        eff, fl = ctx.wall_effort(zid, int(pe[EFF_IDIR]), win, xyz_ref, level=0)
        fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
        for q, v in enumerate(EFF_VARS):
            node = Internal.getNodeFromName1(fs, v)
            node[1][...] = fl[q].reshape(node[1].shape, order="F")
        effort[:] = effort + eff[:11]
    return effort
