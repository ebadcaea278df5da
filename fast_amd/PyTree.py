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
VelocityX/Y/Z, Temperature and their _M1/_P1 time levels, ghost cells included); the device
copy is authoritative during ``_compute`` and the tree arrays are refreshed in place after it.
Like the reference's module-global ``FastC.HOOK``, one solver state lives per process.
"""
import sys

import numpy as np

from . import Internal
from . import params as P
from .zone import HostZone

VARS = ["Density", "VelocityX", "VelocityY", "VelocityZ", "Temperature"]
HOOK = {}
_MODELS = {"NSLaminar": "lamin", "Euler": "euler"}
_STATE_KEYS = ("gamma", "cv", "ro", "p", "u", "T", "mach", "reynolds", "prandtl", "mus", "ts",
               "cs", "vx", "vy", "vz")


# ---------------------------------------------------------------------------------------
# numerics (FastC._setNum2Base / _setNum2Zones)
def _set_define(node, num):
    d = Internal.createUniqueChild(node, ".Solver#define", "UserDefinedData_t")
    for k, v in num.items():
        Internal.createUniqueChild(d, k, "DataArray_t", v)


def _setNum2Base(t, numb):
    for b in Internal.getBases(t):
        _set_define(b, numb)


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


def _init_metric(t, metrics, hook=None):
    """Face vectors and volumes (computed on the GPU when the zone was registered)."""
    from . import _lib
    ctx = HOOK["ctx"]
    for (z, hz), zid, m in zip(HOOK["zones"], HOOK["ids"], metrics):
        ctx.get(zid, _lib.FIELD_METRIC, m[0])


def warmup(t, tc=None, graph=None, infos_ale=None, Adjoint=False, tmy=None, list_graph=None,
           Padding=None, verbose=0):
    """Build the solver state: metrics, time levels _M1/_P1, ghost cells.  -> (t, tc, metrics)

    Zones are joined by the GridConnectivity1to1 nodes of t (injection transfers of the ghost
    cells, as the reference's tc of a matching mesh); ``tc`` may instead be a list of
    fast_amd.connect.Raccord.  A Cassiopee connectivity tree is not accepted (it needs
    Connector)."""
    from . import _lib
    if tc is not None and not (isinstance(tc, list) and all(hasattr(r, "rcv") for r in tc)):
        raise NotImplementedError("Cassiopee connectivity trees are not supported: give the matches as "
                                  "GridConnectivity1to1 nodes of t or a list of fast_amd.connect.Raccord")
    HOOK.clear()
    metrics = allocate_metric(t, verbose)
    _init_metric(t, metrics, HOOK)
    ctx = HOOK["ctx"]
    racs = list(tc) if tc is not None else _matches(HOOK["zones"])
    ctx.set_transfers(racs)
    HOOK["raccords"] = racs
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
    return t, tc, metrics


def _compute(t, metrics, nitrun, tc=None, graph=None, tc2=None, graph2=None, layer="c", NIT=1,
             ucData=None, vtune=None):
    """NIT time steps (all sub-iterations) of every zone on the GPU; tree fields updated in place."""
    import time
    from . import _lib
    ctx = HOOK["ctx"]
    base = Internal.getBases(t)[0]
    bdef = _get_define(base)
    modulo_verif = int(bdef.get("modulo_verif", 200))
    hz0 = HOOK["zones"][0][1]
    nssiter = int(hz0.param_int[P.NSSITER])
    t0 = time.time()
    ctx.compute(nitrun, 1, nssiter, nit_c=NIT, modulo_verif=modulo_verif, first_it=HOOK["first_it"])
    ctx.sync()
    t1 = time.time()
    HOOK["first_it"] = 1
    for (z, hz), zid in zip(HOOK["zones"], HOOK["ids"]):
        for sfx, f, a in (("", _lib.FIELD_RO, hz.ro), ("_M1", _lib.FIELD_RO_M1, hz.ro_m1),
                          ("_P1", _lib.FIELD_RO_P1, hz.ro_p1)):
            ctx.get(zid, f, a)
            _level_to_tree(z, hz, sfx, a)
        # CFL max / min / mean of the last verification step (computePT.cpp:629-645)
        try:
            cfl = ctx.cfl(zid)
        except _lib.FastMIError:
            cfl = None
        if cfl is not None:
            o = Internal.createUniqueChild(z, ".Solver#ownData", "UserDefinedData_t")
            Internal.createUniqueChild(o, "CFL_minmaxmoy", "DataArray_t", cfl)
    if vtune is not None:
        return t1 - t0, 0.0
    return None


def _fillGhostcells(zones, tc, metrics, timelevel_target, vars, nstep, omp_mode=0, hook1=None,
                    nitmax=1, rk=1, exploc=0, num_passage=1):
    """Transfers between zones then physical boundary conditions on the N level of every zone
    (Fast/FastS/FastS/PyTree.py:1162-1230 _fillGhostcells: setInterpTransfers then _applyBC)."""
    from . import _lib
    ctx = HOOK["ctx"]
    for (z, hz), zid in zip(HOOK["zones"], HOOK["ids"]):
        _tree_to_level(z, hz, "", hz.ro)
        ctx.set(zid, _lib.FIELD_RO, hz.ro)
    ctx.transfer(0)
    for (z, hz), zid in zip(HOOK["zones"], HOOK["ids"]):
        ctx.apply_bc(zid, 0, nstep)
        ctx.get(zid, _lib.FIELD_RO, hz.ro)
        _level_to_tree(z, hz, "", hz.ro)


def _applyBC(t, metrics, hook1=None, nstep=1, var="Density"):
    from . import _lib
    ctx = HOOK["ctx"]
    for (z, hz), zid in zip(HOOK["zones"], HOOK["ids"]):
        _tree_to_level(z, hz, "", hz.ro)
        ctx.set(zid, _lib.FIELD_RO, hz.ro)
        ctx.apply_bc(zid, 0, nstep)
        ctx.get(zid, _lib.FIELD_RO, hz.ro)
        _level_to_tree(z, hz, "", hz.ro)


def display_temporal_criteria(t, metrics, nitrun, format=None, gmres=None, verbose="firstlast",
                              stopAtNan=True):
    """Residual L2 / Linf per zone and equation of the last verified sub-iteration
    (display_ss_iteration.cpp:112-145).  format "store" records them in the tree's
    ZoneConvergenceHistory; returns {"L2": ..., "Loo": ...} arrays [zone, equation]."""
    ctx = HOOK["ctx"]
    l2, loo = [], []
    for (z, hz), zid in zip(HOOK["zones"], HOOK["ids"]):
        try:
            r = ctx.residual_norms(zid)
        except Exception:
            r = np.full(10, np.nan)
        l2.append(r[:5]); loo.append(r[5:])
        if format == "store":
            h = Internal.getNodeFromName1(z, "ZoneConvergenceHistory")
            if h is not None:
                it = Internal.getNodeFromName1(h, "IterationValues")[1]
                rl2 = Internal.getNodeFromName1(h, "RSD_L2")[1]
                rlo = Internal.getNodeFromName1(h, "RSD_oo")[1]
                k = int(Internal.getNodeFromName1(h, "LastRecordedIteration")[1][0])
                if k < it.size:
                    it[k] = nitrun; rl2[k * 5:(k + 1) * 5] = r[:5]; rlo[k * 5:(k + 1) * 5] = r[5:]
                    Internal.getNodeFromName1(h, "LastRecordedIteration")[1][0] = k + 1
        elif format is not None or verbose != "none":
            fmt = "%.16e" if format == "double" else "%.6e"
            print("iteration %d zone %s: L2 %s Loo %s" % (nitrun, z[0], " ".join(fmt % v for v in r[:5]),
                                                         " ".join(fmt % v for v in r[5:])))
    if stopAtNan:
        for z, hz in HOOK["zones"]:
            if not np.isfinite(hz.ro).all():
                sys.exit(1)
    out = {"L2": np.array(l2), "Loo": np.array(loo)}
    return out if gmres is not None else out


displayTemporalCriteria = display_temporal_criteria


def _createConvergenceHistory(t, nrecs):
    for z in Internal.getZones(t):
        h = Internal.createUniqueChild(z, "ZoneConvergenceHistory", "ConvergenceHistory_t")
        Internal.createUniqueChild(h, "LastRecordedIteration", "DataArray_t", np.zeros(1, np.int32))
        Internal.createUniqueChild(h, "IterationValues", "DataArray_t", np.zeros(nrecs, np.int32))
        Internal.createUniqueChild(h, "RSD_L2", "DataArray_t", np.zeros(5 * nrecs))
        Internal.createUniqueChild(h, "RSD_oo", "DataArray_t", np.zeros(5 * nrecs))


def _computeEnstrophy(t, metrics):
    """(enstrophy, tke) volume averages over the interior cells of all zones
    (computePT_enstrophy.cpp:206-227): sum 1/2 rho |curl u|^2 vol / sum vol and
    sum 1/2 rho |u|^2 vol / sum vol, from the N level."""
    ctx = HOOK["ctx"]
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


def _compute_dpJ_dpW(t, teff, metrics, cosAoA, sinAoA, surfinv):
    """Gradients of the pressure drag and lift coefficients of the slip-wall windows of teff with
    respect to the conservative state, assigned into the fields dpCDp_dp* and dpCLp_dp* of the
    wall cells (compute_dpJ_dpW.cpp:33-261 via PyTree.py:3719-3735; the fields are created by
    warmup(..., Adjoint=True)).  AUSM+(P) o3 Euler zones."""
    ctx = HOOK["ctx"]
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


def _computeStress(t, teff, metrics, xyz_ref=(0., 0., 0.)):
    """Efforts on the windows of teff from the N level of every zone, fields stored in teff
    (compute_effort.cpp -> bceffort.for).  Returns the 11 sums of the reference: forces x, y, z
    and moments about xyz_ref (pressure minus the free-stream pressure plus viscous stress,
    scaled by 2 / (rho_inf V_inf^2)), the wetted area, and the mass, x-, y-, z-momentum flux
    sums."""
    ctx = HOOK["ctx"]
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
