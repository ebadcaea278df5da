"""Zone descriptor: the ``Parameter_int`` / ``Parameter_real`` vectors of a structured zone.

The kernels of the reference receive all of their inputs through two flat vectors per zone,
``param_int`` (int32) and ``param_real`` (float64), whose slot numbers are fixed by
``Fast/FastS/FastS/param_solver.h:80-312``.  This module keeps the same slot numbers so that a zone
built here can be handed unchanged to the reference kernels (test oracle) and to this package's HIP
kernels through the C ABI.

Construction follows ``FastC._buildOwnData`` (``Fast/FastC/FastC/PyTree.py:1754-2047``) for the
numerics keys, ``allocate_metric`` (``Fast/FastS/FastS/Metric/allocate_metric.cpp:120-176``) for the
dimensions, and ``FastC._BCcompact`` / ``getRange`` (``Fast/FastC/FastC/PyTree.py:2966-3040``,
``Fast/FastC/FastC/getRange.cpp:57-110``) for the boundary-condition table.
"""
import math
import numpy as np

# ---- param_int slots (param_solver.h) ------------------------------------------------------
NIJK = 0; NIJK_MTR = 5; NIJK_XYZ = 10; NIJK_VENT = 15; IJKV = 20; MXSSDOM_LU = 24
ITYPZONE = 25; ITYPVENT = 26; IFLOW = 27; ILES = 28; ITYPCP = 29; SIZE_SSDOM = 30
KFLUDOM = 33; LALE = 34; IFLAGFLT = 35; NEQ = 36; NEQ_IJ = 37; NEQ_K = 38; NEQ_COE = 39
NEQ_VENT = 40; NDIMDX = 41; NDIMDX_XYZ = 42; NDIMDX_MTR = 43; NDIMDX_VENT = 44; IO_THREAD = 45
DTLOC = 46; SA_INT = 47; RK = 52; LEVEL = 53; EXPLOC = 54; ITEXPLOC = 55; LEVELG = 56
LEVELD = 57; NSSITER = 58; CACHEBLCKI = 59; CACHEBLCKJ = 60; CACHEBLCKK = 61; SFD = 62
SFD_INIT_IT = 63; SLOPE = 64; NIT_INFLOW = 65; SHIFTVAR = 66; EXTRACT_RES = 67; SA_DEBUG = 68
PT_OMP = 69; PT_BC = 70; NB_RELAX = 71; NB_RESTART = 72; NB_KRYLOV = 73; IMPLICITSOLVER = 74
LU_MATCH = 75; IBC = 76; SRC = 83; MESHTYPE = 84; SENSORTYPE = 85; SCHEDULER = 86
WM_FUNCTION = 87; WM_SAMPLING = 88; LBM_SPONGE = 102; IBC_PT_FLUX = 129; SA_LOW_RE = 130
SA_ROT_CORR = 131; SA_DIST = 132; PT_INTERP = 133; NONZ = 134; PT_VOL = 135
N_PARAM_INT = 137          # number_of_defines_param_int + 1 (FastC/PyTree.py:1754-1755)

BC_TYPE = 0; BC_IDIR = 1; BC_FEN = 2; BC_NBDATA = 8

# ---- param_real slots ---------------------------------------------------------------------
DTC = 0; GAMMA = 1; STAREF = 1; CVINF = 2; ROINF = 3; PINF = 4; VINF = 5; TINF = 6; MINF = 7
RONUTILDEINF = 8; REYNOLDS = 9; PRANDT = 10; XMUL0 = 11; TEMP0 = 12; CS = 13; EPSI_NEWTON = 14
CFL = 15; VXINF = 19; VYINF = 20; VZINF = 21; TEMPS = 22; PSIROE = 34; SFD_CHI = 35
SFD_DELTA = 36; EPSI_INFLOW = 37; PRANDT_TUR = 38; EPSI_LINEAR = 39; HPC_CUPS = 40; BC_DATA = 41
NUDGING_EQ2 = 67; WIG_AMP = 72
N_PARAM_REAL = 73

# ---- enumerations (FastC/PyTree.py:1686-1749, tagBCDict :30-84) ---------------------------
ZONE_3DFULL, ZONE_3DHOMO, ZONE_3DCART, ZONE_2D = 0, 1, 2, 3
SCHEMES = {"ausmpred": (1, None), "senseur": (2, None), "roe_kap": (5, 2), "roe_min": (5, 3),
           "roe_nul": (5, 1), "roe": (5, None)}
SLOPES = {"o1": 1, "o3": 2, "minmod": 3}
BC_TAGS = {"BCExtrapolate": 0, "BCDegenerateLine": 0, "BCFarfield": 1, "BCInflowSupersonic": 2,
           "BCWallInviscid": 3, "BCSymmetryPlane": 3, "BCWall": 4, "BCWallViscous": 6,
           "BCOutflow": 10, "BCautoperiod": 11, "BCInflow": 13, "BCOutpres": 16}


class Numerics(dict):
    """User numerics with the reference defaults (FastC/PyTree.py:1117-1331), cache blocking
    included (2048 / 2 / 7, FastC/PyTree.py:1299-1301): the block sizes decide in which order a
    cell receives the fluxes of its faces, and the GPU residual follows that order for any j and
    k block sizes (one block in i)."""
    DEFAULTS = dict(temporal_scheme="implicit", ss_iteration=30, rk=3, modulo_verif=200,
                    scheme="ausmpred", slope="o3", psiroe=0.1, epsi_newton=0.1, cfl=1.0,
                    time_step=1.0e-4, time_step_nature="global", cache_blocking_I=2048,
                    cache_blocking_J=2, cache_blocking_K=7, sensor_type=1,
                    ausm_wiggle=1.0, io_thread=0, nit_inflow=10, epsi_inflow=1.0e-5,
                    prandtltb=None, implicit_solver="lussor", nb_relax=1, lu_match=0)

    def __init__(self, **kw):
        d = dict(self.DEFAULTS)
        d.update(kw)
        super().__init__(d)


def nssiter_of(num):
    """Sub-iterations per time step (FastC/PyTree.py:1159-1174)."""
    if num["temporal_scheme"] == "explicit":
        return int(num["rk"])
    return int(num["ss_iteration"]) + 1


def build_param_int(ni, nj, nk, num, flow="lamin", nfic=2):
    """Return (param_int, ...) for a zone of ni x nj x nk interior cells.

    Dimensions follow allocate_metric.cpp:126-151 with ``nfic`` ghost layers in every
    direction; metric-dependent slots (NIJK_MTR, NDIMDX_MTR, ITYPZONE, NEQ_IJ, NEQ_K) are filled
    later by :func:`fast_amd.metric.classify_zone`.
    """
    p = np.full(N_PARAM_INT, -1000, dtype=np.int32)
    p[0:25] = -1
    scheme = num["scheme"]
    if scheme not in SCHEMES:
        raise ValueError("unsupported scheme %r" % scheme)
    kfludom, islope = SCHEMES[scheme]
    if islope is None:
        islope = SLOPES[num["slope"]]
    iflow = {"euler": 1, "lamin": 2}[flow]
    itypcp = 2 if num["temporal_scheme"] == "explicit" else 1
    p[ITYPZONE] = 0; p[ITYPVENT] = 2; p[IFLOW] = iflow; p[ILES] = 0; p[ITYPCP] = itypcp
    p[SIZE_SSDOM:SIZE_SSDOM + 3] = (100, 10, 10)
    p[KFLUDOM] = kfludom; p[LALE] = 0; p[IFLAGFLT] = 0
    p[IO_THREAD] = num["io_thread"]
    p[DTLOC] = 1 if num["time_step_nature"] == "local" else 0
    p[SA_INT:SA_INT + 5] = 0
    p[RK] = num["rk"]; p[LEVEL] = 1; p[EXPLOC] = 0; p[ITEXPLOC] = 0; p[LEVELG] = 1; p[LEVELD] = 0
    p[NSSITER] = nssiter_of(num)
    p[CACHEBLCKI] = num["cache_blocking_I"]; p[CACHEBLCKJ] = num["cache_blocking_J"]
    p[CACHEBLCKK] = num["cache_blocking_K"]
    p[SFD] = 0; p[SFD_INIT_IT] = 0; p[SLOPE] = islope; p[NIT_INFLOW] = num["nit_inflow"]
    p[SHIFTVAR] = 0; p[EXTRACT_RES] = 0; p[SA_DEBUG] = 0; p[PT_OMP] = 0; p[PT_BC] = 0
    p[NB_RELAX] = num["nb_relax"]; p[NB_RESTART] = 1; p[NB_KRYLOV] = 1
    p[IMPLICITSOLVER] = 0; p[LU_MATCH] = num["lu_match"]; p[IBC:IBC + 7] = 0
    p[SRC] = 0; p[MESHTYPE] = 1; p[SENSORTYPE] = num["sensor_type"]; p[SCHEDULER] = -1
    p[WM_FUNCTION] = 0; p[WM_SAMPLING] = 0
    p[89:129] = 0
    p[IBC_PT_FLUX] = -1; p[SA_LOW_RE] = 0; p[SA_ROT_CORR] = 0; p[SA_DIST] = 0
    p[PT_INTERP] = 0; p[NONZ] = 0; p[PT_VOL] = 0
    # dimensions (allocate_metric.cpp:126-151), vertices include the ghost layers
    nkfic = nfic
    p[NIJK_XYZ:NIJK_XYZ + 5] = (ni + 2 * nfic + 1, nj + 2 * nfic + 1, nk + 2 * nkfic + 1, nfic, nkfic)
    p[NIJK:NIJK + 5] = (ni + 2 * nfic, nj + 2 * nfic, nk + 2 * nkfic, nfic, nkfic)
    p[IJKV:IJKV + 3] = (ni, nj, nk)
    p[NDIMDX] = int(p[NIJK]) * int(p[NIJK + 1]) * int(p[NIJK + 2]) + int(p[SHIFTVAR])
    p[NDIMDX_XYZ] = int(p[NIJK_XYZ]) * int(p[NIJK_XYZ + 1]) * int(p[NIJK_XYZ + 2])
    p[NEQ] = 5
    p[NEQ_COE] = 1 if itypcp == 2 else 5
    # no mesh motion: vent arrays reduced to one point (allocate_metric.cpp:200-260)
    p[NEQ_VENT] = 3; p[NIJK_VENT:NIJK_VENT + 5] = 0; p[NDIMDX_VENT] = 1
    p[MXSSDOM_LU] = 1
    return p


def build_param_real(num, state, flow="lamin"):
    """param_real from the numerics and a reference state dict.

    ``state`` keys: gamma, cv, ro, p, u (magnitude), T, mach, reynolds, prandtl, mus, ts, cs,
    vx, vy, vz (slots 1-13, 19-21 of FastC/PyTree.py:1959-1977).
    """
    r = np.zeros(N_PARAM_REAL, dtype=np.float64)
    r[DTC] = num["time_step"]
    r[GAMMA] = state["gamma"]; r[CVINF] = state["cv"]; r[ROINF] = state["ro"]
    r[PINF] = state["p"]; r[VINF] = state["u"]; r[TINF] = state["T"]; r[MINF] = state["mach"]
    r[RONUTILDEINF] = 0.0; r[REYNOLDS] = state["reynolds"]; r[PRANDT] = state["prandtl"]
    r[XMUL0] = state["mus"]; r[TEMP0] = state["ts"]; r[CS] = state["cs"]
    r[VXINF] = state["vx"]; r[VYINF] = state["vy"]; r[VZINF] = state["vz"]
    r[EPSI_NEWTON] = num["epsi_newton"]; r[CFL] = num["cfl"]; r[TEMPS] = 0.0
    r[PSIROE] = num["psiroe"]; r[EPSI_INFLOW] = num["epsi_inflow"]
    prt = num["prandtltb"]
    r[PRANDT_TUR] = state["prandtl"] if prt is None else prt
    r[WIG_AMP] = num["ausm_wiggle"]
    return r


def reference_state(mach=0.1, reynolds=1600.0, prandtl=0.72, gamma=1.4, ro=1.0, u=1.0, T=1.0,
                    length=1.0, direction=(1.0, 0.0, 0.0), cs=0.3831337844872463):
    """A non-dimensional reference state.

    ``cv`` is chosen so that the sound speed at (ro, T) gives the requested Mach number with
    velocity magnitude ``u``; Sutherland's law is anchored so that mu(T) = ro*u*length/Re
    (mu_s at T_s = T).  Cassiopee's ``KCore.Adim`` is not available, so the reference state is
    defined explicitly here and stored in the fixtures with every case.
    """
    c = u / mach
    p = ro * c * c / gamma
    R = p / (ro * T)
    cv = R / (gamma - 1.0)
    mus = ro * u * length / reynolds
    n = np.asarray(direction, dtype=np.float64)
    n = n / np.linalg.norm(n)
    return dict(gamma=gamma, cv=cv, ro=ro, p=p, u=u, T=T, mach=mach, reynolds=reynolds,
                prandtl=prandtl, mus=mus, ts=T, cs=cs * T, vx=u * n[0], vy=u * n[1], vz=u * n[2])


def get_range(in_bc, p):
    """CGNS vertex range -> cell window and direction (getRange.cpp:57-110)."""
    fen = [0] * 6
    nf, nfk = int(p[NIJK + 3]), int(p[NIJK + 4])
    ijkv = [int(p[IJKV]), int(p[IJKV + 1]), int(p[IJKV + 2])]
    if in_bc[0] == in_bc[1]:
        fen[0] = in_bc[0] - 1 if in_bc[0] > 1 else in_bc[0]
        fen[0] = max(1, fen[0] - 2 * nf); fen[1] = fen[0]
        fen[2] = max(in_bc[2] - nf, 1); fen[3] = min(in_bc[3] - nf - 1, ijkv[1])
        fen[4] = max(in_bc[4] - nfk, 1); fen[5] = min(in_bc[5] - nfk - 1, ijkv[2])
    elif in_bc[2] == in_bc[3]:
        fen[2] = in_bc[2] - 1 if in_bc[2] > 1 else in_bc[2]
        fen[2] = max(1, fen[2] - 2 * nf); fen[3] = fen[2]
        fen[0] = max(in_bc[0] - nf, 1); fen[1] = min(in_bc[1] - nf - 1, ijkv[0])
        fen[4] = max(in_bc[4] - nfk, 1); fen[5] = min(in_bc[5] - nfk - 1, ijkv[2])
    elif in_bc[4] == in_bc[5]:
        fen[4] = in_bc[4] - 1 if in_bc[4] > 1 else in_bc[4]
        fen[4] = max(1, fen[4] - 2 * nfk); fen[5] = fen[4]
        fen[0] = max(in_bc[0] - nf, 1); fen[1] = min(in_bc[1] - nf - 1, ijkv[0])
        fen[2] = max(in_bc[2] - nf, 1); fen[3] = min(in_bc[3] - nf - 1, ijkv[1])
    idir = 0
    if fen[0] == fen[1] and fen[0] == 1: idir = 1
    elif fen[0] == fen[1] and fen[0] == ijkv[0]: idir = 2
    elif fen[2] == fen[3] and fen[2] == 1: idir = 3
    elif fen[2] == fen[3] and fen[2] == ijkv[1]: idir = 4
    elif fen[4] == fen[5] and fen[4] == 1: idir = 5
    elif fen[4] == fen[5] and fen[4] == ijkv[2]: idir = 6
    return idir, fen


def face_window(p, face):
    """Cell window (1-based, interior) covering a whole face 'imin'..'kmax'."""
    ni, nj, nk = int(p[IJKV]), int(p[IJKV + 1]), int(p[IJKV + 2])
    idir = ["imin", "imax", "jmin", "jmax", "kmin", "kmax"].index(face) + 1
    fen = [1, ni, 1, nj, 1, nk]
    if idir == 1: fen[1] = 1
    elif idir == 2: fen[0] = ni
    elif idir == 3: fen[3] = 1
    elif idir == 4: fen[2] = nj
    elif idir == 5: fen[5] = 1
    else: fen[4] = nk
    return idir, fen


def append_bcs(p, r, bcs):
    """Append the boundary-condition table to param_int / param_real (FastC._BCcompact).

    ``bcs`` is a list of ``(bctype_name, face_or_window, data)``: ``face_or_window`` is either a
    face name ('imin' ... 'kmax') or an ``(idir, fen[6])`` tuple; ``data`` a list of 1-D float
    arrays (the BC's DataArray_t children) or None.
    Returns the extended (param_int, param_real).
    """
    nb = len(bcs)
    pt_bcs_int = len(p)
    pt_bcs_real = len(r)
    ints = [nb] + [0] * (2 * nb)
    reals = []
    size_int = 1 + 2 * nb
    for n, (name, where, data) in enumerate(bcs):
        ints[1 + n] = size_int + pt_bcs_int
        ints[1 + n + nb] = len(reals) + pt_bcs_real
        if isinstance(where, str):
            idir, fen = face_window(p, where)
        else:
            idir, fen = where
        data = [] if data is None else [np.asarray(d, dtype=np.float64).ravel() for d in data]
        rec = [BC_TAGS[name], idir] + list(fen) + [len(data)]
        ctot = 0
        for d in data:
            ctot += d.size
            rec.append(ctot)
            reals.extend(d.tolist())
        ints.extend(rec)
        size_int += 9 + len(data)
    p2 = np.concatenate([p, np.asarray(ints, dtype=np.int32)])
    p2[PT_BC] = pt_bcs_int
    r2 = np.concatenate([r, np.asarray(reals, dtype=np.float64)])
    return p2, r2


def ndimdx(p):
    return int(p[NDIMDX])


def inddm(p, i, j, k):
    """0-based linear index of cell (i, j, k) (1-based interior numbering; formule_param.h)."""
    return ((i + int(p[NIJK + 3]) - 1) + (j + int(p[NIJK + 3]) - 1) * int(p[NIJK])
            + (k + int(p[NIJK + 4]) - 1) * int(p[NIJK]) * int(p[NIJK + 1]))
