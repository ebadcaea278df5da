"""Zone connectivity: 1-to-1 matching faces between structured zones, as injection transfers.

The reference fills the ghost cells of a zone face that abuts another zone from the donor
zone's interior cells (FastC/Com/setInterpTransfersFast.cpp:440-760, setInterpTransfersIntra:
for each raccord, rcv[rcvPts[n]] = sum coef * dnr[donorPts[...]]; a 1-to-1 match is an
injection: interpolation type 1 with coefficient 1, as the reference's hand-built connectivity
tree of Fast/FastS/test/tgvMpi.py:135-140 writes it).  Raccords come from the face adjacency of
the zones, from the GridConnectivity1to1 nodes of a tree, from the ``ID_*`` subregions of a
Cassiopee connectivity tree (``tc``), or from its flattened form ``param_int_tc`` /
``param_real_tc``; Connector, which builds ``tc`` trees from meshes, is not part of this build.

A raccord is (receiver zone, donor zone, receiver cell indices, donor cell indices), the indices
being 0-based linear indices into the zones' field arrays (``inddm``).  The ghost slab of the
receiver face (all ghost layers, over the face's interior extent) receives the donor cells at
the same distance from the interface on the other side.
"""
import numpy as np

from . import params as P

FACES = ("imin", "imax", "jmin", "jmax", "kmin", "kmax")


class Raccord:
    def __init__(self, zr, zd, rcv, dnr):
        self.zr, self.zd = int(zr), int(zd)
        self.rcv = np.ascontiguousarray(rcv, dtype=np.int64)
        self.dnr = np.ascontiguousarray(dnr, dtype=np.int64)
        assert self.rcv.shape == self.dnr.shape


def _lin(p, i, j, k):
    """0-based linear field index of 1-based cells (arrays broadcast)."""
    ni, nj, nf, nfk = int(p[P.NIJK]), int(p[P.NIJK + 1]), int(p[P.NIJK + 3]), int(p[P.NIJK + 4])
    return (i + nf - 1) + (j + nf - 1) * ni + (k + nfk - 1) * ni * nj


def match(pr, face_r, pd, face_d, offset=(0, 0), window=None):
    """Injection raccord filling the ghost slab of face ``face_r`` of the receiver (descriptor
    ``pr``) from the donor (``pd``) whose face ``face_d`` touches it.  Axes are aligned (CGNS
    Transform = identity); ``offset`` shifts the two tangential indices from receiver to donor;
    ``window`` restricts the two tangential ranges of the receiver ((lo, hi) 1-based cells,
    default the whole face).  Returns (rcv, dnr) index arrays."""
    dr, dd = FACES.index(face_r) // 2, FACES.index(face_d) // 2
    if dr != dd or (FACES.index(face_r) % 2) == (FACES.index(face_d) % 2):
        raise NotImplementedError("only opposite faces of aligned zones are matched")
    nr = [int(pr[P.IJKV]), int(pr[P.IJKV + 1]), int(pr[P.IJKV + 2])]
    nd_ = [int(pd[P.IJKV]), int(pd[P.IJKV + 1]), int(pd[P.IJKV + 2])]
    ng = int(pr[P.NIJK + 3]) if dr < 2 else int(pr[P.NIJK + 4])
    tang = [a for a in range(3) if a != dr]
    rng = [np.arange(1, nr[a] + 1) for a in range(3)]
    if window is not None:
        for t, (lo, hi) in zip(tang, window):
            if lo < 1 or hi > nr[t] or lo > hi:
                raise ValueError("receiver window outside the face")
            rng[t] = np.arange(lo, hi + 1)
    up = FACES.index(face_r) % 2 == 1
    # ghost layers of the receiver, nearest first
    rng[dr] = np.arange(nr[dr] + 1, nr[dr] + 1 + ng) if up else np.arange(0, -ng, -1)
    I, J, K = np.meshgrid(rng[0], rng[1], rng[2], indexing="ij")
    R = [I, J, K]
    D = [I.copy(), J.copy(), K.copy()]
    # distance into the donor: ghost g (1-based from the interface) -> donor cell g from its face
    if up:
        D[dr] = R[dr] - nr[dr]                  # receiver imax ghost ni+g -> donor cell g
    else:
        D[dr] = nd_[dr] + R[dr]                 # receiver imin ghost 1-g  -> donor cell nd-g+1
    for t, o in zip(tang, offset):
        D[t] = R[t] + o
    for a in range(3):
        if D[a].min() < 1 - (ng if a == dr else 0) or D[a].max() > nd_[a]:
            raise ValueError("donor window outside the donor zone")
    rcv = _lin(pr, R[0], R[1], R[2]).ravel(order="F")
    dnr = _lin(pd, D[0], D[1], D[2]).ravel(order="F")
    return rcv, dnr


def tiling(nx, ny, nz, periodic=(False, False, False)):
    """Blocks of a Cartesian tiling: block (a, b, c) has nx[a] x ny[b] x nz[c] cells.
    Returns (blocks, raccords, outer) where blocks[n] = (a, b, c), raccords is a list of
    (receiver, receiver face, donor, donor face) and outer lists the (block, face) pairs on
    the outer boundary (not periodic)."""
    dims = (len(nx), len(ny), len(nz))
    blocks = [(a, b, c) for c in range(dims[2]) for b in range(dims[1]) for a in range(dims[0])]
    num = {blk: n for n, blk in enumerate(blocks)}
    rac, outer = [], []
    for n, blk in enumerate(blocks):
        for f, face in enumerate(FACES):
            d, up = f // 2, f % 2 == 1
            nb = list(blk)
            nb[d] += 1 if up else -1
            if 0 <= nb[d] < dims[d]:
                rac.append((n, face, num[tuple(nb)], FACES[f ^ 1]))
            elif periodic[d]:
                nb[d] %= dims[d]
                rac.append((n, face, num[tuple(nb)], FACES[f ^ 1]))
            else:
                outer.append((n, face))
    return blocks, rac, outer


def from_pointranges(pr, rng_r, pd, rng_d, transform=(1, 2, 3)):
    """Raccord of a CGNS GridConnectivity1to1 (PointRange / PointRangeDonor in the ghost
    inclusive vertex numbering of FastS trees, identity Transform): (rcv, dnr)."""
    if tuple(int(v) for v in transform) != (1, 2, 3):
        raise NotImplementedError("only GridConnectivity1to1 with Transform [1, 2, 3] is supported")
    flat = lambda r: [int(r[0][0]), int(r[0][1]), int(r[1][0]), int(r[1][1]), int(r[2][0]), int(r[2][1])]
    idr, fenr = P.get_range(flat(rng_r), pr)
    idd, fend = P.get_range(flat(rng_d), pd)
    if idr == 0 or idd == 0:
        raise ValueError("GridConnectivity1to1 window is not on a zone face")
    dr = (idr - 1) // 2
    tang = [a for a in range(3) if a != dr]
    window = [(fenr[2 * t], fenr[2 * t + 1]) for t in tang]
    offset = [fend[2 * t] - fenr[2 * t] for t in tang]
    return match(pr, FACES[idr - 1], pd, FACES[idd - 1], offset, window)


def build(zones, pairs):
    """Raccords for a list of HostZones and (receiver, face, donor, face) pairs."""
    out = []
    for zr, fr, zd, fd in pairs:
        rcv, dnr = match(zones[zr].param_int, fr, zones[zd].param_int, fd)
        out.append(Raccord(zr, zd, rcv, dnr))
    return out


def apply(zones, raccords, levels):
    """Host-side injection (test reference): levels[n] is the field array of zone n."""
    for r in raccords:
        ndr, ndd = zones[r.zr].nd, zones[r.zd].nd
        for e in range(5):
            levels[r.zr][r.rcv + e * ndr] = levels[r.zd][r.dnr + e * ndd]


# ---------------------------------------------------------------------------------------------
# Cassiopee connectivity trees (tc) and their flattened form param_int_tc / param_real_tc

INJECTION = 1          # InterpolantsType of an injection point (tgvMpi.py:138-139)
NTAB = 18              # integer tables per raccord (ntab_int, setInterpTransfersFast.cpp:497)
NEXPLOC = 27           # integers per raccord after the tables, read for local explicit steps
# table numbers read by setInterpTransfersIntra (setInterpTransfersFast.cpp:640-760); table t of
# raccord r sits at ech + 4 + 2 * timelevel + r + nrac * t
T_NPTS, T_NINTERP, T_IBCTYPE, T_NOD, T_DONOR, T_NTYPE, T_COEF = 0, 1, 3, 5, 6, 7, 8
T_NRCV, T_NOR, T_RCV, T_NVARS, T_ROT, T_NFLU = 10, 11, 12, 13, 14, 16


def from_tc(tc, index, sizes):
    """Injection raccords of a Cassiopee connectivity tree.

    The tree holds one zone per donor zone (same names as in t).  Each ``ID_<name>``
    ZoneSubRegion_t of a donor zone carries the receiver zone's name as its value, ZoneRole
    'Donor', PointList = donor cells and PointListDonor = receiver cells (storage on the donor
    side, setInterpTransfersFast.cpp:753-756), both 0-based linear indices into the zones'
    field arrays, and InterpolantsType / InterpolantsDonor, the type and the coefficients of
    each point (Fast/FastS/test/tgvMpi.py:132-195 builds such a tree by hand).  ``index`` maps
    zone names to zone numbers, ``sizes[n]`` is the field-array length of zone n.  Immersed-
    boundary subregions (``IBCD_*``) and interpolation types other than injection with
    coefficient 1 raise NotImplementedError."""
    from . import Internal
    racs = []
    for zd in Internal.getZones(tc):
        if zd[0] not in index:
            raise ValueError("connectivity tree zone %r is not a zone of t" % zd[0])
        d = index[zd[0]]
        for sr in Internal.getNodesFromType1(zd, "ZoneSubRegion_t"):
            if sr[0].startswith("IBCD_"):
                raise NotImplementedError("immersed-boundary raccord %s of zone %s is not supported"
                                          % (sr[0], zd[0]))
            if not sr[0].startswith("ID_"):
                continue
            role = Internal.getNodeFromName1(sr, "ZoneRole")
            if role is not None and Internal.getValue(role) != "Donor":
                continue
            rname = Internal.getValue(sr)
            if rname not in index:
                raise ValueError("raccord %s of zone %s: unknown receiver zone %r" % (sr[0], zd[0], rname))
            r = index[rname]
            get = lambda name: Internal.getNodeFromName1(sr, name)
            if get("PointList") is None or get("PointListDonor") is None:
                raise ValueError("raccord %s of zone %s has no PointList / PointListDonor" % (sr[0], zd[0]))
            dnr = np.asarray(get("PointList")[1]).ravel().astype(np.int64)
            rcv = np.asarray(get("PointListDonor")[1]).ravel().astype(np.int64)
            typ = get("InterpolantsType")
            typ = np.full(dnr.size, INJECTION) if typ is None else np.asarray(typ[1]).ravel()
            coef = get("InterpolantsDonor")
            coef = np.ones(dnr.size) if coef is None else np.asarray(coef[1], dtype=np.float64).ravel()
            _check_injection("raccord %s of zone %s" % (sr[0], zd[0]), rcv, dnr, typ, coef,
                             sizes[r], sizes[d])
            racs.append(Raccord(r, d, rcv, dnr))
    return racs


def _check_injection(what, rcv, dnr, typ, coef, nd_r, nd_d):
    if rcv.size != dnr.size or typ.size != dnr.size:
        raise ValueError("%s: %d receiver points, %d donor points, %d types"
                         % (what, rcv.size, dnr.size, typ.size))
    bad = typ[typ != INJECTION]
    if bad.size:
        raise NotImplementedError("%s: interpolation type %d (only injection, type %d, is supported)"
                                  % (what, int(bad[0]), INJECTION))
    if coef.size != dnr.size or not np.all(coef == 1.0):
        raise NotImplementedError("%s: injection coefficients other than 1" % what)
    if rcv.size and (rcv.min() < 0 or rcv.max() >= nd_r or dnr.min() < 0 or dnr.max() >= nd_d):
        raise ValueError("%s: point index outside the zone's field arrays" % what)


def pack_tc(raccords):
    """(param_int_tc, param_real_tc) of injection raccords, laid out where
    setInterpTransfersIntra reads them (setInterpTransfersFast.cpp:440-760): no immersed-boundary
    or MPI part (sizecomIBC = sizecomID = 0), the intra exchange header at ech = param_int_tc[4]
    (nrac, no unsteady raccord, timelevel 0), then NTAB tables of nrac entries, NEXPLOC zeros per
    raccord, and per raccord its type groups [1, n, types...], donor and receiver points; the
    coefficients (1 per point) in param_real_tc.  Entries the reader does not use are 0."""
    nrac = len(raccords)
    ech = 5
    tab0 = ech + 4
    data = tab0 + NTAB * nrac + NEXPLOC * nrac
    size = data + sum(2 + 3 * r.rcv.size for r in raccords)
    pit = np.zeros(size, dtype=np.int64)
    prt = np.zeros(sum(r.rcv.size for r in raccords), dtype=np.float64)
    pit[4] = ech
    pit[ech + 1] = nrac
    pit[ech + 2] = 0
    pit[ech + 3] = 0
    pos, rpos = data, 0
    for n, r in enumerate(raccords):
        npts = r.rcv.size
        tab = lambda t, v: pit.__setitem__(tab0 + n + nrac * t, v)
        tab(T_NPTS, npts); tab(T_NINTERP, npts); tab(T_IBCTYPE, -1)
        tab(T_NOD, r.zd); tab(T_NOR, r.zr); tab(T_NRCV, npts)
        tab(T_NVARS, 5); tab(T_ROT, 0); tab(T_NFLU, 0)
        tab(T_NTYPE, pos)
        pit[pos] = 1; pit[pos + 1] = npts
        pit[pos + 2:pos + 2 + npts] = INJECTION
        pos += 2 + npts
        tab(T_DONOR, pos); pit[pos:pos + npts] = r.dnr; pos += npts
        tab(T_RCV, pos); pit[pos:pos + npts] = r.rcv; pos += npts
        tab(T_COEF, rpos); prt[rpos:rpos + npts] = 1.0; rpos += npts
    if pit.max(initial=0) < 2 ** 31:
        pit = pit.astype(np.int32)
    return pit, prt


def from_param_int_tc(param_int_tc, param_real_tc, sizes):
    """Injection raccords of a flattened connectivity, read as setInterpTransfersIntra reads it
    for the intra-process transfers (setInterpTransfersFast.cpp:461-470 for the header,
    :635-760 for a raccord).  ``sizes[n]`` is the field-array length of zone n.  Unsteady raccords,
    immersed-boundary raccords, azimuthal periodicity, other variable counts, conservative
    flux corrections and types other than injection raise NotImplementedError."""
    p = np.asarray(param_int_tc).ravel().astype(np.int64)
    c = np.asarray(param_real_tc, dtype=np.float64).ravel()
    size_ibc = int(p[2])
    size_id = int(p[3 + size_ibc])
    ech = int(p[1 + size_ibc + size_id + 3])
    nrac, nrac_inst, timelevel = int(p[ech + 1]), int(p[ech + 2]), int(p[ech + 3])
    if nrac_inst:
        raise NotImplementedError("unsteady raccords are not supported")
    racs = []
    for n in range(nrac):
        sh = ech + 4 + 2 * timelevel + n
        tab = lambda t: int(p[sh + nrac * t])
        what = "raccord %d" % n
        if tab(T_IBCTYPE) >= 0:
            raise NotImplementedError("%s: immersed-boundary raccord (type %d)" % (what, tab(T_IBCTYPE)))
        if tab(T_NVARS) != 5:
            raise NotImplementedError("%s: %d transferred variables (5 are supported)" % (what, tab(T_NVARS)))
        if tab(T_ROT):
            raise NotImplementedError("%s: azimuthal periodicity" % what)
        if tab(T_NFLU):
            raise NotImplementedError("%s: conservative flux correction" % what)
        npts = tab(T_NRCV)
        pos = tab(T_NTYPE)
        ngroups = int(p[pos])
        counts = p[pos + 1:pos + 1 + ngroups]
        if int(counts.sum()) != npts:
            raise ValueError("%s: type groups hold %d points, the raccord %d" % (what, int(counts.sum()), npts))
        typ = p[pos + 1 + ngroups:pos + 1 + ngroups + npts]
        dnr = p[tab(T_DONOR):tab(T_DONOR) + npts].copy()
        rcv = p[tab(T_RCV):tab(T_RCV) + npts].copy()
        coef = c[tab(T_COEF):tab(T_COEF) + tab(T_NINTERP)]
        zr, zd = tab(T_NOR), tab(T_NOD)
        if not (0 <= zr < len(sizes) and 0 <= zd < len(sizes)):
            raise ValueError("%s: zone numbers %d <- %d outside the %d zones" % (what, zr, zd, len(sizes)))
        _check_injection(what, rcv, dnr, typ, coef, sizes[zr], sizes[zd])
        racs.append(Raccord(zr, zd, rcv, dnr))
    return racs
