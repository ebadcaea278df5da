"""Zone connectivity: 1-to-1 matching faces between structured zones, as injection transfers.

The reference fills the ghost cells of a zone face that abuts another zone from the donor
zone's interior cells (FastC/Com/setInterpTransfersFast.cpp:440-760, setInterpTransfersIntra:
for each raccord, rcv[rcvPts[n]] = sum coef * dnr[donorPts[...]]; a 1-to-1 match is an
injection with coefficient 1, interpolation type 0).  The raccords themselves are built by
Cassiopee's Connector (not part of this build); here they are built directly from the face
adjacency of the zones.

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
