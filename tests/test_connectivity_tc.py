"""Connectivity from Cassiopee ``tc`` trees and from the flattened ``param_int_tc`` /
``param_real_tc`` (FastC/Com/setInterpTransfersFast.cpp:440-760).

The ``tc`` trees here are built the way the reference's own test builds one by hand
(Fast/FastS/test/tgvMpi.py:132-195: an ``ID_<receiver>`` ZoneSubRegion_t under each donor zone,
PointList = donor cells, PointListDonor = receiver ghost cells, InterpolantsType 1,
InterpolantsDonor 1), with the points computed from global cell coordinates of the box rather
than from the face matching code, so that they check it."""
import numpy as np
import pytest

import oracle
from fast_amd import Internal, cases, connect, PyTree as FP
from tests import helpers as H
from tests.test_multizone import blocks_of, init, oracle_step

NX, NY, NZ = (3, 5), (2, 4), (4,)


def ghost_points(nx, ny, nz, ng=2):
    """For every block of a periodic box split into nx x ny x nz blocks: the ghost cells of
    its six faces (ng layers, over the face's interior extent) and the cells that fill them,
    found from the global cell they stand for.  -> {(receiver, donor): (rcv, dnr)}."""
    dims = [nx, ny, nz]
    off = [np.concatenate([[0], np.cumsum(n)]) for n in dims]
    tot = [o[-1] for o in off]
    blocks = [(a, b, c) for c in range(len(nz)) for b in range(len(ny)) for a in range(len(nx))]
    num = {blk: n for n, blk in enumerate(blocks)}
    out = {}
    for n, blk in enumerate(blocks):
        ext = [dims[d][blk[d]] for d in range(3)]
        for d in range(3):
            for layer in list(range(-ng, 0)) + list(range(ext[d], ext[d] + ng)):
                rng = [np.arange(ext[0]), np.arange(ext[1]), np.arange(ext[2])]
                rng[d] = np.array([layer])
                I, J, K = np.meshgrid(*rng, indexing="ij")
                loc = [I.ravel(), J.ravel(), K.ravel()]            # 0-based local cells
                glob = [(loc[e] + off[e][blk[e]]) % tot[e] for e in range(3)]
                owner = [np.searchsorted(off[e], glob[e], side="right") - 1 for e in range(3)]
                for p in range(loc[0].size):
                    ob = tuple(int(owner[e][p]) for e in range(3))
                    m = num[ob]
                    dl = [int(glob[e][p] - off[e][ob[e]]) for e in range(3)]
                    edn = [dims[e][ob[e]] + 2 * ng for e in range(3)]
                    er = [ext[e] + 2 * ng for e in range(3)]
                    ir = (loc[0][p] + ng) + (loc[1][p] + ng) * er[0] + (loc[2][p] + ng) * er[0] * er[1]
                    idn = (dl[0] + ng) + (dl[1] + ng) * edn[0] + (dl[2] + ng) * edn[0] * edn[1]
                    r, dd = out.setdefault((n, m), ([], []))
                    r.append(int(ir)); dd.append(int(idn))
    return {k: (np.array(v[0]), np.array(v[1])) for k, v in out.items()}


def tc_tree(names, points, typ=1, coef=1.0):
    """A connectivity tree: zone names of t, one ID_<receiver> subregion per (receiver, donor)."""
    tc = Internal.newCGNSTree()
    b = Internal.newCGNSBase("Base", parent=tc)
    zs = [Internal.createChild(b, nm, "Zone_t", np.zeros((3, 3), dtype=np.int32)) for nm in names]
    for (r, d), (rcv, dnr) in sorted(points.items()):
        sr = Internal.createChild(zs[d], "ID_" + names[r], "ZoneSubRegion_t", names[r])
        Internal.createChild(sr, "ZoneRole", "DataArray_t", "Donor")
        Internal.createChild(sr, "GridLocation", "GridLocation_t", "CellCenter")
        Internal.createChild(sr, "InterpolantsDonor", "DataArray_t", np.full(rcv.size, coef))
        Internal.createChild(sr, "InterpolantsType", "DataArray_t", np.full(rcv.size, typ, dtype=np.int32))
        Internal.createChild(sr, "PointListDonor", "DataArray_t", rcv.astype(np.int32))
        Internal.createChild(sr, "PointList", "DataArray_t", dnr.astype(np.int32))
    return tc


def pairs_of(racs):
    """The (receiver zone, donor zone, receiver cell, donor cell) set of a raccord list."""
    return {(r.zr, r.zd, int(a), int(b)) for r in racs for a, b in zip(r.rcv, r.dnr)}


def setup():
    num = H.numerics("ausmpred", "explicit", **H.SINGLE_BLOCK)
    zones, pairs, W = cases.tiled_box(NX, NY, NZ, num, flow="euler")
    names = [z.name for z in zones]
    return num, zones, pairs, W, names


def test_tc_tree_gives_the_tiling_raccords():
    """The raccords read from a tc tree are, point for point, those of the face matching."""
    num, zones, pairs, W, names = setup()
    tc = tc_tree(names, ghost_points(NX, NY, NZ))
    racs = connect.from_tc(tc, {nm: n for n, nm in enumerate(names)}, [z.nd for z in zones])
    assert len(racs) == len(ghost_points(NX, NY, NZ))
    ref = pairs_of(connect.build(zones, pairs))
    assert pairs_of(racs) == ref
    # every receiver ghost is filled once
    assert len({(a, b) for a, _, b, _ in ref}) == len(ref)


def test_param_int_tc_roundtrip_and_header():
    """pack_tc lays the raccords out where setInterpTransfersIntra reads them; reading the
    flat arrays back gives the same raccords."""
    num, zones, pairs, W, names = setup()
    racs = connect.build(zones, pairs)
    pit, prt = connect.pack_tc(racs)
    assert pit[2] == 0 and pit[3] == 0                  # sizecomIBC, sizecomID
    ech = int(pit[1 + 0 + 0 + 3])
    nrac = int(pit[ech + 1])
    assert nrac == len(racs) and pit[ech + 2] == 0 and pit[ech + 3] == 0
    for n, r in enumerate(racs):
        sh = ech + 4 + n
        assert pit[sh + nrac * connect.T_NOD] == r.zd and pit[sh + nrac * connect.T_NOR] == r.zr
        assert pit[sh + nrac * connect.T_IBCTYPE] < 0
        pos = int(pit[sh + nrac * connect.T_RCV])
        assert np.array_equal(pit[pos:pos + r.rcv.size], r.rcv)
        pos = int(pit[sh + nrac * connect.T_DONOR])
        assert np.array_equal(pit[pos:pos + r.dnr.size], r.dnr)
    back = connect.from_param_int_tc(pit, prt, [z.nd for z in zones])
    assert [(r.zr, r.zd) for r in back] == [(r.zr, r.zd) for r in racs]
    for a, b in zip(back, racs):
        assert np.array_equal(a.rcv, b.rcv) and np.array_equal(a.dnr, b.dnr)


def test_tc_unsupported_forms_are_refused():
    num, zones, pairs, W, names = setup()
    index = {nm: n for n, nm in enumerate(names)}
    sizes = [z.nd for z in zones]
    pts = ghost_points(NX, NY, NZ)
    with pytest.raises(NotImplementedError, match="interpolation type 2"):
        connect.from_tc(tc_tree(names, pts, typ=2), index, sizes)
    with pytest.raises(NotImplementedError, match="coefficients"):
        connect.from_tc(tc_tree(names, pts, coef=0.5), index, sizes)
    tc = tc_tree(names, pts)
    z0 = Internal.getZones(tc)[0]
    Internal.createChild(z0, "IBCD_3_" + names[1], "ZoneSubRegion_t", names[1])
    with pytest.raises(NotImplementedError, match="immersed"):
        connect.from_tc(tc, index, sizes)
    (r, d), (rcv, dnr) = next(iter(sorted(pts.items())))
    bad = {(r, d): (rcv, dnr + sizes[d])}
    with pytest.raises(ValueError, match="outside"):
        connect.from_tc(tc_tree(names, bad), index, sizes)
    racs = connect.build(zones, pairs)
    pit, prt = connect.pack_tc(racs)
    ech, nrac = int(pit[4]), len(racs)
    for table, value, msg in ((connect.T_IBCTYPE, 0, "immersed"), (connect.T_ROT, 1, "azimuthal"),
                              (connect.T_NVARS, 6, "variables"), (connect.T_NFLU, 1, "flux")):
        p = pit.copy()
        p[ech + 4 + nrac * table] = value
        with pytest.raises(NotImplementedError, match=msg):
            connect.from_param_int_tc(p, prt, sizes)
    p = pit.copy(); p[ech + 2] = 1
    with pytest.raises(NotImplementedError, match="unsteady"):
        connect.from_param_int_tc(p, prt, sizes)


def test_split_box_from_tc_equals_single_zone():
    """Oracle: the split box joined by the raccords of a tc tree steps exactly as one zone."""
    orc = oracle.Oracle()
    num, zones, pairs, W, names = setup()
    tc = tc_tree(names, ghost_points(NX, NY, NZ))
    racs = connect.from_tc(tc, {nm: n for n, nm in enumerate(names)}, [z.nd for z in zones])
    init(orc, zones, racs)
    whole = cases.whole_box(NX, NY, NZ, W, num, flow="euler")
    orc.init_metric(whole)
    orc.apply_bc(whole, whole.ro)
    whole.ro_m1[:] = whole.ro; whole.ro_p1[:] = whole.ro
    for step in range(2):
        oracle_step(orc, zones, racs, 3, 0 if step == 0 else 1)
        for it in range(1, 4):
            orc.substep(whole, it, first_it=0 if step == 0 else 1)
        H.rotate(whole)
    Wg = whole.interior(whole.ro)
    for z, blk in zip(zones, blocks_of(Wg, NX, NY, NZ)):
        assert np.array_equal(z.interior(z.ro), blk), z.name


@pytest.mark.gpu
def test_warmup_with_tc_tree_matches_single_zone():
    """warmup(t, tc) + _compute with the zones joined only by a tc tree (explicit RK3, Euler)
    equal the single periodic zone stepped by the oracle, bit for bit; HOOK keeps the
    flattened connectivity as the reference does."""
    from tests.test_pytree import tree_of
    orc = oracle.Oracle()
    num, zones, pairs, W, names = setup()
    t = tree_of(zones, num, model="Euler")
    tc = tc_tree(names, ghost_points(NX, NY, NZ))
    t, tc2, metrics = FP.warmup(t, tc)
    assert tc2 is tc
    racs = connect.from_param_int_tc(FP.HOOK["param_int_tc"], FP.HOOK["param_real_tc"],
                                     [z.nd for z in zones])
    assert pairs_of(racs) == pairs_of(connect.build(zones, pairs))
    nit = 2
    FP._compute(t, metrics, 0, NIT=nit)
    whole = cases.whole_box(NX, NY, NZ, W, num, flow="euler")
    orc.init_metric(whole)
    orc.apply_bc(whole, whole.ro)
    whole.ro_m1[:] = whole.ro; whole.ro_p1[:] = whole.ro
    for step in range(nit):
        for it in range(1, 4):
            orc.substep(whole, it, first_it=0 if step == 0 else 1)
        H.rotate(whole)
    Wg = whole.interior(whole.ro)
    for z, blk in zip(Internal.getZones(t), blocks_of(Wg, NX, NY, NZ)):
        fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
        got = np.stack([Internal.getNodeFromName1(fs, v)[1][2:-2, 2:-2, 2:-2].transpose(2, 1, 0)
                        for v in FP.VARS])
        assert np.array_equal(got, blk), z[0]
