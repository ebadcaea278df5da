"""Several zones joined by 1-to-1 matches (fast_amd.connect), driven in the order of gsdr3.cpp
(Compute/gsdr3.cpp:340-444): the right-hand side and LU-SGS of every zone, then the transfers
(FastC/Com/setInterpTransfersFast.cpp), then the physical boundary conditions and the edge/corner
fill of every zone.

A periodic box cut into blocks must give, with an explicit RK3 Euler scheme (whose stencils
only reach the face ghosts, never the edge ghosts), exactly the single-zone result."""
import numpy as np
import pytest

import oracle
from fast_amd import cases, connect, params as P
from tests import helpers as H

NX, NY, NZ = (3, 5), (2, 4), (4,)


@pytest.fixture(scope="module")
def orc():
    return oracle.Oracle()


def ro_cl(z, nitcfg):
    """gsdr3.cpp:152-173: the level that receives the boundary conditions."""
    if nitcfg == 1 or int(z.param_int[P.ITYPCP]) < 2 or nitcfg % 2 == 1:
        return z.ro_p1
    return z.ro


def oracle_step(orc, zones, racs, nsub, first_it):
    for it in range(1, nsub + 1):
        for z in zones:
            orc.substep(z, it, first_it=first_it, stop=3)
        levels = [ro_cl(z, it) for z in zones]
        connect.apply(zones, racs, levels)
        for z, lv in zip(zones, levels):
            orc.apply_bc(z, lv)
    for z in zones:
        H.rotate(z)


def init(orc, zones, racs):
    for z in zones:
        orc.init_metric(z)
    connect.apply(zones, racs, [z.ro for z in zones])
    for z in zones:
        orc.apply_bc(z, z.ro)
        z.ro_m1[:] = z.ro; z.ro_p1[:] = z.ro


def blocks_of(W_single, nx, ny, nz):
    ox = np.concatenate([[0], np.cumsum(nx)]); oy = np.concatenate([[0], np.cumsum(ny)])
    oz = np.concatenate([[0], np.cumsum(nz)])
    for c in range(len(nz)):
        for b in range(len(ny)):
            for a in range(len(nx)):
                yield W_single[:, oz[c]:oz[c + 1], oy[b]:oy[b + 1], ox[a]:ox[a + 1]]


def test_match_indices():
    """Receiver ghosts of a face map to the donor cells on the other side of the interface."""
    num = H.numerics("roe_min", "explicit")
    zones, pairs, _ = cases.tiled_box((3, 4), (2,), (2,), num)
    r, d = connect.match(zones[0].param_int, "imax", zones[1].param_int, "imin")
    p0, p1 = zones[0].param_int, zones[1].param_int
    assert r[0] == connect._lin(p0, 4, 1, 1) and d[0] == connect._lin(p1, 1, 1, 1)
    r, d = connect.match(zones[1].param_int, "imin", zones[0].param_int, "imax")
    assert r[0] == connect._lin(p1, 0, 1, 1) and d[0] == connect._lin(p0, 3, 1, 1)
    assert len(r) == 2 * 2 * 2


@pytest.mark.parametrize("scheme", ["roe_min", "roe_kap", "ausmpred", "senseur"])
def test_split_box_equals_single_zone(orc, scheme):
    num = H.numerics(scheme, "explicit")
    zones, pairs, W = cases.tiled_box(NX, NY, NZ, num, flow="euler")
    racs = connect.build(zones, pairs)
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
