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
    num = H.numerics(scheme, "explicit", **H.SINGLE_BLOCK)
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


# ---------------------------------------------------------------------------------------
# GPU: fm_transfers_set / fm_compute with several zones against the oracle driven the same way
def gpu_run(lib, zones, racs, nsub, nsteps, modulo_verif=0):
    ctx = lib.Context(0)
    ids = [ctx.add_zone(z) for z in zones]
    ctx.set_transfers(racs)
    ctx.transfer(0)
    for i in ids:
        ctx.apply_bc(i, 0)
    for i, z in zip(ids, zones):
        t = np.zeros_like(z.ro)
        ctx.get(i, lib.FIELD_RO, t)
        ctx.set(i, lib.FIELD_RO_M1, t); ctx.set(i, lib.FIELD_RO_P1, t)
    ctx.compute(0, 1, nsub, nit_c=nsteps, modulo_verif=modulo_verif, first_it=0)
    ctx.sync()
    for i, z in zip(ids, zones):
        ctx.pull(i, z)
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["roe_min", "ausmpred", "senseur"])
def test_gpu_split_box_explicit(orc, scheme):
    """Euler RK3 on 6 blocks of a periodic box: GPU zones joined by transfers equal the oracle
    driven the same way, which equals the single zone bit for bit."""
    from fast_amd import _lib
    num = H.numerics(scheme, "explicit")
    zo, pairs, _ = cases.tiled_box(NX, NY, NZ, num, flow="euler")
    zg, _, _ = cases.tiled_box(NX, NY, NZ, num, flow="euler")
    racs = connect.build(zo, pairs)
    init(orc, zo, racs)
    for step in range(2):
        oracle_step(orc, zo, racs, 3, 0 if step == 0 else 1)
    gpu_run(_lib, zg, racs, 3, 2)
    for a, b in zip(zo, zg):
        assert np.array_equal(a.ro, b.ro), a.name


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["bdf", "bdf_ss2"])
def test_gpu_tiles_implicit_walls(orc, mode):
    """Laminar implicit BDF with LU-SGS per zone (block-Jacobi across zones, as across the
    reference's thread boxes) on uneven blocks, periodic in i and k, viscous walls in j."""
    from fast_amd import _lib
    num = H.numerics("ausmpred", mode)
    nx, ny, nz = (5, 7), (6, 5), (4, 3)
    mk = lambda: cases.tiled_box(nx, ny, nz, num, flow="lamin", periodic=(True, False, True),
                                 bcs_outer="BCWallViscous", seed=3)
    zo, pairs, _ = mk()
    zg, _, _ = mk()
    racs = connect.build(zo, pairs)
    init(orc, zo, racs)
    nsub = int(zo[0].param_int[P.NSSITER])
    nsteps = 2
    for step in range(nsteps):
        for it in range(1, nsub + 1):
            verif = step % 2 == 0
            lexit = 1 if it == nsub else 0
            if lexit and not verif:
                continue
            for z in zo:
                orc.substep(z, it, first_it=0 if step == 0 else 1, lssiter_verif=int(verif),
                            lexit_lu=lexit, stop=3)
            if lexit:
                continue
            levels = [ro_cl(z, it) for z in zo]
            connect.apply(zo, racs, levels)
            for z, lv in zip(zo, levels):
                orc.apply_bc(z, lv)
        for z in zo:
            H.rotate(z)
    gpu_run(_lib, zg, racs, nsub, nsteps, modulo_verif=2)
    for a, b in zip(zo, zg):
        assert np.array_equal(a.ro, b.ro), a.name
        assert np.array_equal(a.ro_m1, b.ro_m1), a.name


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["bdf", "bdf_ss2", "local"])
def test_gpu_zone_batch_equals_zone_loop(mode):
    """The batched LU-SGS over all zones (one work list of (zone, pencil) pairs, the default
    streaming schedule) gives the per-zone sweeps bit for bit, on uneven blocks and on unconnected
    zones of both mesh kinds.  Compared: the three time levels, drodm, mu and coe after seven time
    steps (run eagerly, then captured, then replayed; dW stays in the rows inside fm_compute and is
    copied back when drodm is read), a state edit of one zone and three more steps."""
    from fast_amd import _lib
    num = H.numerics("ausmpred", mode)

    def run(batch, make):
        zones, racs = make()
        ctx = _lib.Context(0)
        ctx.set_option(_lib.OPT_ZONE_BATCH, batch)
        ids = [ctx.add_zone(z) for z in zones]
        ctx.set_transfers(racs)
        ctx.transfer(0)
        for i in ids:
            ctx.apply_bc(i, 0)
        for i, z in zip(ids, zones):
            t = np.zeros_like(z.ro)
            ctx.get(i, _lib.FIELD_RO, t)
            ctx.set(i, _lib.FIELD_RO_M1, t); ctx.set(i, _lib.FIELD_RO_P1, t)
        nsub = int(zones[0].param_int[P.NSSITER])
        ctx.compute(0, 1, nsub, nit_c=7, modulo_verif=2, first_it=0)
        # an edit of one zone's state between calls: its rows no longer hold it
        t = np.zeros_like(zones[0].ro)
        ctx.get(ids[0], _lib.FIELD_RO, t)
        ctx.set(ids[0], _lib.FIELD_RO, t * (1.0 + 1e-3))
        ctx.compute(1, 1, nsub, nit_c=3, modulo_verif=2, first_it=0)
        ctx.sync()
        for i, z in zip(ids, zones):
            ctx.pull(i, z)
        ctx.close()
        return zones

    def tiles():
        zones, pairs, _ = cases.tiled_box((9, 17, 6), (12, 7), (10, 8), num, flow="lamin",
                                          periodic=(True, False, True), bcs_outer="BCWallViscous", seed=4)
        return zones, connect.build(zones, pairs)

    def mixed():
        zones = [cases.tgv_zone((19, 11, 9), num, perturb=0.02, name="a"),
                 H.jittered_box((10, 13, 9), num, bcs=H.WALL_J),
                 cases.channel_zone(17, 10, 5, num)]
        return zones, []

    for make in (tiles, mixed):
        a = run(0, make)
        b = run(1, make)
        for za, zb in zip(a, b):
            for f in ("ro", "ro_m1", "ro_p1", "drodm", "xmut", "coe"):
                assert np.array_equal(getattr(za, f), getattr(zb, f)), (za.name, f)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["tiles", "single"])
def test_gpu_graph_replays_equal_eager(case):
    """Time steps captured in HIP graphs: every replay (one step per call, with the state read
    back and compared in between) equals the eagerly launched run bit for bit; the three level
    rotations and the verification steps each get their own graph."""
    from fast_amd import _lib
    num = H.numerics("ausmpred", "bdf")

    def make():
        if case == "tiles":
            zones, pairs, _ = cases.tiled_box((9, 12), (10, 7), (8,), num, flow="lamin",
                                              periodic=(True, False, True), bcs_outer="BCWallViscous", seed=6)
            return zones, connect.build(zones, pairs)
        return [cases.tgv_zone((20, 14, 11), num, perturb=0.03)], []

    def start(graph):
        zones, racs = make()
        ctx = _lib.Context(0)
        ctx.set_option(_lib.OPT_GRAPH, graph)
        ids = [ctx.add_zone(z) for z in zones]
        ctx.set_transfers(racs)
        ctx.transfer(0)
        for i, z in zip(ids, zones):
            ctx.apply_bc(i, 0)
            t = np.zeros_like(z.ro)
            ctx.get(i, _lib.FIELD_RO, t)
            ctx.set(i, _lib.FIELD_RO_M1, t); ctx.set(i, _lib.FIELD_RO_P1, t)
        return ctx, ids, zones

    ce, ie, ze = start(0)
    cg, ig, zg = start(1)
    nsub = int(ze[0].param_int[P.NSSITER])
    for step in range(9):
        for ctx, ids in ((ce, ie), (cg, ig)):
            ctx.compute(step, 1, nsub, nit_c=1, modulo_verif=3, first_it=0 if step == 0 else 1)
            ctx.sync()
        for i, j in zip(ie, ig):
            for f in (_lib.FIELD_RO, _lib.FIELD_RO_M1):
                ta = np.zeros_like(ze[ie.index(i)].ro); tb = np.zeros_like(ta)
                ce.get(i, f, ta); cg.get(j, f, tb)
                assert np.array_equal(ta, tb), (step, i, f)
    ce.close(); cg.close()
