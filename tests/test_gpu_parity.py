"""HIP kernels against the oracle, bit for bit (FP64, contraction off on both sides).

Each case builds the same zone twice, runs one or more sub-iterations through the C ABI on
the GPU and through the oracle's C restatement on the CPU, and compares every interior field
(and, after a full sub-iteration, the ghost cells) with exact equality.
"""
import numpy as np
import pytest

from fast_amd import params as P, cases
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctxmod():
    from fast_amd import _lib
    return _lib


@pytest.fixture(scope="module")
def orc():
    import oracle
    return oracle.Oracle()


def run_pair(ctxmod, orc, make, nsub, nsteps=1, first_it=0, stop=0, verif=True, schedule=None,
             options=()):
    zo = make()
    zg = make()
    orc.init_metric(zo)
    orc.apply_bc(zo, zo.ro)
    zo.ro_m1[:] = zo.ro; zo.ro_p1[:] = zo.ro
    ctx = ctxmod.Context(0)
    if schedule is not None:
        ctx.set_option(ctxmod.OPT_LU_SCHEDULE, schedule)
    for o, v in options:
        ctx.set_option(o, v)
    zid = ctx.add_zone(zg)
    ctx.apply_bc(zid, 0)
    # same N-1 / N+1 initialisation as the oracle side
    tmp = np.zeros_like(zg.ro)
    ctx.get(zid, ctxmod.FIELD_RO, tmp)
    ctx.set(zid, ctxmod.FIELD_RO_M1, tmp)
    ctx.set(zid, ctxmod.FIELD_RO_P1, tmp)
    rdm = np.zeros(zo.rdm_size())
    for step in range(nsteps):
        fi = first_it if step == 0 else 1
        for it in range(1, nsub + 1):
            orc.substep(zo, it, first_it=fi, lssiter_verif=1 if verif else 0, rdm=rdm, stop=stop)
            ctx.substep(zid, it, first_it=fi, lssiter_verif=1 if verif else 0, stop=stop)
        if nsteps > 1:
            H.rotate(zo)
            # rotate the device levels the same way (N-1 <- N <- N+1)
            a = [np.zeros_like(zo.ro) for _ in range(3)]
            for f, buf in zip((ctxmod.FIELD_RO, ctxmod.FIELD_RO_M1, ctxmod.FIELD_RO_P1), a):
                ctx.get(zid, f, buf)
            ctx.set(zid, ctxmod.FIELD_RO_M1, a[0])
            ctx.set(zid, ctxmod.FIELD_RO, a[2])
            ctx.set(zid, ctxmod.FIELD_RO_P1, a[1])
    ctx.pull(zid, zg)
    norms = ctx.residual_norms(zid) if (verif and zo.param_int[P.ITYPCP] <= 1 and stop != 1) else None
    ctx.close()
    return zo, zg, rdm, norms


def assert_same(zo, zg, fields=("ro", "ro_m1", "ro_p1", "drodm"), ghosts=False):
    assert np.array_equal(zo.metric, zg.metric)
    for f in fields:
        a, b = getattr(zo, f), getattr(zg, f)
        if ghosts and f != "drodm":
            # the reference also writes -F into the drodm of the first ghost layer; those
            # values are never used, so drodm is compared on the interior only
            ok = np.array_equal(a, b)
        else:
            ok = np.array_equal(zo.interior(a), zg.interior(b))
        if not ok:
            d = np.abs(zo.interior(a) - zg.interior(b))
            raise AssertionError("%s differs: max %.3e at %s" % (f, d.max(), np.unravel_index(d.argmax(), d.shape)))


SCHEMES = ["roe_min", "roe_kap", "roe_nul", "ausmpred", "senseur"]


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("flow", ["lamin", "euler"])
@pytest.mark.parametrize("stop", [1, 3, 0])
def test_tgv_cart_implicit_one_substep(ctxmod, orc, scheme, flow, stop):
    mk = lambda: cases.tgv_zone((13, 11, 9), H.numerics(scheme, "bdf"), perturb=0.02, flow=flow)
    zo, zg, _, _ = run_pair(ctxmod, orc, mk, 1, stop=stop)
    assert_same(zo, zg, ghosts=(stop == 0))
    assert np.array_equal(zo.interior(zo.xmut, 1), zg.interior(zg.xmut, 1))
    assert np.array_equal(zo.interior(zo.coe), zg.interior(zg.coe))


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("mode", ["explicit", "bdf", "bdf_ss2", "local"])
def test_tgv_cart_modes_two_steps(ctxmod, orc, scheme, mode):
    nsub = {"explicit": 3, "bdf": 1, "bdf_ss2": 2, "local": 1}[mode]
    mk = lambda: cases.tgv_zone((12, 10, 8), H.numerics(scheme, mode), perturb=0.02)
    zo, zg, _, _ = run_pair(ctxmod, orc, mk, nsub, nsteps=2)
    assert_same(zo, zg, ghosts=True)


@pytest.mark.parametrize("schedule", ["blocks", "pipeline", "dataflow", "dataflow_g7", "stream"])
@pytest.mark.parametrize("mesh", ["cart", "full", "wall", "cartwall", "cartbig"])
@pytest.mark.parametrize("mode", ["bdf", "bdf_ss2"])
def test_lu_schedules(ctxmod, orc, schedule, mesh, mode):
    """The LU-SGS schedules (one launch per block diagonal; persistent trailing tiles; blocks
    handed out in diagonal order with a face exchange) on zones of several blocks in i, j and k,
    in-place sub-iterations included."""
    sch = {"blocks": ctxmod.LU_BLOCKS, "pipeline": ctxmod.LU_PIPELINE, "dataflow": ctxmod.LU_DATAFLOW,
           "dataflow_g7": ctxmod.LU_DATAFLOW, "stream": ctxmod.LU_STREAM}[schedule]
    # dataflow_g7: the dataflow sweeps on 7 workgroups (FM_OPT_LU_FLOW_GRID)
    opts = ((ctxmod.OPT_LU_FLOW_GRID, 7),) if schedule == "dataflow_g7" else ()
    nsub = 2 if mode == "bdf_ss2" else 1
    if mesh == "cart":
        mk = lambda: cases.tgv_zone((21, 19, 26), H.numerics("roe_kap", mode), perturb=0.02)
    elif mesh == "cartwall":
        # Cartesian with walls: the LU box extends one ghost layer at the walls
        mk = lambda: cases.tgv_zone((19, 23, 17), H.numerics("ausmpred", mode), perturb=0.02, bcs=H.WALL_J)
    elif mesh == "cartbig":
        # many pencils, ragged in every direction
        mk = lambda: cases.tgv_zone((45, 37, 29), H.numerics("ausmpred", mode), perturb=0.02)
    elif mesh == "full":
        mk = lambda: H.jittered_box((11, 18, 17), H.numerics("ausmpred", mode))
    else:
        mk = lambda: H.jittered_box((10, 17, 20), H.numerics("senseur", mode), bcs=H.WALL_J)
    zo, zg, _, _ = run_pair(ctxmod, orc, mk, nsub, nsteps=2, schedule=sch, options=opts)
    assert_same(zo, zg, ghosts=True)


@pytest.mark.parametrize("scheme", ["roe_min", "ausmpred", "senseur"])
@pytest.mark.parametrize("mode", ["explicit", "bdf", "bdf_ss2", "local"])
def test_3dfull_jittered(ctxmod, orc, scheme, mode):
    nsub = {"explicit": 3, "bdf": 1, "bdf_ss2": 2, "local": 1}[mode]
    mk = lambda: H.jittered_box((11, 9, 8), H.numerics(scheme, mode))
    zo, zg, _, _ = run_pair(ctxmod, orc, mk, nsub, nsteps=2)
    assert zo.zone_type == P.ZONE_3DFULL
    assert_same(zo, zg, ghosts=True)


@pytest.mark.parametrize("bcs", ["wall_j", "extrap_i"])
@pytest.mark.parametrize("mode", ["bdf", "local", "bdf_ss2"])
def test_walls_lu_extension(ctxmod, orc, bcs, mode):
    """Viscous walls extend the LU range by one ghost layer with the RHS mirrored there
    (BCzone.cpp:115-253); extrapolated faces zero the RHS ghosts."""
    b = H.WALL_J if bcs == "wall_j" else H.EXTRAP_I
    nsub = 2 if mode == "bdf_ss2" else 1
    mk = lambda: H.jittered_box((10, 12, 7), H.numerics("ausmpred", mode), bcs=b)
    zo, zg, _, _ = run_pair(ctxmod, orc, mk, nsub, nsteps=2)
    assert_same(zo, zg, ghosts=True)


BC_CASES = {
    # channel with a bump: farfield in/out, viscous walls, periodic k (3dfull)
    "channel": lambda sch, mode: cases.channel_zone(18, 12, 6, H.numerics(sch, mode)),
    # Euler zone with inviscid (slip) walls
    "inviscid": lambda sch, mode: H.jittered_box((9, 10, 6), H.numerics(sch, mode), bcs=H.INVISCID_J, flow="euler"),
    # farfield inlet, outflow exit
    "outflow": lambda sch, mode: H.jittered_box((11, 7, 5), H.numerics(sch, mode), bcs=H.FARFIELD_OUTFLOW),
    # free-stream inflow, exit pressure per face point
    "outpres": lambda sch, mode: H.jittered_box((9, 6, 5), H.numerics(sch, mode), bcs=H.inflow_outpres((9, 6, 5))),
}


@pytest.mark.parametrize("case", sorted(BC_CASES))
@pytest.mark.parametrize("mode", ["explicit", "bdf", "local", "bdf_ss2"])
@pytest.mark.parametrize("scheme", ["ausmpred", "senseur"])
def test_boundary_conditions(ctxmod, orc, case, mode, scheme):
    """Farfield / inflow (characteristic, BC_nextrank), outflow and outpres (BC_extrap), and
    inviscid walls against the oracle, ghost cells included, over two time steps."""
    nsub = {"explicit": 3, "bdf": 1, "bdf_ss2": 2, "local": 1}[mode]
    mk = lambda: BC_CASES[case](scheme, mode)
    zo, zg, _, _ = run_pair(ctxmod, orc, mk, nsub, nsteps=2)
    assert_same(zo, zg, ghosts=True)


@pytest.mark.parametrize("case", ["channel", "inviscid", "wall_j"])
@pytest.mark.parametrize("mode", ["explicit", "bdf", "local"])
@pytest.mark.parametrize("scheme", ["roe_min", "roe_kap", "roe_nul"])
def test_roe_wall_flux_correction(ctxmod, orc, case, mode, scheme):
    """Roe with walls: the Euler flux of every wall face is replaced by the wall flux
    (correct_flux.for: corr_fluroe_* then bflwall / bflwallslip), ghost cells included."""
    nsub = {"explicit": 3, "bdf": 1, "local": 1}[mode]
    if case == "wall_j":
        mk = lambda: H.jittered_box((10, 9, 6), H.numerics(scheme, mode), bcs=H.WALL_J)
    else:
        mk = lambda: BC_CASES[case](scheme, mode)
    zo, zg, _, _ = run_pair(ctxmod, orc, mk, nsub, nsteps=2)
    assert_same(zo, zg, ghosts=True)


@pytest.mark.parametrize("schedule", ["faces", "kcol"])
def test_roe_wall_correction_other_schedules(ctxmod, orc, schedule):
    """The wall correction under the other residual schedules (faces falls back to cells)."""
    num = H.numerics("roe_kap", "bdf")
    mk = lambda: cases.channel_zone(18, 12, 6, num)
    zo, zg = mk(), mk()
    orc.init_metric(zo); orc.apply_bc(zo, zo.ro)
    zo.ro_m1[:] = zo.ro; zo.ro_p1[:] = zo.ro
    orc.substep(zo, 1, first_it=0, stop=1)
    ctx = ctxmod.Context(0)
    set_res(ctxmod, ctx, schedule)
    zid = ctx.add_zone(zg)
    ctx.apply_bc(zid, 0)
    t = np.zeros_like(zg.ro); ctx.get(zid, ctxmod.FIELD_RO, t)
    ctx.set(zid, ctxmod.FIELD_RO_M1, t); ctx.set(zid, ctxmod.FIELD_RO_P1, t)
    ctx.substep(zid, 1, first_it=0, stop=1)
    ctx.pull(zid, zg)
    ctx.close()
    assert np.array_equal(zo.interior(zo.drodm), zg.interior(zg.drodm))


def test_norms_match_residual_blocks(ctxmod, orc):
    """GPU L2/Linf against the oracle's per-block sums (cprdu3s1.for), relative 1e-12."""
    mk = lambda: cases.tgv_zone((24, 20, 16), H.numerics("roe_min", "bdf"), perturb=0.02)
    zo, zg, rdm, norms = run_pair(ctxmod, orc, mk, 1, stop=2)
    nss = int(zo.param_int[P.NSSITER])
    r = rdm.reshape(-1, 2, 5, nss)      # (block, sum/max, eq, it) in Fortran order reversed
    s2 = r[:, 0, :, 0].sum(axis=0)
    mx = r[:, 1, :, 0].max(axis=0)
    n = 24 * 20 * 16
    dtc = zo.param_real[P.DTC]
    l2 = np.sqrt(s2 / n) / dtc
    linf = mx / dtc
    np.testing.assert_allclose(norms[:5], l2, rtol=1e-12, atol=0)
    assert np.array_equal(norms[5:], linf)


@pytest.mark.parametrize("n", [(2, 2, 2), (2, 3, 2), (67, 5, 3)])
def test_ragged_sizes(ctxmod, orc, n):
    """Tiny and odd sizes (two cells per direction, sizes not multiple of the 64x4 blocks)."""
    mk = lambda: cases.tgv_zone(n, H.numerics("roe_min", "bdf"), perturb=0.02)
    zo, zg, _, _ = run_pair(ctxmod, orc, mk, 1, nsteps=2)
    assert_same(zo, zg, ghosts=True)


def test_compute_loop_matches_oracle(ctxmod, orc):
    """fm_compute (time-step loop with level rotation) against the oracle driven the same way."""
    num = H.numerics("roe_min", "bdf")
    zo = cases.tgv_zone((16, 12, 10), num, perturb=0.02)
    zg = cases.tgv_zone((16, 12, 10), num, perturb=0.02)
    orc.init_metric(zo); orc.apply_bc(zo, zo.ro); zo.ro_m1[:] = zo.ro; zo.ro_p1[:] = zo.ro
    ctx = ctxmod.Context(0)
    zid = ctx.add_zone(zg)
    ctx.apply_bc(zid, 0)
    t = np.zeros_like(zg.ro); ctx.get(zid, ctxmod.FIELD_RO, t)
    ctx.set(zid, ctxmod.FIELD_RO_M1, t); ctx.set(zid, ctxmod.FIELD_RO_P1, t)
    nsteps = 4
    ctx.compute(0, 1, 2, nit_c=nsteps, modulo_verif=2, first_it=0)
    ctx.sync()
    for step in range(nsteps):
        orc.substep(zo, 1, first_it=0 if step == 0 else 1)
        H.rotate(zo)
    ctx.pull(zid, zg)
    ctx.close()
    assert np.array_equal(zo.interior(zo.ro), zg.interior(zg.ro))
    assert np.array_equal(zo.ro, zg.ro)


@pytest.mark.parametrize("graph", [0, 1])
def test_compute_leaves_dw_like_eager_sweeps(ctxmod, graph):
    """Inside fm_compute the streaming LU-SGS leaves dW in its skewed rows and copies it to drodm
    once, after the last time step, and reuses the state rows its upper sweep wrote (graph replays
    included; an edit of the state in between invalidates them): every field, drodm too, equals the
    dataflow schedule's, which writes dW at every sweep and copies nothing."""
    num = H.numerics("ausmpred", "bdf")
    out = {}
    for sched in (ctxmod.LU_STREAM, ctxmod.LU_DATAFLOW):
        zg = cases.tgv_zone((20, 17, 11), num, perturb=0.02)
        ctx = ctxmod.Context(0)
        ctx.set_option(ctxmod.OPT_LU_SCHEDULE, sched)
        ctx.set_option(ctxmod.OPT_GRAPH, graph)
        zid = ctx.add_zone(zg)
        ctx.apply_bc(zid, 0)
        t = np.zeros_like(zg.ro); ctx.get(zid, ctxmod.FIELD_RO, t)
        ctx.set(zid, ctxmod.FIELD_RO_M1, t); ctx.set(zid, ctxmod.FIELD_RO_P1, t)
        ctx.compute(0, 1, 1, nit_c=5, modulo_verif=0, first_it=0)
        # an edit of the state between two calls: the rows the streaming sweeps kept of it are
        # stale and must be copied again
        w = np.zeros_like(zg.ro); ctx.get(zid, ctxmod.FIELD_RO, w)
        w *= 1.0 + 1e-3 * np.sin(np.arange(w.size))
        ctx.set(zid, ctxmod.FIELD_RO, w)
        ctx.compute(5, 1, 1, nit_c=3, modulo_verif=2, first_it=1)
        ctx.compute(8, 1, 1, nit_c=4, modulo_verif=0, first_it=1)
        ctx.sync()
        ctx.pull(zid, zg)
        ctx.close()
        out[sched] = zg
    a, b = out[ctxmod.LU_STREAM], out[ctxmod.LU_DATAFLOW]
    for f in ("ro", "ro_m1", "ro_p1", "drodm"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


@pytest.mark.parametrize("mesh", ["cart", "full"])
@pytest.mark.parametrize("flow", ["lamin", "euler"])
def test_cfl_monitor(ctxmod, orc, mesh, flow):
    """cpcfl3.for on the GPU: max and min exact, the mean to 1e-12 (the sum is reduced in
    another order)."""
    num = H.numerics("ausmpred", "bdf")
    if mesh == "cart":
        mk = lambda: cases.tgv_zone((23, 17, 9), num, perturb=0.05, flow=flow)
    else:
        mk = lambda: H.jittered_box((12, 11, 9), num, flow=flow)
    zo, zg = mk(), mk()
    orc.init_metric(zo); orc.apply_bc(zo, zo.ro)
    zo.ro_m1[:] = zo.ro; zo.ro_p1[:] = zo.ro
    cfl = orc.substep(zo, 1, first_it=0, lssiter_verif=1, stop=1)
    ctx = ctxmod.Context(0)
    zid = ctx.add_zone(zg)
    ctx.apply_bc(zid, 0)
    ctx.substep(zid, 1, first_it=0, lssiter_verif=1, stop=1)
    got = ctx.cfl(zid)
    ctx.close()
    n = np.prod(zo.dims)
    assert got[0] == cfl[0] and got[1] == cfl[1]
    np.testing.assert_allclose(got[2], cfl[2] / n, rtol=1e-12)


# schedule -> (FM_OPT_RESIDUAL_SCHEDULE, FM_OPT_RES_KCOL_N)
RES = {"cells": (0, 0), "faces": (1, 0), "kcol": (2, 0), "kcol3": (2, 3), "auto": (3, 0)}


def set_res(ctxmod, ctx, schedule):
    ctx.set_option(ctxmod.OPT_RESIDUAL_SCHEDULE, RES[schedule][0])
    ctx.set_option(ctxmod.OPT_RES_KCOL_N, RES[schedule][1])


@pytest.mark.parametrize("schedule", ["cells", "faces", "kcol", "kcol3", "auto"])
@pytest.mark.parametrize("mesh", ["cart", "full"])
@pytest.mark.parametrize("scheme", ["roe_min", "ausmpred", "senseur"])
def test_residual_schedules(ctxmod, orc, schedule, mesh, scheme):
    """Face balance with six faces per cell, with every face once (63-cell i-chunks, 4-row
    j-chunks, 32-plane k-chunks), with 4-cell and 3-cell k columns and with the per-mesh default:
    the sizes straddle every chunk boundary (35 and 34 planes are not multiples of 3, 4 or 32)."""
    num = H.numerics(scheme, "bdf")
    if mesh == "cart":
        mk = lambda: cases.tgv_zone((70, 9, 35), num, perturb=0.02)
    else:
        mk = lambda: H.jittered_box((66, 7, 34), num)
    zo, zg = mk(), mk()
    orc.init_metric(zo); orc.apply_bc(zo, zo.ro)
    zo.ro_m1[:] = zo.ro; zo.ro_p1[:] = zo.ro
    orc.substep(zo, 1, first_it=0, stop=1)
    ctx = ctxmod.Context(0)
    set_res(ctxmod, ctx, schedule)
    zid = ctx.add_zone(zg)
    ctx.apply_bc(zid, 0)
    t = np.zeros_like(zg.ro); ctx.get(zid, ctxmod.FIELD_RO, t)
    ctx.set(zid, ctxmod.FIELD_RO_M1, t); ctx.set(zid, ctxmod.FIELD_RO_P1, t)
    ctx.substep(zid, 1, first_it=0, stop=1)
    ctx.pull(zid, zg)
    ctx.close()
    assert np.array_equal(zo.interior(zo.drodm), zg.interior(zg.drodm))


@pytest.mark.parametrize("mesh", ["cart", "full"])
def test_enstrophy_tke(ctxmod, orc, mesh):
    """POST/post.for on the GPU: kinetic energy, enstrophy and volume sums against the oracle
    (sums reduced in another order: 1e-12)."""
    num = H.numerics("ausmpred", "bdf")
    if mesh == "cart":
        mk = lambda: cases.tgv_zone((31, 17, 12), num, perturb=0.05)
    else:
        mk = lambda: H.jittered_box((19, 13, 11), num)
    zo, zg = mk(), mk()
    orc.init_metric(zo); orc.apply_bc(zo, zo.ro)
    e = orc.enstrophy(zo, zo.ro)
    ctx = ctxmod.Context(0)
    zid = ctx.add_zone(zg)
    ctx.apply_bc(zid, 0)
    got = ctx.tke(zid, 0)
    ctx.close()
    np.testing.assert_allclose([got[0], got[2], got[1]], e, rtol=1e-12)


EFFORT_CASES = {
    "cart": lambda num: cases.tgv_zone((13, 9, 7), num, perturb=0.05),
    "wall": lambda num: H.jittered_box((10, 9, 6), num, bcs=H.WALL_J),
    "channel": lambda num: cases.channel_zone(16, 9, 5, num),
    "inviscid": lambda num: H.jittered_box((9, 8, 5), num, bcs=H.INVISCID_J, flow="euler"),
}


@pytest.mark.parametrize("case", sorted(EFFORT_CASES))
@pytest.mark.parametrize("scheme", ["ausmpred", "senseur", "roe_min"])
def test_wall_effort(ctxmod, orc, case, scheme):
    """fm_wall_effort (compute_effort.cpp / bceffort.for) against the oracle: per-face fields
    and the 12 sums bit for bit (both sum in the reference's loop order)."""
    num = H.numerics(scheme, "bdf")
    zo, zg = EFFORT_CASES[case](num), EFFORT_CASES[case](num)
    orc.init_metric(zo); orc.apply_bc(zo, zo.ro)
    zo.ro_m1[:] = zo.ro; zo.ro_p1[:] = zo.ro
    orc.substep(zo, 1, first_it=0)
    orc.apply_bc(zo, zo.ro)
    ctx = ctxmod.Context(0)
    zid = ctx.add_zone(zg)
    ctx.apply_bc(zid, 0)
    t = np.zeros_like(zg.ro); ctx.get(zid, ctxmod.FIELD_RO, t)
    ctx.set(zid, ctxmod.FIELD_RO_M1, t); ctx.set(zid, ctxmod.FIELD_RO_P1, t)
    ctx.substep(zid, 1, first_it=0)
    ctx.apply_bc(zid, 0)
    ni, nj, nk = zo.dims
    pos = (0.3, -0.2, 0.1)
    for idir, ind in [(3, [1, ni, 1, 1, 1, nk]), (4, [1, ni, nj + 1, nj + 1, 1, nk]),
                      (1, [1, 1, 1, nj, 1, nk]), (6, [2, ni - 1, 1, nj, nk + 1, nk + 1])]:
        eo, fo = orc.effort(zo, idir, ind, pos)
        eg, fg = ctx.wall_effort(zid, idir, ind, pos, level=0)
        assert np.array_equal(fo, fg), (case, idir)
        assert np.array_equal(eo, eg), (case, idir)
    ctx.close()


def test_c0_tgv64_roe_minmod_rk3(ctxmod, orc):
    """The CPU-runnable case at its own size: TGV 64^3 periodic, Roe with the minmod slope,
    laminar, explicit RK3 with dt = 0.5 h / max(|u| + c) of the initial field; two time steps
    (six stages) on the GPU equal the oracle bit for bit, ghost cells included."""
    num = H.numerics("roe_min", "explicit")

    def make():
        z = cases.tgv_zone(64, num)
        z.param_real[P.DTC] = cases.tgv_dt(z, cfl=0.5)
        return z
    zo, zg, _, _ = run_pair(ctxmod, orc, make, 3, nsteps=2)
    assert_same(zo, zg, fields=("ro", "ro_m1", "ro_p1"), ghosts=True)


def test_tgv256_one_iteration_matches_oracle(ctxmod, orc):
    """The benchmark's own size and case (TGV 256^3, laminar AUSM+(P) o3, implicit BDF, global
    dt at CFL 2, one Newton sub-iteration): one full iteration on the GPU equals the oracle's bit
    for bit (about 2 s of oracle time, 9 GB of host memory for the two copies)."""
    num = P.Numerics(scheme="ausmpred", slope="o3", temporal_scheme="implicit", ss_iteration=1,
                     time_step_nature="global", time_step=1.0)

    def make():
        z = cases.tgv_zone(256, num)
        z.param_real[P.DTC] = cases.tgv_dt(z, cfl=2.0)
        return z
    zo, zg, _, _ = run_pair(ctxmod, orc, make, 1)
    assert_same(zo, zg)


CB = {"default": (2048, 2, 7), "j3k2": (2048, 3, 2), "j5k1": (2048, 5, 1), "single": (1000000,) * 3}


@pytest.mark.parametrize("cb", sorted(CB))
@pytest.mark.parametrize("schedule", ["cells", "faces", "kcol"])
@pytest.mark.parametrize("case", ["cart", "channel", "inviscid"])
@pytest.mark.parametrize("scheme", ["roe_min", "ausmpred", "senseur"])
def test_cache_blocking(ctxmod, orc, cb, schedule, case, scheme):
    """The reference's flux pass runs per cache block, which fixes the order in which a cell
    receives its max faces and its Roe wall corrections.  Every residual schedule follows it for
    any j and k block sizes: two implicit time steps equal the oracle bit for bit, ghost cells
    included."""
    bi, bj, bk = CB[cb]
    num = H.numerics(scheme, "bdf", cache_blocking_I=bi, cache_blocking_J=bj, cache_blocking_K=bk)
    if case == "cart":
        mk = lambda: cases.tgv_zone((19, 10, 16), num, perturb=0.02)
    elif case == "channel":
        mk = lambda: cases.channel_zone(18, 12, 15, num)
    else:
        mk = lambda: H.jittered_box((9, 10, 9), num, bcs=H.INVISCID_J, flow="euler")
    zo, zg = mk(), mk()
    orc.init_metric(zo); orc.apply_bc(zo, zo.ro)
    zo.ro_m1[:] = zo.ro; zo.ro_p1[:] = zo.ro
    ctx = ctxmod.Context(0)
    set_res(ctxmod, ctx, schedule)
    zid = ctx.add_zone(zg)
    ctx.apply_bc(zid, 0)
    t = np.zeros_like(zg.ro); ctx.get(zid, ctxmod.FIELD_RO, t)
    ctx.set(zid, ctxmod.FIELD_RO_M1, t); ctx.set(zid, ctxmod.FIELD_RO_P1, t)
    for step in range(2):
        fi = 0 if step == 0 else 1
        orc.substep(zo, 1, first_it=fi)
        H.rotate(zo)
        ctx.compute(step, 1, 1, nit_c=1, modulo_verif=0, first_it=fi)
    ctx.sync()
    ctx.pull(zid, zg)
    ctx.close()
    assert_same(zo, zg, fields=("ro", "ro_m1", "ro_p1"), ghosts=True)


def test_cache_blocking_limits(ctxmod):
    """More than one cache block in i is refused (FM_E_UNSUPPORTED), as is a block size below 1."""
    z = cases.tgv_zone((20, 6, 5), H.numerics("ausmpred", "bdf", cache_blocking_I=8))
    ctx = ctxmod.Context(0)
    with pytest.raises(ctxmod.FastMIError, match="cache block in i"):
        ctx.add_zone(z)
    z = cases.tgv_zone((20, 6, 5), H.numerics("ausmpred", "bdf", cache_blocking_J=0))
    with pytest.raises(ctxmod.FastMIError, match="positive"):
        ctx.add_zone(z)
    ctx.close()


def test_lu_wait_give_up_is_reported(ctxmod):
    """The failure path of the bounded LU-SGS waits, forced from the host (FM_OPT_LU_GIVE_UP sets
    the error word a wait that gives up would set): the dataflow sweeps still finish, the next
    call that waits for the stream reports FM_E_HIP instead of handing out the state, the word is
    cleared, and the following steps run normally."""
    num = H.numerics("ausmpred", "bdf")
    z = cases.tgv_zone((24, 20, 18), num, perturb=0.02)
    ctx = ctxmod.Context(0)
    ctx.set_option(ctxmod.OPT_LU_SCHEDULE, ctxmod.LU_DATAFLOW)
    zid = ctx.add_zone(z)
    ctx.apply_bc(zid, 0)
    ctx.compute(0, 1, 1, nit_c=1, modulo_verif=0, first_it=0)
    ctx.sync()
    ctx.set_option(ctxmod.OPT_LU_GIVE_UP, 1)
    ctx.compute(1, 1, 1, nit_c=2, modulo_verif=0, first_it=1)
    buf = np.zeros_like(z.ro)
    with pytest.raises(ctxmod.FastMIError, match="timed out"):
        ctx.get(zid, ctxmod.FIELD_RO, buf)
    ctx.get(zid, ctxmod.FIELD_RO, buf)          # reported once, then cleared
    ctx.compute(3, 1, 1, nit_c=1, modulo_verif=0, first_it=1)
    ctx.sync()
    ctx.close()


def test_fp32_and_pipeline_exclude_each_other(ctxmod):
    ctx = ctxmod.Context(0)
    ctx.set_option(ctxmod.OPT_LU_SCHEDULE, ctxmod.LU_PIPELINE)
    with pytest.raises(ctxmod.FastMIError, match="FP32"):
        ctx.set_option(ctxmod.OPT_LU_PRECISION, ctxmod.LU_FP32)
    ctx.set_option(ctxmod.OPT_LU_SCHEDULE, ctxmod.LU_DATAFLOW)
    ctx.set_option(ctxmod.OPT_LU_PRECISION, ctxmod.LU_FP32)
    with pytest.raises(ctxmod.FastMIError, match="FP32"):
        ctx.set_option(ctxmod.OPT_LU_SCHEDULE, ctxmod.LU_PIPELINE)
    ctx.close()
