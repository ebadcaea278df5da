"""Pinning the oracle (CPU, no GPU).

1. Golden fixtures written from the reference FastS Fortran kernels (tests/golden/*.npz,
   produced by tests/golden/gen_golden.py): the oracle must reproduce them bit for bit.
2. When oracle/_ref/libfastref.so is present (the reference kernels compiled from their own
   sources), a wider matrix of cases is compared live, bit for bit.
"""
import copy
import glob
import os

import numpy as np
import pytest

from fast_amd import params as P, cases
from fast_amd.zone import HostZone
from tests import helpers as H
import oracle

# the case fixtures of gen_golden.py (gen_histories.py writes the *_history / *_ke ones)
GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not os.path.basename(p).startswith(("channel257_history", "tgv64_rk3_ke")))


@pytest.fixture(scope="module")
def orc():
    return oracle.Oracle()


def zone_from_fixture(d):
    z = HostZone.from_arrays(d["param_int"], d["param_real"], d["x"], d["y"], d["z"])
    z.ro[:] = d["in_ro"]; z.ro_m1[:] = d["in_ro_m1"]; z.ro_p1[:] = d["in_ro_p1"]
    return z


def interior_equal(z, a, b, nvar=5):
    return np.array_equal(z.interior(a, nvar), z.interior(b, nvar))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_golden_fixture(orc, path):
    d = np.load(path, allow_pickle=False)
    z = zone_from_fixture(d)
    orc.init_metric(z)
    assert np.array_equal(z.metric, d["metric"]), "metric"
    # first sub-iteration RHS
    zr = copy.deepcopy(z)
    rdm = np.zeros(zr.rdm_size())
    explicit = int(zr.param_int[P.ITYPCP]) == 2
    orc.substep(zr, 1, first_it=0, lssiter_verif=1, rdm=rdm, stop=1 if explicit else 2)
    assert interior_equal(zr, zr.drodm, d["rhs_drodm"]), "RHS"
    assert interior_equal(zr, zr.xmut, d["rhs_xmut"], 1), "viscosity"
    assert interior_equal(zr, zr.coe, d["rhs_coe"], int(zr.param_int[P.NEQ_COE])), "coefficients"
    if not explicit:
        assert np.array_equal(rdm, d["rdm"]), "residual blocks"
    # full time steps
    nsub, nsteps = int(d["nsub"]), int(d["nsteps"])
    for step in range(nsteps):
        for it in range(1, nsub + 1):
            orc.substep(z, it, first_it=0 if step == 0 else 1)
        H.rotate(z)
    for f in ("ro", "ro_m1", "ro_p1"):
        assert np.array_equal(getattr(z, f), d["step_" + f]), f
    assert interior_equal(z, z.drodm, d["step_drodm"]), "increment"


requires_ref = pytest.mark.skipif(not oracle.Reference.available(),
                                  reason="reference kernels not built (make -C oracle ref)")


@requires_ref
@pytest.mark.parametrize("scheme", ["roe_min", "roe_kap", "roe_nul", "ausmpred", "senseur"])
@pytest.mark.parametrize("mode", ["explicit", "bdf", "bdf_ss2", "local"])
@pytest.mark.parametrize("mesh", ["cart", "full", "wall", "channel", "inviscid", "outflow", "outpres"])
def test_oracle_vs_reference(orc, scheme, mode, mesh):
    ref = oracle.Reference()
    nsub = {"explicit": 3, "bdf": 1, "bdf_ss2": 2, "local": 1}[mode]
    if mesh == "cart":
        mk = lambda: cases.tgv_zone((9, 8, 7), H.numerics(scheme, mode), perturb=0.02)
    elif mesh == "full":
        mk = lambda: H.jittered_box((8, 7, 6), H.numerics(scheme, mode))
    elif mesh == "wall":
        mk = lambda: H.jittered_box((8, 9, 5), H.numerics(scheme, mode), bcs=H.WALL_J)
    elif mesh == "channel":
        mk = lambda: cases.channel_zone(10, 8, 4, H.numerics(scheme, mode))
    elif mesh == "inviscid":
        mk = lambda: H.jittered_box((7, 8, 5), H.numerics(scheme, mode), bcs=H.INVISCID_J, flow="euler")
    elif mesh == "outflow":
        mk = lambda: H.jittered_box((8, 7, 5), H.numerics(scheme, mode), bcs=H.FARFIELD_OUTFLOW)
    else:
        mk = lambda: H.jittered_box((8, 6, 5), H.numerics(scheme, mode), bcs=H.inflow_outpres((8, 6, 5)))
    zr, zo = mk(), mk()
    p_ref, _ = ref.classify(zr)
    assert np.array_equal(p_ref[:P.N_PARAM_INT], zr.param_int[:P.N_PARAM_INT])
    ref.init_metric(zr); orc.init_metric(zo)
    assert np.array_equal(zr.metric, zo.metric)
    ref.apply_bc(zr, zr.ro); orc.apply_bc(zo, zo.ro)
    assert np.array_equal(zr.ro, zo.ro)
    for z in (zr, zo):
        z.ro_m1[:] = z.ro; z.ro_p1[:] = z.ro
    for step in range(2):
        for it in range(1, nsub + 1):
            ref.substep(zr, it, first_it=0 if step == 0 else 1)
            orc.substep(zo, it, first_it=0 if step == 0 else 1)
        H.rotate(zr); H.rotate(zo)
    for f in ("ro", "ro_m1", "ro_p1"):
        assert np.array_equal(getattr(zr, f), getattr(zo, f)), f
    assert interior_equal(zr, zr.drodm, zo.drodm)
    assert interior_equal(zr, zr.coe, zo.coe, int(zr.param_int[P.NEQ_COE]))


CACHE_BLOCKINGS = {"default": (2048, 2, 7), "j3k2": (2048, 3, 2), "i4j2k3": (4, 2, 3), "j5k1": (2048, 5, 1)}


@requires_ref
@pytest.mark.parametrize("cb", sorted(CACHE_BLOCKINGS))
@pytest.mark.parametrize("scheme", ["roe_min", "ausmpred", "senseur"])
@pytest.mark.parametrize("mode", ["explicit", "bdf_ss2", "local"])
@pytest.mark.parametrize("mesh", ["cart", "wall", "channel", "inviscid"])
def test_cache_blocking_vs_reference(orc, cb, scheme, mode, mesh):
    """The reference's flux pass runs per cache block (navier_stokes_struct.for LOOP_CACHE), and
    a block edge changes the order in which a cell receives its max faces and its Roe wall
    corrections.  With several blocks in i, j and k (the reference default 2048 / 2 / 7 and
    smaller ones) the oracle equals the compiled reference bit for bit, including the CFL sums."""
    ref = oracle.Reference()
    bi, bj, bk = CACHE_BLOCKINGS[cb]
    num = H.numerics(scheme, mode, cache_blocking_I=bi, cache_blocking_J=bj, cache_blocking_K=bk)
    nsub = {"explicit": 3, "bdf_ss2": 2, "local": 1}[mode]
    if mesh == "cart":
        mk = lambda: cases.tgv_zone((9, 10, 16), num, perturb=0.02)
    elif mesh == "wall":
        mk = lambda: H.jittered_box((8, 9, 15), num, bcs=H.WALL_J)
    elif mesh == "channel":
        mk = lambda: cases.channel_zone(10, 8, 9, num)
    else:
        mk = lambda: H.jittered_box((7, 8, 9), num, bcs=H.INVISCID_J, flow="euler")
    zr, zo = mk(), mk()
    ref.init_metric(zr); orc.init_metric(zo)
    for be, z in ((ref, zr), (orc, zo)):
        be.apply_bc(z, z.ro)
        z.ro_m1[:] = z.ro; z.ro_p1[:] = z.ro
    for step in range(2):
        for it in range(1, nsub + 1):
            verif = 1 if it == 1 else 0
            cr = ref.substep(zr, it, first_it=0 if step == 0 else 1, lssiter_verif=verif)
            co = orc.substep(zo, it, first_it=0 if step == 0 else 1, lssiter_verif=verif)
            assert np.array_equal(cr, co), "CFL monitor"
        H.rotate(zr); H.rotate(zo)
    for f in ("ro", "ro_m1", "ro_p1"):
        assert np.array_equal(getattr(zr, f), getattr(zo, f)), f
    # the blocking does change the result: against one block the state differs in the last bits
    if cb == "default" and mesh == "cart" and mode == "bdf_ss2":
        z1 = cases.tgv_zone((9, 10, 16), H.numerics(scheme, mode, **H.SINGLE_BLOCK), perturb=0.02)
        orc.init_metric(z1); orc.apply_bc(z1, z1.ro); z1.ro_m1[:] = z1.ro; z1.ro_p1[:] = z1.ro
        for step in range(2):
            for it in range(1, nsub + 1):
                orc.substep(z1, it, first_it=0 if step == 0 else 1)
            H.rotate(z1)
        assert not np.array_equal(z1.ro, zo.ro)


@requires_ref
@pytest.mark.parametrize("cb", sorted(CACHE_BLOCKINGS))
def test_cache_blocking_enstrophy_vs_reference(orc, cb):
    """post.for runs its gradients per cache block too: per-cell velocity gradients of a
    curvilinear zone bit for bit."""
    ref = oracle.Reference()
    bi, bj, bk = CACHE_BLOCKINGS[cb]
    num = H.numerics("ausmpred", "bdf", cache_blocking_I=bi, cache_blocking_J=bj, cache_blocking_K=bk)
    zr, zo = H.jittered_box((9, 8, 15), num), H.jittered_box((9, 8, 15), num)
    ref.init_metric(zr); orc.init_metric(zo)
    ref.apply_bc(zr, zr.ro); orc.apply_bc(zo, zo.ro)
    gr, go = np.zeros(9 * zr.nd), np.zeros(9 * zo.nd)
    er = ref.enstrophy(zr, zr.ro, gr)
    eo = orc.enstrophy(zo, zo.ro, go)
    assert np.array_equal(zr.interior(gr, 9), zo.interior(go, 9))
    np.testing.assert_allclose(eo, er, rtol=1e-13)


@requires_ref
def test_classification_vs_reference():
    """Zone type and metric sizing (nature_geom_dom) for a Cartesian box, a jittered box and
    the channel-with-bump mesh."""
    ref = oracle.Reference()
    for z in (cases.tgv_zone((7, 6, 5)), H.jittered_box((7, 6, 5), H.numerics("roe_min", "bdf")),
              cases.channel_zone(12, 10, 4)):
        p_ref, degen = ref.classify(z)
        assert np.array_equal(p_ref[:P.N_PARAM_INT], z.param_int[:P.N_PARAM_INT])
        assert not degen.any()


@requires_ref
@pytest.mark.parametrize("mesh", ["cart", "full"])
@pytest.mark.parametrize("flow", ["lamin", "euler"])
@pytest.mark.parametrize("mode", ["explicit", "bdf", "local"])
def test_cfl_monitor_vs_reference(orc, mesh, flow, mode):
    """cpcfl3.for: max / min / sum of the local CFL number at the first sub-iteration of a
    verification step, bit for bit."""
    ref = oracle.Reference()
    num = H.numerics("ausmpred", mode)
    if mesh == "cart":
        mk = lambda: cases.tgv_zone((9, 8, 7), num, perturb=0.05, flow=flow)
    else:
        mk = lambda: H.jittered_box((8, 7, 6), num, flow=flow)
    zr, zo = mk(), mk()
    ref.init_metric(zr); orc.init_metric(zo)
    for be, z in ((ref, zr), (orc, zo)):
        be.apply_bc(z, z.ro)
        z.ro_m1[:] = z.ro; z.ro_p1[:] = z.ro
    cr = ref.substep(zr, 1, first_it=0, lssiter_verif=1)
    co = orc.substep(zo, 1, first_it=0, lssiter_verif=1)
    assert cr[0] > 0 and cr[1] < 1e8
    assert np.array_equal(cr, co)


@requires_ref
@pytest.mark.parametrize("mesh", ["cart", "full"])
def test_enstrophy_vs_reference(orc, mesh):
    """POST/post.for (4th-order Green-Gauss gradient + cptaylor.for sums), bit for bit."""
    ref = oracle.Reference()
    num = H.numerics("ausmpred", "bdf")
    if mesh == "cart":
        mk = lambda: cases.tgv_zone((10, 9, 8), num, perturb=0.05)
    else:
        mk = lambda: H.jittered_box((9, 8, 7), num)
    zr, zo = mk(), mk()
    ref.init_metric(zr); orc.init_metric(zo)
    ref.apply_bc(zr, zr.ro); orc.apply_bc(zo, zo.ro)
    gr, go = np.zeros(9 * zr.nd), np.zeros(9 * zo.nd)
    er = ref.enstrophy(zr, zr.ro, gr)
    eo = orc.enstrophy(zo, zo.ro, go)
    assert er[1] > 0 and er[0] > 0
    assert np.array_equal(zr.interior(gr, 9), zo.interior(go, 9))
    # the per-cell gradients agree bit for bit; the reference's sums run under !$OMP SIMD
    # (POST/loopI_begin.for), whose partial sums the sequential restatement need not follow
    np.testing.assert_allclose(eo, er, rtol=1e-13)


def wall_windows(z):
    """Face windows (face index, cells across) of the j faces and of the imin face."""
    ni, nj, nk = z.dims
    return [(3, [1, ni, 1, 1, 1, nk]), (4, [1, ni, nj + 1, nj + 1, 1, nk]),
            (1, [1, 1, 1, nj, 1, nk]), (6, [1, ni, 1, nj, nk + 1, nk + 1])]


@requires_ref
@pytest.mark.parametrize("scheme", ["ausmpred", "senseur", "roe_min", "roe_kap"])
@pytest.mark.parametrize("mesh", ["cart", "wall", "channel", "inviscid"])
def test_effort_vs_reference(orc, scheme, mesh):
    """_computeStress (compute_effort.cpp -> bceffort.for -> eff_flu*): per-face fields bit for
    bit; the 12 sums (reduced under !$OMP SIMD in the reference) to 1e-13."""
    ref = oracle.Reference()
    num = H.numerics(scheme, "bdf")
    if mesh == "cart":
        mk = lambda: cases.tgv_zone((9, 8, 7), num, perturb=0.05)
    elif mesh == "wall":
        mk = lambda: H.jittered_box((8, 9, 5), num, bcs=H.WALL_J)
    elif mesh == "channel":
        mk = lambda: cases.channel_zone(10, 8, 4, num)
    else:
        mk = lambda: H.jittered_box((7, 8, 5), num, bcs=H.INVISCID_J, flow="euler")
    zr, zo = mk(), mk()
    for be, z in ((ref, zr), (orc, zo)):
        be.init_metric(z)
        be.apply_bc(z, z.ro)
        z.ro_m1[:] = z.ro; z.ro_p1[:] = z.ro
        be.substep(z, 1, first_it=0)        # viscosity and SENSEUR sensor of this state
        be.apply_bc(z, z.ro)
    pos = (0.3, -0.2, 0.1)
    for idir, ind in wall_windows(zr):
        er, fr = ref.effort(zr, idir, ind, pos)
        eo, fo = orc.effort(zo, idir, ind, pos)
        assert np.array_equal(fr, fo), (idir, int(np.argmax(np.any(fr != fo, axis=1))))
        np.testing.assert_allclose(eo, er, rtol=1e-13, atol=1e-13 * np.abs(er).max())
        assert er[6] > 0


def test_effort_closed_form_uniform_flow(orc):
    """A uniform state on a Cartesian box: no viscous stress, the pressure part of the force
    on a j face is (p - p_inf) * area * 2 / (rho_inf V_inf^2) along -+y, the wetted area is the
    face area, and the moment about the face centre is zero."""
    num = H.numerics("ausmpred", "bdf")
    z = cases.tgv_zone((6, 5, 4), num)
    st = z.state
    nd = z.nd
    w = [st["ro"] * 1.1, 0.0, 0.0, 0.0, st["T"]]
    for e in range(5):
        z.ro[e * nd:(e + 1) * nd] = w[e]
    z.ro_m1[:] = z.ro; z.ro_p1[:] = z.ro
    orc.init_metric(z)
    orc.substep(z, 1, first_it=0)
    ni, nj, nk = z.dims
    p = w[0] * w[4] * st["cv"] * (st["gamma"] - 1.)
    norm = 2. / (z.param_real[P.ROINF] * z.param_real[P.VINF] ** 2)
    area = (2 * np.pi / ni) * (2 * np.pi / nk)
    ind = [1, ni, nj + 1, nj + 1, 1, nk]
    eff, flu = orc.effort(z, 4, ind, pos=(np.pi, 2 * np.pi, np.pi))
    np.testing.assert_allclose(eff[1], (p - z.param_real[P.PINF]) * area * ni * nk * norm, rtol=1e-12)
    np.testing.assert_allclose(eff[6], area * ni * nk, rtol=1e-12)
    assert abs(eff[0]) < 1e-12 and abs(eff[2]) < 1e-12
    np.testing.assert_allclose(eff[3:6], 0., atol=1e-11 * abs(eff[1]))
