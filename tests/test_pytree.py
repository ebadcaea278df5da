"""The FastS-style tree API (fast_amd.PyTree): tree <-> zone descriptors on the CPU, and
warmup + _compute on the GPU against the oracle driven the way FastS.PyTree._compute drives
the kernels (Fast/FastS/FastS/PyTree.py:42-236)."""
import numpy as np
import pytest

from fast_amd import Internal, params as P, cases, PyTree as FP
from tests import helpers as H


def tree_of(zones, num, model="NSLaminar"):
    t = Internal.newCGNSTree()
    b = Internal.newCGNSBase("Base", parent=t)
    for hz in zones:
        FP.newZoneFromHost(hz, b)
    FP._setNum2Zones(t, dict(num))
    FP._addState(t, zones[0].state, model=model)
    return t


CASES = {
    "tgv_periodic": lambda num: cases.tgv_zone((8, 6, 5), num, perturb=0.02),
    "channel_farfield_walls": lambda num: cases.channel_zone(12, 8, 4, num),
    "jbox_extrap_walls": lambda num: H.jittered_box((8, 7, 6), num, bcs=H.EXTRAP_I),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_tree_rebuilds_descriptors(case):
    """A zone written to a tree and read back gives the same Parameter_int / Parameter_real
    (BC windows go through PointRange and getRange, getRange.cpp:57-110)."""
    num = H.numerics("ausmpred", "local")
    hz = CASES[case](num)
    t = tree_of([hz], num)
    b = Internal.getBases(t)[0]
    z = Internal.getZones(t)[0]
    hz2 = FP._host_zone(b, z)
    assert np.array_equal(hz.param_int, hz2.param_int)
    assert np.array_equal(hz.param_real, hz2.param_real)
    assert np.array_equal(hz.x, hz2.x) and np.array_equal(hz.z, hz2.z)
    fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
    rho = Internal.getNodeFromName1(fs, "Density")[1]
    assert rho.shape == tuple(d + 4 for d in hz.dims)
    assert np.array_equal(rho.ravel(order="F"), hz.ro[:hz.nd])


def test_euler_model_and_unsupported_model():
    num = H.numerics("roe_min", "bdf")
    hz = cases.tgv_zone((6, 6, 6), num, flow="euler")
    t = tree_of([hz], num, model="Euler")
    hz2 = FP._host_zone(Internal.getBases(t)[0], Internal.getZones(t)[0])
    assert np.array_equal(hz.param_int, hz2.param_int)
    with pytest.raises(NotImplementedError):
        FP._addState(t, hz.state, model="NSTurbulent")


def test_numerics_roundtrip_through_define_nodes():
    num = P.Numerics(scheme="roe_kap", temporal_scheme="implicit", ss_iteration=3, cfl=4.0,
                     time_step_nature="local")
    t = Internal.newCGNSTree()
    b = Internal.newCGNSBase("Base", parent=t)
    z = Internal.newZone("z", 9, 9, 9, parent=b)
    FP._setNum2Base(t, {"temporal_scheme": "implicit", "ss_iteration": 3})
    FP._setNum2Zones(t, {"scheme": "roe_kap", "cfl": 4.0, "time_step_nature": "local"})
    got = FP._numerics(b, z)
    for k in ("scheme", "temporal_scheme", "ss_iteration", "cfl", "time_step_nature"):
        assert got[k] == num[k]
    assert got["slope"] == "o3"


@pytest.mark.gpu
def test_warmup_compute_matches_oracle():
    """warmup + _compute(NIT=3) on a two-zone tree equals the oracle stepping each zone."""
    import oracle
    orc = oracle.Oracle()
    num = H.numerics("ausmpred", "bdf")
    mk = [lambda: cases.tgv_zone((12, 10, 8), num, perturb=0.02, name="a"),
          lambda: H.jittered_box((8, 9, 7), num, bcs=H.WALL_J)]
    t = tree_of([m() for m in mk], num)
    Internal.getZones(t)[1][0] = "b"
    t, tc, metrics = FP.warmup(t, None)
    assert FP.HOOK["first_it"] == 0
    rho_node = Internal.getNodeFromName1(Internal.getNodeFromName1(Internal.getZones(t)[0],
                                         "FlowSolution#Centers"), "Density")
    rho_obj = rho_node[1]
    FP._createConvergenceHistory(t, 4)
    nit = 3
    FP._compute(t, metrics, 0, NIT=nit)
    FP.display_temporal_criteria(t, metrics, nit, format="store")
    assert rho_node[1] is rho_obj          # fields refreshed in place
    for m, z in zip(mk, Internal.getZones(t)):
        zo = m()
        orc.init_metric(zo)
        orc.apply_bc(zo, zo.ro)
        zo.ro_m1[:] = zo.ro; zo.ro_p1[:] = zo.ro
        for step in range(nit):
            orc.substep(zo, 1, first_it=0 if step == 0 else 1)
            H.rotate(zo)
        fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
        for sfx, a in (("", zo.ro), ("_M1", zo.ro_m1)):
            got = np.concatenate([Internal.getNodeFromName1(fs, v + sfx)[1].ravel(order="F")
                                  for v in FP.VARS])
            assert np.array_equal(zo.interior(got), zo.interior(a)), (z[0], sfx)
        h = Internal.getNodeFromName1(z, "ZoneConvergenceHistory")
        assert int(Internal.getValue(h)) == 1          # the record count is the node's value
        assert int(Internal.getNodeFromName1(h, "IterationNumber")[1][0]) == nit
    for z in Internal.getZones(t):
        cfl = Internal.getValue(Internal.getNodeFromName(z, "CFL_minmaxmoy"))
        assert cfl[0] >= cfl[2] >= cfl[1] > 0
    enst, tke = FP._computeEnstrophy(t, metrics)
    assert np.isfinite(tke) and tke > 0 and np.isfinite(enst) and enst > 0


def tiled_tree(num, flow="euler"):
    from fast_amd import connect
    zones, pairs, _ = cases.tiled_box((3, 5), (4, 2), (4,), num, flow=flow)
    t = tree_of(zones, num, model="Euler" if flow == "euler" else "NSLaminar")
    tz = Internal.getZones(t)
    for n, (zr, fr, zd, fd) in enumerate(pairs):
        pr, pd = zones[zr].param_int, zones[zd].param_int
        idr = connect.FACES.index(fr) + 1
        idd = connect.FACES.index(fd) + 1
        FP._addMatch(tz[zr], "match%d" % n, tz[zd][0],
                     FP._fen_to_pointrange(pr, idr, P.face_window(pr, fr)[1]),
                     FP._fen_to_pointrange(pd, idd, P.face_window(pd, fd)[1]))
    return t, zones, pairs


def test_gridconnectivity1to1_gives_the_tiling_raccords():
    """GridConnectivity1to1 nodes (PointRange / PointRangeDonor) read back as the same
    injection raccords as the tiling itself."""
    from fast_amd import connect
    num = H.numerics("roe_min", "explicit")
    t, zones, pairs = tiled_tree(num)
    b = Internal.getBases(t)[0]
    tz = Internal.getZones(t)
    got = FP._matches([(z, FP._host_zone(b, z)) for z in tz])
    ref = connect.build(zones, pairs)
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        assert (g.zr, g.zd) == (r.zr, r.zd)
        assert np.array_equal(g.rcv, r.rcv) and np.array_equal(g.dnr, r.dnr)


def test_partial_window_match():
    """A window covering part of a face fills only that part of the ghost slab."""
    from fast_amd import connect
    num = H.numerics("roe_min", "explicit")
    zones, _, _ = cases.tiled_box((3, 4), (6,), (5,), num)
    pr, pd = zones[0].param_int, zones[1].param_int
    rcv, dnr = connect.match(pr, "imax", pd, "imin", offset=(0, 0), window=((2, 4), (1, 3)))
    assert len(rcv) == 2 * 3 * 3
    assert rcv[0] == connect._lin(pr, 4, 2, 1) and dnr[0] == connect._lin(pd, 1, 2, 1)
    with pytest.raises(NotImplementedError):
        connect.from_pointranges(pr, [[1, 1], [1, 2], [1, 2]], pd, [[1, 1], [1, 2], [1, 2]], (2, 1, 3))


@pytest.mark.gpu
def test_warmup_compute_tiled_tree_matches_single_zone():
    """warmup + _compute on 6 blocks joined by GridConnectivity1to1 (explicit RK3, Euler)
    equal the single periodic zone stepped by the oracle, bit for bit."""
    import oracle
    orc = oracle.Oracle()
    num = H.numerics("ausmpred", "explicit", **H.SINGLE_BLOCK)
    t, zones, pairs = tiled_tree(num)
    t, tc, metrics = FP.warmup(t, None)
    assert len(FP.HOOK["raccords"]) == len(pairs)
    nit = 2
    FP._compute(t, metrics, 0, NIT=nit)
    _, _, W = cases.tiled_box((3, 5), (4, 2), (4,), num, flow="euler")
    whole = cases.whole_box((3, 5), (4, 2), (4,), W, num, flow="euler")
    orc.init_metric(whole)
    orc.apply_bc(whole, whole.ro)
    whole.ro_m1[:] = whole.ro; whole.ro_p1[:] = whole.ro
    for step in range(nit):
        for it in range(1, 4):
            orc.substep(whole, it, first_it=0 if step == 0 else 1)
        H.rotate(whole)
    Wg = whole.interior(whole.ro)
    ox, oy = np.cumsum([0, 3, 5]), np.cumsum([0, 4, 2])
    for n, z in enumerate(Internal.getZones(t)):
        a, bb = n % 2, n // 2
        fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
        got = np.stack([Internal.getNodeFromName1(fs, v)[1][2:-2, 2:-2, 2:-2].transpose(2, 1, 0)
                        for v in FP.VARS])
        assert np.array_equal(got, Wg[:, :, oy[bb]:oy[bb + 1], ox[a]:ox[a + 1]]), z[0]


def _with_own(t):
    """Descriptor vectors under .Solver#ownData, as allocate_metric leaves them."""
    b = Internal.getBases(t)[0]
    hz = []
    for z in Internal.getZones(t):
        h = FP._host_zone(b, z)
        o = Internal.createUniqueChild(z, ".Solver#ownData", "UserDefinedData_t")
        Internal.createUniqueChild(o, "Parameter_int", "DataArray_t", h.param_int)
        hz.append(h)
    return hz


def test_create_stress_nodes_windows():
    """One surface zone per selected wall BC, with its face window (face index n+1 on a max
    face), the face vertices of the NS zone and the 21 effort fields."""
    num = H.numerics("ausmpred", "bdf")
    t = tree_of([cases.channel_zone(12, 8, 4, num)], num)
    hz = _with_own(t)[0]
    teff = FP.createStressNodes(t, BC=["BCWallViscous"])
    zs = Internal.getZones(teff)
    assert len(zs) == 2
    ni, nj, nk = hz.dims
    wins = []
    for z in zs:
        pe = Internal.getNodeFromName1(Internal.getNodeFromName1(z, ".Solver#ownData"), "Parameter_int")[1]
        wins.append((int(pe[FP.EFF_IDIR]), [int(v) for v in pe[FP.EFF_LOOP:FP.EFF_LOOP + 6]]))
        fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
        assert Internal.getNodeFromName1(fs, "CoefPressure")[1].shape == (ni, nk, 1)
        x = Internal.getNodeFromName1(Internal.getNodeFromName1(z, "GridCoordinates"), "CoordinateY")[1]
        assert x.shape == (ni + 1, nk + 1, 1)
    assert wins == [(3, [1, ni, 1, 1, 1, nk]), (4, [1, ni, nj + 1, nj + 1, 1, nk])]
    # the surface vertices are the NS zone's wall vertices (ghost-inclusive numbering, 2 layers)
    g = Internal.getNodeFromName1(Internal.getZones(t)[0], "GridCoordinates")
    y = Internal.getNodeFromName1(g, "CoordinateY")[1]
    ys = Internal.getNodeFromName1(Internal.getNodeFromName1(zs[1], "GridCoordinates"), "CoordinateY")[1]
    assert np.array_equal(ys[:, :, 0], y[2:ni + 3, nj + 2, 2:nk + 3])


@pytest.mark.gpu
def test_compute_stress_matches_oracle():
    """_computeStress after warmup + _compute: the sums of the two channel walls equal the
    oracle's efforts of the same state, and the teff fields hold the per-face values."""
    import oracle
    orc = oracle.Oracle()
    num = H.numerics("ausmpred", "bdf")
    mk = lambda: cases.channel_zone(14, 8, 4, num)
    t = tree_of([mk()], num)
    t, tc, metrics = FP.warmup(t, None)
    FP._compute(t, metrics, 0, NIT=2)
    teff = FP.createStressNodes(t, BC=["BCWallViscous"])
    eff = FP._computeStress(t, teff, metrics, xyz_ref=(1.0, 0.0, 0.0))
    zo = mk()
    orc.init_metric(zo); orc.apply_bc(zo, zo.ro)
    zo.ro_m1[:] = zo.ro; zo.ro_p1[:] = zo.ro
    for step in range(2):
        orc.substep(zo, 1, first_it=0 if step == 0 else 1)
        H.rotate(zo)
    ni, nj, nk = zo.dims
    tot = np.zeros(11)
    for z, (idir, win) in zip(Internal.getZones(teff), [(3, [1, ni, 1, 1, 1, nk]), (4, [1, ni, nj + 1, nj + 1, 1, nk])]):
        e, f = orc.effort(zo, idir, win, (1.0, 0.0, 0.0))
        tot = tot + e[:11]
        fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
        cp = Internal.getNodeFromName1(fs, "CoefPressure")[1]
        assert np.array_equal(cp.ravel(order="F"), f[17])
    assert np.array_equal(eff, tot)


# ---------------------------------------------------------------------------------------
# Convergence history and restart files (CPU)
def test_create_convergence_history_nodes():
    """_createConvergenceHistory (PyTree.py:2152-2190): a GlobalConvergenceHistory per base and
    a ZoneConvergenceHistory per zone, valued 0, with IterationNumber and the four residual
    arrays (global Linf filled with -inf)."""
    num = H.numerics("ausmpred", "bdf")
    t = tree_of([cases.tgv_zone((6, 5, 4), num, name="a"), cases.tgv_zone((5, 5, 4), num, name="b")], num)
    Internal.getZones(t)[1][0] = "b"
    FP._createConvergenceHistory(t, 7)
    b = Internal.getBases(t)[0]
    g = Internal.getNodeFromName1(b, "GlobalConvergenceHistory")
    assert g[3] == "ConvergenceHistory_t" and Internal.getValue(g) == 0
    assert [c[0] for c in g[2]] == ["IterationNumber", "RSD_L2", "RSD_L2_diff", "RSD_oo", "RSD_oo_diff"]
    assert Internal.getNodeFromName1(g, "IterationNumber")[1].shape == (7,)
    assert np.all(Internal.getNodeFromName1(g, "RSD_oo")[1] == -np.inf)
    for z in Internal.getZones(t):
        h = Internal.getNodeFromName1(z, "ZoneConvergenceHistory")
        assert [c[0] for c in h[2]] == ["IterationNumber", "RSD_L2", "RSD_L2_diff", "RSD_oo", "RSD_oo_diff"]
        assert Internal.getNodeFromName1(h, "RSD_L2")[1].shape == (35,)


def _fill_history(t, nrec, seed):
    rng = np.random.default_rng(seed)
    for z in Internal.getZones(t):
        h = Internal.getNodeFromName1(z, "ZoneConvergenceHistory")
        h[1] = np.array([nrec], dtype=np.int32)
        Internal.getNodeFromName1(h, "IterationNumber")[1][:nrec] = np.arange(nrec) * 10
        for v in ("RSD_L2", "RSD_L2_diff", "RSD_oo", "RSD_oo_diff"):
            Internal.getNodeFromName1(h, v)[1][:5 * nrec] = rng.uniform(-8, 0, 5 * nrec)


def test_global_convergence_and_extract(tmp_path):
    """_calc_global_convergence (PyTree.py:3852-3910): Linf the maximum over the zones, L2 the
    average weighted by the zones' point counts; _extractConvergenceHistory writes one tecplot
    block per zone and per base (PyTree.py:2196-2270) with the records as fixed-point values."""
    num = H.numerics("ausmpred", "bdf")
    t = tree_of([cases.tgv_zone((6, 5, 4), num, name="a"), cases.tgv_zone((5, 5, 4), num, name="b")], num)
    Internal.getZones(t)[1][0] = "b"
    FP._createConvergenceHistory(t, 6)
    _fill_history(t, 6, 3)
    FP._calc_global_convergence(t)
    b = Internal.getBases(t)[0]
    g = Internal.getNodeFromName1(b, "GlobalConvergenceHistory")
    zs = Internal.getZones(t)
    hs = [Internal.getNodeFromName1(z, "ZoneConvergenceHistory") for z in zs]
    w = [int(np.prod(z[1][:, 0])) for z in zs]
    for v in ("RSD_oo", "RSD_oo_diff"):
        want = np.maximum(*[Internal.getNodeFromName1(h, v)[1] for h in hs])
        assert np.array_equal(Internal.getNodeFromName1(g, v)[1], want)
    for v in ("RSD_L2", "RSD_L2_diff"):
        want = (Internal.getNodeFromName1(hs[0], v)[1] * w[0] + Internal.getNodeFromName1(hs[1], v)[1] * w[1]) / (w[0] + w[1])
        np.testing.assert_allclose(Internal.getNodeFromName1(g, v)[1], want, rtol=1e-15)
    assert Internal.getValue(g) == 6
    out = tmp_path / "cvg.dat"
    FP._extractConvergenceHistory(t, str(out))
    lines = out.read_text().splitlines()
    assert lines[0].startswith("VARIABLES = it RO_l2")
    blocks = [l for l in lines if l.startswith("ZONE")]
    assert blocks == ['ZONE T="bloc Base/a " I=6 F=POINT', 'ZONE T="bloc Base/b " I=6 F=POINT',
                      'ZONE T="bloc Base " I=6 F=POINT']
    first = lines[2].split()
    assert first[0] == "1" and len(first) == 21
    h0 = hs[0]
    assert float(first[1]) == pytest.approx(Internal.getNodeFromName1(h0, "RSD_L2")[1][0], abs=1e-6)
    assert float(first[6]) == pytest.approx(Internal.getNodeFromName1(h0, "RSD_oo")[1][0], abs=1e-6)


def test_fortran_e_format():
    """conv2pytree.for writes its ascii records with Fortran E15.8."""
    assert FP._fortran_e(1.0) == " 0.10000000E+01"
    assert FP._fortran_e(-123.456) == "-0.12345600E+03"
    assert FP._fortran_e(2.5e-7) == " 0.25000000E-06"
    assert FP._fortran_e(0.0) == " 0.00000000E+00"


def test_save_load_tree_roundtrip(tmp_path):
    """saveTree / loadTree: the tree without its _P1 level and .Solver#ownData (as the
    reference's saveTree writes it), values and Fortran order kept, nothing pickled."""
    num = H.numerics("ausmpred", "bdf")
    t = tree_of([cases.tgv_zone((6, 5, 4), num, perturb=0.02)], num)
    z = Internal.getZones(t)[0]
    fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
    for v in FP.VARS:
        a = Internal.getNodeFromName1(fs, v)[1]
        Internal.createChild(fs, v + "_M1", "DataArray_t", a * 1.5)
        Internal.createChild(fs, v + "_P1", "DataArray_t", a * 2.0)
    Internal.createChild(z, ".Solver#ownData", "UserDefinedData_t")
    path = tmp_path / "restart.cgns"
    FP.saveTree(t, str(path))
    t2 = FP.loadTree(str(path))
    z2 = Internal.getZones(t2)[0]
    fs2 = Internal.getNodeFromName1(z2, "FlowSolution#Centers")
    assert Internal.getNodeFromName1(z2, ".Solver#ownData") is None
    assert all(Internal.getNodeFromName1(fs2, v + "_P1") is None for v in FP.VARS)
    for v in FP.VARS + [v + "_M1" for v in FP.VARS]:
        a, b = Internal.getNodeFromName1(fs, v)[1], Internal.getNodeFromName1(fs2, v)[1]
        assert a.shape == b.shape and np.array_equal(a, b) and b.flags.f_contiguous
    g, g2 = Internal.getNodeFromName1(z, "GridCoordinates"), Internal.getNodeFromName1(z2, "GridCoordinates")
    assert np.array_equal(Internal.getNodeFromName1(g, "CoordinateX")[1], Internal.getNodeFromName1(g2, "CoordinateX")[1])
    b2 = Internal.getBases(t2)[0]
    assert Internal.getValue(Internal.getNodeFromName(b2, "scheme")) == "ausmpred"
    t3, tc, ts, graph = FP.load(str(path))
    assert tc is None and ts is None and graph is None and Internal.getZones(t3)[0][0] == z[0]


# ---------------------------------------------------------------------------------------
# GPU: display, lazy tree refresh, restart
@pytest.mark.gpu
def test_display_temporal_criteria_contract(capsys):
    """display_temporal_criteria as the reference: None without gmres; with gmres the sum of
    the L2 norms over zones, equations and recorded sub-iterations and the largest Linf; the
    zone history records log10 of the first recorded sub-iteration's L2 / Linf and the drop to
    the last one (dpssiter.for, conv2pytree.for); format "store" prints nothing."""
    num = H.numerics("ausmpred", "bdf")
    t = tree_of([cases.tgv_zone((12, 10, 8), num, perturb=0.02, name="a"),
                 H.jittered_box((8, 9, 7), num, bcs=H.WALL_J)], num)
    Internal.getZones(t)[1][0] = "b"
    FP._setNum2Base(t, {"modulo_verif": 1})
    t, tc, metrics = FP.warmup(t, None)
    FP._createConvergenceHistory(t, 3)
    ctx = FP.HOOK["ctx"]
    for it in range(2):
        FP._compute(t, metrics, it)
        capsys.readouterr()
        assert FP.display_temporal_criteria(t, metrics, it, format="store") is None
        assert capsys.readouterr().out == ""
        hist = [ctx.residual_history(zid) for zid in FP.HOOK["ids"]]
        assert all(h.shape == (2, 10) for h in hist)
        r = FP.display_temporal_criteria(t, metrics, it, format="store", gmres=1)
        assert r["L2"] == pytest.approx(sum(h[:, :5].sum() for h in hist), rel=1e-15)
        assert r["Loo"] == max(h[:, 5:].max() for h in hist)
    for z, zid in zip(Internal.getZones(t), FP.HOOK["ids"]):
        h = Internal.getNodeFromName1(z, "ZoneConvergenceHistory")
        assert Internal.getValue(h) == 3                   # 2 iterations x 2 calls, capped at 3 records
        assert list(Internal.getNodeFromName1(h, "IterationNumber")[1]) == [0, 0, 1]
        last = ctx.residual_history(zid)
        rl2 = Internal.getNodeFromName1(h, "RSD_L2")[1][10:15]
        rl2d = Internal.getNodeFromName1(h, "RSD_L2_diff")[1][10:15]
        roo = Internal.getNodeFromName1(h, "RSD_oo")[1][10:15]
        np.testing.assert_array_equal(rl2, np.log10(np.maximum(last[0, :5], 1e-30)))
        np.testing.assert_array_equal(roo, np.log10(np.maximum(last[0, 5:], 1e-30)))
        np.testing.assert_array_equal(rl2d, np.log10(np.maximum(last[1, :5], 1e-30)) - rl2)
    FP.display_temporal_criteria(t, metrics, 1)
    out = capsys.readouterr().out
    assert "Res_L2 (zone=a, ssiter=  1   1)" in out and "Diff Ssi 1-  2" in out and "CFL_MOY_MAX_MIN" in out


@pytest.mark.gpu
def test_tree_edit_between_computes_takes_effect():
    """The state stays on the GPU between _compute calls; reading the tree refreshes it, and an
    edit made to the tree arrays then is the state the next _compute starts from."""
    import oracle
    orc = oracle.Oracle()
    num = H.numerics("ausmpred", "bdf")
    mk = lambda: cases.tgv_zone((12, 10, 8), num, perturb=0.02, name="a")
    t = tree_of([mk()], num)
    t, tc, metrics = FP.warmup(t, None)
    FP._compute(t, metrics, 0)
    z = Internal.getZones(t)[0]                 # reading the tree refreshes its arrays
    fs = Internal.getNodeFromName1(z, "FlowSolution#Centers")
    rho = Internal.getNodeFromName1(fs, "Density")[1]
    rho *= 1.01                                 # an edit of the N level, ghost cells included
    FP._compute(t, metrics, 1)
    zo = mk()
    orc.init_metric(zo); orc.apply_bc(zo, zo.ro)
    zo.ro_m1[:] = zo.ro; zo.ro_p1[:] = zo.ro
    orc.substep(zo, 1, first_it=0)
    H.rotate(zo)
    zo.ro[:zo.nd] *= 1.01
    orc.substep(zo, 1, first_it=1)
    H.rotate(zo)
    got = np.concatenate([Internal.getNodeFromName1(fs, v)[1].ravel(order="F") for v in FP.VARS])
    assert np.array_equal(zo.interior(got), zo.interior(zo.ro))


@pytest.mark.gpu
def test_compute_loop_keeps_the_state_on_the_gpu():
    """A loop of _compute calls that does not read the tree moves no field between host and
    GPU: after the first call (which uploads the tree warmup handed out) no field copy runs."""
    num = H.numerics("ausmpred", "bdf")
    t = tree_of([cases.tgv_zone((12, 10, 8), num, perturb=0.02, name="a")], num)
    t, tc, metrics = FP.warmup(t, None)
    FP._compute(t, metrics, 0)
    ctx = FP.HOOK["ctx"]
    calls = []
    get, set_ = ctx.get, ctx.set
    ctx.get = lambda *a: (calls.append("get"), get(*a))[1]
    ctx.set = lambda *a: (calls.append("set"), set_(*a))[1]
    try:
        for it in range(1, 6):
            FP._compute(t, metrics, it)
        assert calls == []
        FP._syncTree(t)
        assert calls.count("get") == 3
    finally:
        ctx.get, ctx.set = get, set_


@pytest.mark.gpu
def test_restart_from_saved_tree(tmp_path):
    """saveTree after some steps, loadTree, warmup: the restart has the N-1 level (FIRST_IT = 1,
    BDF2 from its first step) and continues exactly as the uninterrupted run."""
    num = H.numerics("ausmpred", "bdf")
    mk = lambda: cases.tgv_zone((12, 10, 8), num, perturb=0.02, name="a")
    t = tree_of([mk()], num)
    t, tc, metrics = FP.warmup(t, None)
    FP._compute(t, metrics, 0, NIT=2)
    FP.saveTree(t, str(tmp_path / "restart.cgns"))
    FP._compute(t, metrics, 2, NIT=2)
    ref = np.concatenate([Internal.getNodeFromName(t, v)[1].ravel(order="F") for v in FP.VARS])
    t2 = FP.loadTree(str(tmp_path / "restart.cgns"))
    t2, tc2, metrics2 = FP.warmup(t2, None)
    assert FP.HOOK["first_it"] == 1
    FP._compute(t2, metrics2, 2, NIT=2)
    got = np.concatenate([Internal.getNodeFromName(t2, v)[1].ravel(order="F") for v in FP.VARS])
    assert np.array_equal(got, ref)


@pytest.mark.gpu
def test_fasts_metric_name():
    """FastS.FastS.metric (dead in the reference) computes the metric of every zone."""
    from fast_amd import FastS
    num = H.numerics("ausmpred", "bdf")
    hz = H.jittered_box((8, 7, 6), num)
    t = tree_of([hz], num)
    FP.HOOK.clear()
    m = FastS.metric(t, {"scheme": "ausmpred"})
    import oracle
    orc = oracle.Oracle()
    orc.init_metric(hz)
    assert np.array_equal(m[0][0], hz.metric)
