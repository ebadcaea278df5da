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
        assert int(Internal.getNodeFromName1(h, "LastRecordedIteration")[1][0]) == 1
    for z in Internal.getZones(t):
        cfl = Internal.getValue(Internal.getNodeFromName(z, "CFL_minmaxmoy"))
        assert cfl[0] >= cfl[2] >= cfl[1] > 0
    enst, tke = FP._computeEnstrophy(t, metrics)
    assert np.isfinite(tke) and tke > 0 and np.isfinite(enst) and enst > 0
