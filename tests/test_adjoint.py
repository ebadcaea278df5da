"""Adjoint pieces: gradients of the wall pressure-force coefficients (compute_dpJ_dpW).

The reference codes dpJ_dpW_fluausm_euler_o3_2d.for only (AUSM+(P) o3, Euler, 2-D zone,
dpJ_dpW_fluausm_select.for option 213).  Its 3-D statement here is pinned three ways:
* against the compiled reference routine on extruded meshes, whose k planes are independent 2-D
  problems (w = 0, face vectors in the (x, y) plane), for the four i and j walls;
* against central differences of the wall effort (the oracle's restatement of bceffort.for,
  itself pinned against the compiled reference) with the ghost cells refilled by the slip-wall
  condition, on genuinely 3-D curvilinear and Cartesian zones, walls in i, j and k;
* the GPU kernel against the oracle, bit for bit.
"""
import numpy as np
import pytest

import oracle
from fast_amd import cases
from tests import helpers as H

requires_ref = pytest.mark.skipif(not oracle.Reference.available(),
                                  reason="reference kernels not built (make -C oracle ref)")

SLIP_IJ = [("BCWallInviscid", f, None) for f in ("imin", "imax", "jmin", "jmax")] + \
          [("BCautoperiod", "kmin", None), ("BCautoperiod", "kmax", None)]
SLIP_ALL = [("BCWallInviscid", f, None) for f in ("imin", "imax", "jmin", "jmax", "kmin", "kmax")]
COSA, SINA, SURFINV = 0.8, 0.6, 0.5


@pytest.fixture(scope="module")
def orc():
    return oracle.Oracle()


def windows(z, dirs=(1, 2, 3, 4, 5, 6)):
    ni, nj, nk = z.dims
    w = {1: [1, 1, 1, nj, 1, nk], 2: [ni + 1, ni + 1, 1, nj, 1, nk],
         3: [1, ni, 1, 1, 1, nk], 4: [1, ni, nj + 1, nj + 1, 1, nk],
         5: [1, ni, 1, nj, 1, 1], 6: [1, ni, 1, nj, nk + 1, nk + 1]}
    return [(d, w[d]) for d in dirs]


def prim_to_cons(q, cv):
    r = q[0]
    return np.array([r, r * q[1], r * q[2], r * q[3], r * (cv * q[4] + 0.5 * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2))])


def cons_to_prim(w, cv):
    r = w[0]
    u, v, ww = w[1] / r, w[2] / r, w[3] / r
    return np.array([r, u, v, ww, (w[4] / r - 0.5 * (u * u + v * v + ww * ww)) / cv])


@requires_ref
def test_dpj_dpw_vs_reference_extruded(orc):
    """The reference's 2-D routine, run plane by plane on an extruded mesh with slip walls on
    the i and j faces: the same wall cells are set, values to 1e-12 of each component's largest
    (the reference chains forward partials, the restatement runs a reverse pass: the rounding
    differs; the density component sums terms of opposite signs)."""
    ref = oracle.Reference()
    num = H.numerics("ausmpred", "bdf")
    n = (7, 8, 4)
    zr, zo = H.extruded_box(n, num, bcs=SLIP_IJ), H.extruded_box(n, num, bcs=SLIP_IJ)
    for be, z in ((ref, zr), (orc, zo)):
        be.init_metric(z)
        be.apply_bc(z, z.ro)
    assert zo.zone_type == 0
    assert np.abs(zo.ti.reshape(3, -1)[2]).max() < 1e-14 and np.abs(zo.tj.reshape(3, -1)[2]).max() < 1e-14
    for idir, ind in windows(zo, (1, 2, 3, 4)):
        rd, rl = ref.dpj_dpw(zr, idir, ind, COSA, SINA, SURFINV)
        od, ol = orc.dpj_dpw(zo, idir, ind, COSA, SINA, SURFINV)
        for a, b in ((rd, od), (rl, ol)):
            A, B = a.reshape(5, -1), b.reshape(5, -1)
            assert np.array_equal(np.any(A != 0, axis=0), np.any(B != 0, axis=0)), idir
            assert np.count_nonzero(np.any(B != 0, axis=0)) == 2 * (ind[1] - ind[0] + 1) * \
                (ind[3] - ind[2] + 1) * (ind[5] - ind[4] + 1)
            assert not B[3].any()                      # w = 0 and no z component of the normal
            for e in (0, 1, 2, 4):
                np.testing.assert_allclose(B[e], A[e], rtol=0, atol=1e-12 * np.abs(A[e]).max())


def fd_check(orc, z, idir, ind, ncells=3, hrel=1e-5):
    """Central differences of J = s (cos F_x + sin F_y) (and the lift) over the window's faces,
    the state perturbed in one conservative variable of one wall cell and the ghost cells refilled
    by the zone's boundary conditions; returns the largest error relative to the gradient."""
    nd = z.nd
    cv = z.param_real[2]
    norm = 2. / (z.param_real[3] * z.param_real[5] ** 2)
    dcd, dcl = orc.dpj_dpw(z, idir, ind, COSA, SINA, SURFINV)

    def J(ro):
        r = ro.copy()
        orc.apply_bc(z, r)
        e = orc.effort(z, idir, ind, rop=r, fields=False)
        return norm * SURFINV * np.array([COSA * e[8] + SINA * e[9], COSA * e[9] - SINA * e[8]])

    cells = np.nonzero(np.any(dcd.reshape(5, -1) != 0, axis=0))[0]
    rng = np.random.default_rng(idir)
    worst = 0.
    for c in rng.choice(cells, size=min(ncells, cells.size), replace=False):
        w = prim_to_cons(z.ro[c::nd][:5], cv)
        g = np.zeros((2, 5))
        for e in range(5):
            h = hrel * max(abs(w[e]), 1e-2)
            out = []
            for sgn in (1, -1):
                w2 = w.copy()
                w2[e] += sgn * h
                ro = z.ro.copy()
                ro[c::nd][:5] = cons_to_prim(w2, cv)
                out.append(J(ro))
            g[:, e] = (out[0] - out[1]) / (2 * h)
        an = np.array([dcd[c::nd][:5], dcl[c::nd][:5]])
        worst = max(worst, np.abs(an - g).max() / np.abs(g).max())
    return worst


@pytest.mark.parametrize("mesh", ["jittered", "cart"])
def test_dpj_dpw_central_differences(orc, mesh):
    """3-D zones (w != 0, slip walls all round): the gradient equals central differences
    of the wall force, with the difference error falling as the square of the step (relative
    steps 1e-3 and 1e-4: truncation-bound; below 1e-4 the rounding of the force sums, ~1e-16 of
    the free-stream pressure force, takes over) -- a wrong term would leave a floor instead."""
    num = H.numerics("ausmpred", "bdf")
    if mesh == "cart":
        z = cases.tgv_zone((6, 7, 5), num, perturb=0.05, seed=3, flow="euler", bcs=SLIP_ALL)
    else:
        z = H.jittered_box((6, 7, 5), num, bcs=SLIP_ALL, flow="euler")
    orc.init_metric(z)
    assert z.zone_type == (2 if mesh == "cart" else 0)
    orc.apply_bc(z, z.ro)
    # k walls carry no force in the (x, y) plane of the incidence: J and its gradient vanish there
    for idir, ind in windows(z, (1, 2, 3, 4)):
        # one face row keeps the rounding of the force sums small
        ind = list(ind)
        t = [a for a in range(3) if a != (idir - 1) // 2][0]
        ind[2 * t + 1] = ind[2 * t]
        e3, e4 = fd_check(orc, z, idir, ind, hrel=1e-3), fd_check(orc, z, idir, ind, hrel=1e-4)
        assert e4 < 5e-6 and e3 > 50 * e4, (idir, e3, e4)


def test_dpj_dpw_scope(orc):
    """Schemes other than AUSM+(P) o3 Euler are refused, as the reference's select refuses them."""
    z = H.jittered_box((5, 5, 4), H.numerics("roe_min", "bdf"), bcs=SLIP_ALL, flow="euler")
    orc.init_metric(z)
    with pytest.raises(RuntimeError):
        orc.dpj_dpw(z, 3, windows(z, (3,))[0][1], COSA, SINA, SURFINV)


@pytest.mark.gpu
@pytest.mark.parametrize("mesh", ["jittered", "cart"])
def test_dpj_dpw_gpu(orc, mesh):
    """fm_dpj_dpw equals the oracle bit for bit; cells off the walls keep the caller's values."""
    from fast_amd import _lib
    num = H.numerics("ausmpred", "bdf")
    if mesh == "cart":
        mk = lambda: cases.tgv_zone((9, 8, 7), num, perturb=0.05, seed=3, flow="euler", bcs=SLIP_ALL)
    else:
        mk = lambda: H.jittered_box((9, 8, 7), num, bcs=SLIP_ALL, flow="euler")
    zo, zg = mk(), mk()
    orc.init_metric(zo)
    orc.apply_bc(zo, zo.ro)
    ctx = _lib.Context(0)
    zid = ctx.add_zone(zg)
    ctx.apply_bc(zid, 0)
    for idir, ind in windows(zo):
        od = np.full(5 * zo.nd, 7.0); ol = np.full(5 * zo.nd, -3.0)
        orc.dpj_dpw(zo, idir, ind, COSA, SINA, SURFINV, dcd=od, dcl=ol)
        gd = np.full(5 * zo.nd, 7.0); gl = np.full(5 * zo.nd, -3.0)
        ctx.dpj_dpw(zid, idir, ind, COSA, SINA, SURFINV, gd, gl)
        assert np.array_equal(od, gd), (mesh, idir)
        assert np.array_equal(ol, gl), (mesh, idir)
    ctx.close()


@pytest.mark.gpu
def test_compute_dpj_dpw_pytree(orc):
    """warmup(..., Adjoint=True) creates the dpCDp_dp* / dpCLp_dp* fields (zero) and
    _compute_dpJ_dpW fills the wall cells of the teff windows as the oracle does."""
    from fast_amd import Internal, PyTree as FP
    from tests.test_pytree import tree_of
    num = H.numerics("ausmpred", "bdf")
    mk = lambda: H.jittered_box((8, 7, 5), num, bcs=H.INVISCID_J, flow="euler")
    t = tree_of([mk()], num, model="Euler")
    t, tc, metrics = FP.warmup(t, None, Adjoint=True)
    teff = FP.createStressNodes(t, BC=["BCWallInviscid"])
    FP._compute_dpJ_dpW(t, teff, metrics, COSA, SINA, SURFINV)
    zo = mk()
    orc.init_metric(zo)
    orc.apply_bc(zo, zo.ro)
    dcd, dcl = np.zeros(5 * zo.nd), np.zeros(5 * zo.nd)
    for idir, ind in windows(zo, (3, 4)):
        orc.dpj_dpw(zo, idir, ind, COSA, SINA, SURFINV, dcd=dcd, dcl=dcl)
    fs = Internal.getNodeFromName1(Internal.getZones(t)[0], "FlowSolution#Centers")
    for c, ref in (("cd", dcd), ("cl", dcl)):
        for e, v in enumerate(FP.ADJ_VARS[c]):
            got = Internal.getNodeFromName1(fs, v)[1].ravel(order="F")
            assert np.array_equal(got, ref[e * zo.nd:(e + 1) * zo.nd]), v
    assert np.count_nonzero(dcd) > 0
