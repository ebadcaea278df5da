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


# ------------------------------------------------------------------------------------------
# transposed Jacobian product of the AUSM+(P) o3 Euler face balance
VJP_CASES = {
    "jittered": lambda num: H.jittered_box((6, 5, 4), num, flow="euler"),
    "cart": lambda num: cases.tgv_zone((6, 5, 4), num, perturb=0.05, flow="euler"),
    "walls": lambda num: H.jittered_box((5, 6, 4), num, bcs=SLIP_ALL, flow="euler"),
    # i rows longer than the LDS-staged kernel takes: the two-pass path
    "long_i": lambda num: H.jittered_box((405, 3, 2), num, flow="euler"),
}


def vjp_zone(orc, case):
    z = VJP_CASES[case](H.numerics("ausmpred", "bdf"))
    orc.init_metric(z)
    orc.apply_bc(z, z.ro)
    return z


def scaled_direction(z, rng):
    scale = np.abs(z.ro.reshape(5, -1)).max(axis=1) * np.array([1e-1, 1e-2, 1e-2, 1e-2, 1e-1])
    return rng.standard_normal(5 * z.nd) * np.repeat(scale, z.nd)


@pytest.mark.parametrize("case", sorted(VJP_CASES))
def test_residual_vjp_dot_product(orc, case):
    """<lam, J v> over the interior (J v by central differences of the face balance) equals
    <J^T lam, v> for random directions v over every cell, ghosts included; the difference error
    falls with the step (1e-4 -> 1e-5) to the 1e-9 level."""
    z = vjp_zone(orc, case)
    rng = np.random.default_rng(1)
    lam = rng.standard_normal(5 * z.nd)
    g = orc.residual_vjp(z, lam)
    for _ in range(3):
        v = scaled_direction(z, rng)
        errs = []
        for h in (1e-4, 1e-5):
            jv = (orc.flux_balance(z, z.ro + h * v) - orc.flux_balance(z, z.ro - h * v)) / (2 * h)
            lhs = np.dot(z.interior(lam).ravel(), z.interior(jv).ravel())
            rhs = np.dot(g, v)
            errs.append(abs(lhs - rhs) / abs(rhs))
        assert errs[1] < 1e-9, errs


def test_residual_vjp_entries(orc):
    """Single entries of J^T lam, interior and ghost cells, against central differences of
    <lam, R(W + h e_m)>; cells outside every stencil (edge ghosts) get exactly zero."""
    z = vjp_zone(orc, "jittered")
    nd = z.nd
    rng = np.random.default_rng(2)
    lam = rng.standard_normal(5 * nd)
    g = orc.residual_vjp(z, lam)
    ni, nj, nk = z.dims
    idx = lambda i, j, k: (i + 1) + (j + 1) * (ni + 4) + (k + 1) * (ni + 4) * (nj + 4)
    cells = [idx(1, 1, 1), idx(3, 2, 2), idx(ni, nj, nk), idx(0, 2, 2), idx(-1, 3, 1),
             idx(ni + 2, 1, 3), idx(2, nj + 1, 2), idx(3, 3, -1)]
    for c in cells:
        for e in range(5):
            m = c + e * nd
            h = 1e-5 * max(abs(z.ro[m]), 1e-2)
            w = z.ro.copy(); w[m] += h
            rp = orc.flux_balance(z, w)
            w[m] -= 2 * h
            rm = orc.flux_balance(z, w)
            fd = np.dot(z.interior(lam).ravel(), z.interior((rp - rm) / (2 * h)).ravel())
            assert abs(fd - g[m]) <= 1e-7 * max(abs(g[m]), np.abs(g).max() * 1e-3), (c, e, fd, g[m])
    assert g[idx(0, 0, 1)] == 0. and g[idx(-1, nj + 2, 3)] == 0.


def test_residual_vjp_ignores_ghost_lambda(orc):
    """lam is read on the interior cells only (R is defined there)."""
    z = vjp_zone(orc, "cart")
    rng = np.random.default_rng(3)
    lam = rng.standard_normal(5 * z.nd)
    lam2 = rng.standard_normal(5 * z.nd)
    z.interior(lam2)[...] = z.interior(lam)
    assert np.array_equal(orc.residual_vjp(z, lam), orc.residual_vjp(z, lam2))


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(VJP_CASES))
def test_residual_vjp_gpu(orc, case):
    """fm_residual_vjp and fm_flux_balance equal the oracle bit for bit."""
    from fast_amd import _lib
    zo = vjp_zone(orc, case)
    zg = VJP_CASES[case](H.numerics("ausmpred", "bdf"))
    ctx = _lib.Context(0)
    zid = ctx.add_zone(zg)
    ctx.apply_bc(zid, 0)
    lam = np.random.default_rng(4).standard_normal(5 * zo.nd)
    assert np.array_equal(ctx.residual_vjp(zid, lam), orc.residual_vjp(zo, lam))
    rg, ro_ = ctx.flux_balance(zid), orc.flux_balance(zo)
    assert np.array_equal(zo.interior(rg), zo.interior(ro_))
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["ausmpred", "senseur", "roe_min"])
def test_flux_balance_gpu(orc, scheme):
    """fm_flux_balance (the RHS kernels with the time terms left out) equals the oracle's face
    balance on the interior, bit for bit, for every flux family (laminar: viscous fluxes too)."""
    from fast_amd import _lib
    num = H.numerics(scheme, "bdf")
    mk = lambda: H.jittered_box((7, 6, 5), num, bcs=H.WALL_J)
    zo, zg = mk(), mk()
    orc.init_metric(zo)
    orc.apply_bc(zo, zo.ro)
    zo.ro_m1[:] = zo.ro; zo.ro_p1[:] = zo.ro
    orc.substep(zo, 1, first_it=0, stop=1)        # viscosity and SENSEUR sensor of this state
    ctx = _lib.Context(0)
    zid = ctx.add_zone(zg)
    ctx.apply_bc(zid, 0)
    assert np.array_equal(zo.interior(ctx.flux_balance(zid)), zo.interior(orc.flux_balance(zo)))
    ctx.close()


@pytest.mark.gpu
def test_flux_balance_autograd(orc):
    """torch.autograd through fast_amd.adjoint.flux_balance: the gradient of <lam, R(W)> is the
    oracle's transposed product, bit for bit."""
    import torch
    from fast_amd import _lib, adjoint
    zo = vjp_zone(orc, "walls")
    zg = VJP_CASES["walls"](H.numerics("ausmpred", "bdf"))
    ctx = _lib.Context(0)
    zid = ctx.add_zone(zg)
    ctx.apply_bc(zid, 0)
    # host tensors: the kernels run on the GPU behind the C ABI (torch's own HIP runtime is not
    # initialised in the same process as the library's)
    W = torch.from_numpy(zo.ro.copy()).requires_grad_(True)
    lam = np.random.default_rng(5).standard_normal(5 * zo.nd)
    R = adjoint.flux_balance(W, ctx, zid)
    assert np.array_equal(zo.interior(R.detach().numpy()), zo.interior(orc.flux_balance(zo)))
    (R * torch.from_numpy(lam)).sum().backward()
    assert np.array_equal(W.grad.cpu().numpy(), orc.residual_vjp(zo, lam))
    ctx.close()
