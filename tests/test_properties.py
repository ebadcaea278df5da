"""Properties the mathematics fixes, checked on the oracle (CPU).

These pin parts of the oracle independently of the reference kernels:
* metric: closed Cartesian values, per-cell closure of the face vectors, total volume;
* Sutherland's law in closed form;
* free-stream preservation and telescoping conservation of the face balance on closed meshes;
* LU-SGS: the two sweeps equal a dense solve of (D - L) D^-1 (D - U) dW = D R built from the
  textbook Euler flux Jacobian (Fast/FastS/FastS/LU/invlu_l.for:62-215, invlu_u.for:62-265,
  lu_i_3dcart.for:1-44 define the operator; the matrices here are written from the Jacobian's
  definition, not from those loops).
"""
import numpy as np
import pytest

import oracle
from fast_amd import params as P, cases, mesh as M
from fast_amd.zone import HostZone
from tests import helpers as H


@pytest.fixture(scope="module")
def orc():
    return oracle.Oracle()


def metric_arrays(z):
    """ti, tj, tk as [3, nk, nj, ni] arrays and vol as [nk, nj, ni] on the metric grid of a
    3dfull zone (strides NIJK_MTR; index 0 is cell 1 - NIJK_MTR(4))."""
    p = z.param_int
    mi, mj, mk = int(p[P.NIJK_MTR]), int(p[P.NIJK_MTR + 1]), int(p[P.NIJK_MTR + 2])
    n = int(p[P.NDIMDX_MTR])
    assert mi == 1
    shp = (n // mk, mk // mj, mj)
    ti = z.ti.reshape(3, n).reshape((3,) + shp)
    tj = z.tj.reshape(3, n).reshape((3,) + shp)
    tk = z.tk.reshape(3, n).reshape((3,) + shp)
    return ti, tj, tk, z.vol.reshape(shp), int(p[P.NIJK_MTR + 3]), int(p[P.NIJK_MTR + 4])


def test_cartesian_metric_closed_form(orc):
    n = (8, 6, 5)
    z = cases.tgv_zone(n)
    orc.init_metric(z)
    h = 2 * np.pi / np.array(n)
    assert z.zone_type == P.ZONE_3DCART
    np.testing.assert_allclose(z.ti[0], h[1] * h[2], rtol=1e-14)
    np.testing.assert_allclose(z.tj[0], h[0] * h[2], rtol=1e-14)
    np.testing.assert_allclose(z.tk[0], h[0] * h[1], rtol=1e-14)
    np.testing.assert_allclose(z.vol[0], h.prod(), rtol=1e-14)


def test_metric_closure_and_total_volume(orc):
    """Jittered hexahedra: the six outward face vectors of every cell sum to zero and the cell
    volumes add up to the box volume (boundary vertices are not jittered)."""
    n = (7, 6, 5)
    x, y, zc = M.cart_vertices(n, (0, 0, 0), (2 * np.pi,) * 3)
    nv = (n[0] + 5, n[1] + 5, n[2] + 5)
    rng = np.random.default_rng(11)
    inner = (slice(3, -3),) * 3          # vertices strictly inside the physical box
    for a in (x, y, zc):
        A = a.reshape(nv, order="F")
        A[inner] += 0.05 * rng.uniform(-1, 1, A[inner].shape)
    z = HostZone("jb", n[0], n[1], n[2], H.numerics("ausmpred", "bdf"), cases.tgv_state(), x, y, zc,
                 bcs=cases.ALL_PERIODIC)
    orc.init_metric(z)
    assert z.zone_type == P.ZONE_3DFULL
    ti, tj, tk, vol, g, gk = metric_arrays(z)
    s = (slice(gk, gk + n[2]), slice(g, g + n[1]), slice(g, g + n[0]))
    sip = (slice(gk, gk + n[2]), slice(g, g + n[1]), slice(g + 1, g + 1 + n[0]))
    sjp = (slice(gk, gk + n[2]), slice(g + 1, g + 1 + n[1]), slice(g, g + n[0]))
    skp = (slice(gk + 1, gk + 1 + n[2]), slice(g, g + n[1]), slice(g, g + n[0]))
    for c in range(3):
        close = (ti[c][sip] - ti[c][s]) + (tj[c][sjp] - tj[c][s]) + (tk[c][skp] - tk[c][s])
        scale = np.abs(ti[c][s]).max() + np.abs(tj[c][s]).max() + np.abs(tk[c][s]).max()
        assert np.abs(close).max() < 1e-13 * scale
    np.testing.assert_allclose(vol[s].sum(), (2 * np.pi) ** 3, rtol=1e-12)


def test_sutherland_closed_form(orc):
    z = cases.tgv_zone((6, 5, 4), H.numerics("ausmpred", "bdf"), perturb=0.3)
    orc.init_metric(z)
    orc.apply_bc(z, z.ro)
    z.ro_m1[:] = z.ro; z.ro_p1[:] = z.ro
    orc.substep(z, 1, stop=1)
    st = z.state
    T = z.interior(z.ro)[4]
    mu_s, ts, cs = st["mus"], st["ts"], st["cs"]
    expect = mu_s * (T / ts) ** 1.5 * (ts + cs) / (T + cs)
    np.testing.assert_allclose(z.interior(z.xmut, 1)[0], expect, rtol=1e-13)


def uniform_zone(make):
    z = make()
    nd = z.nd
    w = np.array([1.2, 0.3, -0.2, 0.1, 1.1])
    for e in range(5):
        z.ro[e * nd:(e + 1) * nd] = w[e]
    z.ro_m1[:] = z.ro; z.ro_p1[:] = z.ro
    return z


@pytest.mark.parametrize("scheme", ["roe_min", "roe_kap", "ausmpred", "senseur"])
def test_free_stream_preservation(orc, scheme):
    """A uniform state on a closed curvilinear mesh has a zero face balance."""
    z = uniform_zone(lambda: H.jittered_box((7, 6, 5), H.numerics(scheme, "explicit")))
    orc.init_metric(z)
    orc.apply_bc(z, z.ro)
    orc.substep(z, 1, stop=1)
    r = z.interior(z.drodm)
    # rounding of face fluxes of size rho*c_p*T*|u|*S (energy) cancels to ~1e-16 of them;
    # a scheme that did not preserve the free stream would leave O(|flux| * jitter)
    st = z.state
    scale = np.array([1.2, 1.2, 1.2, 1.2, 1.2 * st["gamma"] * st["cv"] * 1.1]) * 0.4 * (2 * np.pi / 5) ** 2
    for e in range(5):
        assert np.abs(r[e]).max() < 1e-14 * scale[e] * 1e2


@pytest.mark.parametrize("scheme", ["roe_min", "roe_kap", "ausmpred", "senseur"])
@pytest.mark.parametrize("flow", ["euler"])
def test_periodic_conservation(orc, scheme, flow):
    """On a periodic Cartesian box every inviscid face flux enters one cell and leaves another
    (the two cells beside a periodic face see the same stencil): the face balance sums to zero
    up to rounding.  (The viscous flux at a boundary face also reads edge ghosts, which
    correct_coins.for fills by averaging rather than periodically, so it does not telescope
    exactly, in the reference as here.)"""
    z = cases.tgv_zone((9, 7, 6), H.numerics(scheme, "explicit"), perturb=0.05, seed=2, flow=flow)
    orc.init_metric(z)
    orc.apply_bc(z, z.ro)
    z.ro_m1[:] = z.ro; z.ro_p1[:] = z.ro
    orc.substep(z, 1, stop=1)
    r = z.interior(z.drodm).reshape(5, -1)
    tot = np.abs(r).sum(axis=1)
    assert (np.abs(r.sum(axis=1)) < 1e-13 * tot).all()


# --------------------------------------------------------------------------------------
def flux_jacobian(w, n, gam, cv):
    """dF.n/dU for the Euler flux, conservative U = (rho, rho u, rho v, rho w, rho E),
    in terms of the primitive state w = (rho, u, v, w, T)."""
    _, u, v, ww, T = w
    vel = np.array([u, v, ww])
    qn = n @ vel
    q2 = 0.5 * (vel @ vel)
    g1 = gam - 1.0
    H_ = gam * cv * T + q2
    phi = g1 * q2
    A = np.zeros((5, 5))
    A[0, 1:4] = n
    for a in range(3):
        A[1 + a, 0] = n[a] * phi - vel[a] * qn
        for b in range(3):
            A[1 + a, 1 + b] = n[b] * vel[a] - g1 * n[a] * vel[b]
        A[1 + a, 1 + a] += qn
        A[1 + a, 4] = g1 * n[a]
    A[4, 0] = qn * (phi - H_)
    A[4, 1:4] = n * H_ - g1 * vel * qn
    A[4, 4] = gam * qn
    return A


@pytest.mark.parametrize("scheme", ["roe_kap", "ausmpred"])
def test_lu_sgs_equals_dense_solve(orc, scheme):
    """3dcart periodic zone: the oracle's LU-SGS increment equals
    dW = (D - U)^-1 D (D - L)^-1 R, with D = coe5 I, and for a lower neighbour m of cell c
    L[c, m] = coe1_c/2 (A_d(W_m) n_d + coe_{2+d}(m)/2 I), for an upper one
    U[c, m] = -coe1_c/2 (A_d(W_m) n_d - coe_{2+d}(m)/2 I)."""
    n = (4, 3, 3)
    num = H.numerics(scheme, "bdf")
    mk = lambda: cases.tgv_zone(n, num, perturb=0.05, seed=4)
    za, zb = mk(), mk()
    for z in (za, zb):
        orc.init_metric(z)
        orc.apply_bc(z, z.ro)
        z.ro_m1[:] = z.ro; z.ro_p1[:] = z.ro
    orc.substep(za, 1, stop=2)     # R and the coefficients
    orc.substep(zb, 1, stop=3)     # after the two sweeps
    gam, cv = za.state["gamma"], za.state["cv"]
    N = n[0] * n[1] * n[2]
    R = za.interior(za.drodm).reshape(5, -1).T.ravel()     # cell-major (k, j, i), 5 per cell
    W = za.interior(za.ro).reshape(5, -1).T
    coe = za.interior(za.coe).reshape(5, -1).T
    nrm = [np.array([za.ti[0], 0, 0]), np.array([0, za.tj[0], 0]), np.array([0, 0, za.tk[0]])]
    cid = lambda i, j, k: (k * n[1] + j) * n[0] + i
    D = np.zeros((5 * N, 5 * N)); L = np.zeros_like(D); U = np.zeros_like(D)
    I5 = np.eye(5)
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                c = cid(i, j, k)
                D[5 * c:5 * c + 5, 5 * c:5 * c + 5] = coe[c, 4] * I5
                for d in range(3):
                    ijk = [i, j, k]
                    if ijk[d] > 0:            # lower neighbour inside the LU box
                        ijk[d] -= 1
                        m = cid(*ijk)
                        B = flux_jacobian(W[m], nrm[d], gam, cv) + 0.5 * coe[m, 1 + d] * I5
                        L[5 * c:5 * c + 5, 5 * m:5 * m + 5] = 0.5 * coe[c, 0] * B
                    ijk = [i, j, k]
                    if ijk[d] < n[d] - 1:     # upper neighbour inside the LU box
                        ijk[d] += 1
                        m = cid(*ijk)
                        B = flux_jacobian(W[m], nrm[d], gam, cv) - 0.5 * coe[m, 1 + d] * I5
                        U[5 * c:5 * c + 5, 5 * m:5 * m + 5] = -0.5 * coe[c, 0] * B
    dws = np.linalg.solve(D - L, R)
    dw = np.linalg.solve(D - U, D @ dws)
    got = zb.interior(zb.drodm).reshape(5, -1).T.ravel()
    np.testing.assert_allclose(got, dw, rtol=1e-11, atol=1e-13 * np.abs(dw).max())
