"""Shared case builders for the test suite."""
import numpy as np

from fast_amd import params as P, cases, mesh as M
from fast_amd.zone import HostZone


def jittered_box(n, num, flow="lamin", seed=3, amp=0.05, bcs=None, perturb=0.02):
    """TGV field on a [0, 2pi]^3 box whose interior vertices are jittered (a 3dfull zone)."""
    st = cases.tgv_state()
    x, y, z = M.cart_vertices(n, (0, 0, 0), (2 * np.pi,) * 3)
    rng = np.random.default_rng(seed)
    nv = (n[0] + 5, n[1] + 5, n[2] + 5)
    X = x.reshape(nv, order="F"); Y = y.reshape(nv, order="F"); Z = z.reshape(nv, order="F")
    h = 2 * np.pi / np.array(n)
    s = (slice(1, -1),) * 3
    X[s] += amp * h[0] * rng.uniform(-1, 1, X[s].shape)
    Y[s] += amp * h[1] * rng.uniform(-1, 1, Y[s].shape)
    Z[s] += amp * h[2] * rng.uniform(-1, 1, Z[s].shape)
    x, y, z = X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")
    bcs = cases.ALL_PERIODIC if bcs is None else bcs
    zz = HostZone("jbox", n[0], n[1], n[2], num, st, x, y, z, bcs=bcs, flow=flow)
    xc, yc, zc = M.cell_centers_from_vertices(x, y, z, nv)
    w = M.tgv_primitives(xc, yc, zc, st)
    w = w + perturb * rng.uniform(-1, 1, w.size)
    zz.ro[:] = w; zz.ro_m1[:] = w; zz.ro_p1[:] = w
    return zz


WALL_J = [("BCautoperiod", "imin", None), ("BCautoperiod", "imax", None),
          ("BCWallViscous", "jmin", None), ("BCWallViscous", "jmax", None),
          ("BCautoperiod", "kmin", None), ("BCautoperiod", "kmax", None)]
EXTRAP_I = [("BCExtrapolate", "imin", None), ("BCExtrapolate", "imax", None),
            ("BCWallViscous", "jmin", None), ("BCWallViscous", "jmax", None),
            ("BCautoperiod", "kmin", None), ("BCautoperiod", "kmax", None)]


# One cache block per zone: a box cut into zones equals the single zone bit for bit only when no
# cache-block edge changes the order in which a cell receives its face fluxes.
SINGLE_BLOCK = dict(cache_blocking_I=1000000, cache_blocking_J=1000000, cache_blocking_K=1000000)


def numerics(scheme, mode, **kw):
    if mode == "explicit":
        return P.Numerics(scheme=scheme, temporal_scheme="explicit", time_step=0.01, **kw)
    if mode == "bdf":
        return P.Numerics(scheme=scheme, temporal_scheme="implicit", ss_iteration=1, time_step=0.01, **kw)
    if mode == "bdf_ss2":
        return P.Numerics(scheme=scheme, temporal_scheme="implicit", ss_iteration=2, time_step=0.01, **kw)
    if mode == "local":
        return P.Numerics(scheme=scheme, temporal_scheme="implicit", ss_iteration=1,
                          time_step_nature="local", cfl=5.0, **kw)
    raise ValueError(mode)


def rotate(z):
    m1 = z.ro_m1.copy(); z.ro_m1[:] = z.ro; z.ro[:] = z.ro_p1; z.ro_p1[:] = m1


INVISCID_J = [("BCautoperiod", "imin", None), ("BCautoperiod", "imax", None),
              ("BCWallInviscid", "jmin", None), ("BCWallInviscid", "jmax", None),
              ("BCautoperiod", "kmin", None), ("BCautoperiod", "kmax", None)]
FARFIELD_OUTFLOW = [("BCFarfield", "imin", None), ("BCOutflow", "imax", None),
                    ("BCWallViscous", "jmin", None), ("BCFarfield", "jmax", None),
                    ("BCautoperiod", "kmin", None), ("BCautoperiod", "kmax", None)]


def inflow_outpres(n, p_exit=None, seed=5):
    """Inflow (free stream) at imin, outpres at imax with one exit pressure per face cell,
    viscous wall at jmin, farfield at jmax, periodic in k."""
    nj, nk = n[1], n[2]
    if p_exit is None:
        st = cases.tgv_state()
        p0 = st["p"]
        p_exit = p0 * (1.0 + 0.01 * np.random.default_rng(seed).uniform(-1, 1, nj * nk))
    return [("BCInflow", "imin", None), ("BCOutpres", "imax", [p_exit]),
            ("BCWallViscous", "jmin", None), ("BCFarfield", "jmax", None),
            ("BCautoperiod", "kmin", None), ("BCautoperiod", "kmax", None)]


def extruded_box(n, num, seed=7, amp=0.05, bcs=None, perturb=0.02, flow="euler"):
    """A 2-D jittered (x, y) mesh extruded along z with unequal k steps (a 3dfull zone whose
    k planes are independent 2-D problems: i and j face vectors lie in the (x, y) plane), with
    the TGV state perturbed and w = 0."""
    st = cases.tgv_state()
    x, y, z = M.cart_vertices(n, (0, 0, 0), (2 * np.pi,) * 3)
    rng = np.random.default_rng(seed)
    nv = (n[0] + 5, n[1] + 5, n[2] + 5)
    X = x.reshape(nv, order="F"); Y = y.reshape(nv, order="F"); Z = z.reshape(nv, order="F")
    h = 2 * np.pi / np.array(n)
    s = (slice(1, -1), slice(1, -1))
    dx = amp * h[0] * rng.uniform(-1, 1, X[s + (0,)].shape)
    dy = amp * h[1] * rng.uniform(-1, 1, Y[s + (0,)].shape)
    X[s] += dx[:, :, None]; Y[s] += dy[:, :, None]
    zk = Z[0, 0, :].copy()
    zk[1:-1] += amp * h[2] * rng.uniform(-1, 1, zk.size - 2)
    Z[:] = zk[None, None, :]
    x, y, z = X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")
    bcs = INVISCID_J if bcs is None else bcs
    zz = HostZone("ebox", n[0], n[1], n[2], num, st, x, y, z, bcs=bcs, flow=flow)
    xc, yc, zc = M.cell_centers_from_vertices(x, y, z, nv)
    w = M.tgv_primitives(xc, yc, zc, st)
    w = w + perturb * rng.uniform(-1, 1, w.size) * np.repeat([1., 1., 1., 0., 1.], w.size // 5)
    zz.ro[:] = w; zz.ro_m1[:] = w; zz.ro_p1[:] = w
    return zz
