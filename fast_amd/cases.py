"""The benchmark workloads as host zones (meshes, numerics, reference states, BCs, fields).

* Taylor-Green vortex: periodic box [0, 2 pi]^3, Mach 0.1, Re 1600, Pr 0.72, gamma 1.4
  (``Fast/FastS/test/explicitLocalTgv_t1.py:86-90`` for the initial field, autoperiodic BCs as
  ``Fast/FastS/test/tgvMpi.py:66-74``).
* Channel with a bump: lower wall y_w(x) = 0.05 exp(-((x-1.5)/0.2)^2), walls stretched from a
  first-cell height of 1e-4 with ratio 1.1, +-1 % interior node jitter (seed 0), Mach 0.5,
  Re 1e5, viscous walls in j, periodic in k.
"""
import math
import numpy as np

from . import params as P
from . import mesh as M
from .zone import HostZone

TWO_PI = 2.0 * math.pi
ALL_PERIODIC = [("BCautoperiod", f, None) for f in ("imin", "imax", "jmin", "jmax", "kmin", "kmax")]


def tgv_state():
    # mu at T0 equals rho0 U0 L / Re with L = 1, U0 = 1
    return P.reference_state(mach=0.1, reynolds=1600.0, prandtl=0.72, gamma=1.4)


def tgv_zone(n, numerics=None, perturb=0.0, seed=0, flow="lamin", name="tgv", bcs=None):
    """TGV zone of n^3 (or n = (ni,nj,nk)) cells; ``perturb`` adds a seeded relative noise to
    every primitive (used by the parity tests so that no term of the fluxes vanishes); ``bcs``
    replaces the periodic boundary conditions."""
    if np.isscalar(n):
        n = (n, n, n)
    num = P.Numerics() if numerics is None else numerics
    st = tgv_state()
    lo, hi = (0.0, 0.0, 0.0), (TWO_PI, TWO_PI, TWO_PI)
    x, y, z = M.cart_vertices(n, lo, hi)
    zone = HostZone(name, n[0], n[1], n[2], num, st, x, y, z, bcs=ALL_PERIODIC if bcs is None else bcs,
                    flow=flow)
    xc, yc, zc = M.cell_centers_cart(n, lo, hi)
    w = M.tgv_primitives(xc, yc, zc, st)
    if perturb > 0.0:
        rng = np.random.default_rng(seed)
        nd = zone.nd
        scale = np.repeat([st["ro"], st["u"], st["u"], st["u"], st["T"]], nd)
        w = w + perturb * scale * rng.uniform(-1.0, 1.0, w.size)
    zone.ro[:] = w
    zone.ro_m1[:] = w
    zone.ro_p1[:] = w
    return zone


def tgv_dt(zone, cfl=0.5):
    """Explicit time step dt = cfl * h / max(|u| + c) from the field (the TGV 64^3 explicit case)."""
    st = zone.state
    nd = zone.nd
    ro = zone.interior(zone.ro)
    c = np.sqrt(st["gamma"] * st["cv"] * (st["gamma"] - 1.0) * ro[4])
    vmax = np.max(np.abs(ro[1:4]).max(axis=0) + c)
    h = TWO_PI / zone.dims[0]
    return cfl * h / vmax


def channel_state(mach=0.5, reynolds=1.0e5):
    return P.reference_state(mach=mach, reynolds=reynolds, prandtl=0.72, gamma=1.4,
                             length=1.0, direction=(1.0, 0.0, 0.0))


def channel_zone(ni, nj, nk, numerics=None, jitter=0.01, seed=0, name="channel",
                 inflow="BCFarfield", outflow="BCFarfield"):
    num = P.Numerics(scheme="ausmpred", slope="o3", temporal_scheme="implicit",
                     ss_iteration=1, time_step_nature="local", cfl=10.0) if numerics is None else numerics
    st = channel_state()
    x, y, z = M.channel_bump_vertices(ni, nj, nk, jitter=jitter, seed=seed)
    bcs = [(inflow, "imin", None), (outflow, "imax", None),
           ("BCWallViscous", "jmin", None), ("BCWallViscous", "jmax", None),
           ("BCautoperiod", "kmin", None), ("BCautoperiod", "kmax", None)]
    zone = HostZone(name, ni, nj, nk, num, st, x, y, z, bcs=bcs)
    nd = zone.nd
    w = np.concatenate([np.full(nd, st["ro"]), np.full(nd, st["vx"]), np.full(nd, st["vy"]),
                        np.full(nd, st["vz"]), np.full(nd, st["T"])])
    zone.ro[:] = w; zone.ro_m1[:] = w; zone.ro_p1[:] = w
    return zone


def tiled_box(nx, ny, nz, numerics=None, h=0.125, perturb=0.02, seed=0, flow="lamin",
              periodic=(True, True, True), bcs_outer="BCExtrapolate"):
    """A box of nx[a] x ny[b] x nz[c] blocks (Cartesian, spacing h, a power of two so that
    every block sees exactly the same coordinates differences), with a TGV-like field plus a
    seeded perturbation defined on the whole box; faces between blocks (and across the box
    where ``periodic``) are 1-to-1 matches, other outer faces get ``bcs_outer``.
    Returns (zones, pairs) with pairs for fast_amd.connect.build."""
    from . import connect
    num = P.Numerics() if numerics is None else numerics
    st = tgv_state()
    N = (sum(nx), sum(ny), sum(nz))
    blocks, pairs, outer = connect.tiling(nx, ny, nz, periodic)
    rng = np.random.default_rng(seed)
    # global cell-centred field, i fastest
    xc = (np.arange(N[0]) + 0.5) * h
    yc = (np.arange(N[1]) + 0.5) * h
    zc = (np.arange(N[2]) + 0.5) * h
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    L = np.array(N) * h
    sx, sy, sz = 2 * np.pi / L
    W = np.stack([np.full(X.shape, st["ro"]),
                  st["u"] * np.sin(sx * X) * np.cos(sy * Y) * np.cos(sz * Z),
                  -st["u"] * np.cos(sx * X) * np.sin(sy * Y) * np.cos(sz * Z),
                  np.zeros(X.shape), np.full(X.shape, st["T"])])
    scale = np.array([st["ro"], st["u"], st["u"], st["u"], st["T"]])[:, None, None, None]
    W = W + perturb * scale * rng.uniform(-1.0, 1.0, W.shape)
    ox = np.concatenate([[0], np.cumsum(nx)]); oy = np.concatenate([[0], np.cumsum(ny)])
    oz = np.concatenate([[0], np.cumsum(nz)])
    outer_faces = {}
    for b, f in outer:
        outer_faces.setdefault(b, []).append(f)
    zones = []
    for n, (a, b, c) in enumerate(blocks):
        dims = (nx[a], ny[b], nz[c])
        lo = (ox[a] * h, oy[b] * h, oz[c] * h)
        hi = (lo[0] + dims[0] * h, lo[1] + dims[1] * h, lo[2] + dims[2] * h)
        x, y, zz = M.cart_vertices(dims, lo, hi)
        bcs = [(bcs_outer, f, None) for f in outer_faces.get(n, [])]
        zone = HostZone("blk%d" % n, dims[0], dims[1], dims[2], num, st, x, y, zz, bcs=bcs, flow=flow)
        nd = zone.nd
        full = np.zeros((5, dims[2] + 4, dims[1] + 4, dims[0] + 4))
        full[:, 2:-2, 2:-2, 2:-2] = W[:, ox[a]:ox[a + 1], oy[b]:oy[b + 1], oz[c]:oz[c + 1]].transpose(0, 3, 2, 1)
        zone.ro[:] = full.reshape(5 * nd)
        zone.ro_m1[:] = zone.ro
        zone.ro_p1[:] = zone.ro
        zones.append(zone)
    return zones, pairs, W


def whole_box(nx, ny, nz, W, numerics=None, h=0.125, flow="lamin", periodic=True):
    """The single-zone counterpart of tiled_box (same field, periodic in every direction)."""
    num = P.Numerics() if numerics is None else numerics
    st = tgv_state()
    N = (sum(nx), sum(ny), sum(nz))
    x, y, zz = M.cart_vertices(N, (0., 0., 0.), (N[0] * h, N[1] * h, N[2] * h))
    zone = HostZone("box", N[0], N[1], N[2], num, st, x, y, zz, bcs=ALL_PERIODIC, flow=flow)
    full = np.zeros((5, N[2] + 4, N[1] + 4, N[0] + 4))
    full[:, 2:-2, 2:-2, 2:-2] = W.transpose(0, 3, 2, 1)
    zone.ro[:] = full.reshape(5 * zone.nd)
    zone.ro_m1[:] = zone.ro
    zone.ro_p1[:] = zone.ro
    return zone
