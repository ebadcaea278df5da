"""Structured meshes and initial fields for the benchmark cases.

Cassiopee's ``Generator.cart`` / ``Converter.addGhostCells`` / ``Initiator`` are not available;
these helpers build the same objects directly: vertex coordinates *including* the ghost layers
(the layout ``allocate_metric`` expects, ``Fast/FastS/FastS/Metric/allocate_metric.cpp:126-141``),
stored Fortran-ordered (i fastest) as flat float64 arrays.
"""
import numpy as np


def cart_vertices(n, lo, hi, nfic=2):
    """Vertices of a uniform Cartesian block with ``nfic`` ghost layers, i-fastest flat arrays."""
    axes = []
    for d in range(3):
        h = (hi[d] - lo[d]) / n[d]
        iv = np.arange(-nfic, n[d] + nfic + 1, dtype=np.float64)
        axes.append(lo[d] + iv * h)
    X, Y, Z = np.meshgrid(axes[0], axes[1], axes[2], indexing="ij")
    return (np.asfortranarray(X).ravel(order="F"), np.asfortranarray(Y).ravel(order="F"),
            np.asfortranarray(Z).ravel(order="F"))


def cell_centers_cart(n, lo, hi, nfic=2):
    """Cell centres (including ghosts) of the uniform block, i-fastest flat arrays."""
    axes = []
    for d in range(3):
        h = (hi[d] - lo[d]) / n[d]
        ic = np.arange(-nfic, n[d] + nfic, dtype=np.float64) + 0.5
        axes.append(lo[d] + ic * h)
    X, Y, Z = np.meshgrid(axes[0], axes[1], axes[2], indexing="ij")
    return X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")


def stretched(n, y0, y1, h0, ratio):
    """1-D node distribution on [y0, y1] clustered at both ends with first step h0 and growth
    ``ratio`` (capped so the middle stays uniform), n cells."""
    half = n // 2
    steps = []
    h = h0
    for _ in range(half):
        steps.append(h)
        h *= ratio
    steps = np.array(steps)
    mid = n - 2 * half
    full = np.concatenate([steps, np.full(mid, steps[-1]), steps[::-1]])
    # clip the geometric growth where it exceeds the uniform spacing, then rescale to length
    L = y1 - y0
    hmax = L / n * 4.0
    full = np.minimum(full, hmax)
    full *= L / full.sum()
    return y0 + np.concatenate([[0.0], np.cumsum(full)])


def extend_ghosts_1d(v, nfic):
    """Linear extrapolation of a node distribution by ``nfic`` nodes on each side."""
    lo = [v[0] - (k + 1) * (v[1] - v[0]) for k in range(nfic)][::-1]
    hi = [v[-1] + (k + 1) * (v[-1] - v[-2]) for k in range(nfic)]
    return np.concatenate([lo, v, hi])


def channel_bump_vertices(ni, nj, nk, lx=3.0, ly=1.0, lz=0.25, h0=1.0e-4, ratio=1.1,
                          jitter=0.01, seed=0, nfic=2):
    """Channel with a Gaussian bump y_w(x) = 0.05 exp(-((x-1.5)/0.2)^2) on the lower wall,
    wall-normal stretching from both walls, interior node jitter of +-``jitter`` of the local
    spacing (numpy default_rng(seed)), and ``nfic`` ghost layers extrapolated linearly."""
    xs = np.linspace(0.0, lx, ni + 1)
    zs = np.linspace(0.0, lz, nk + 1)
    eta = stretched(nj, 0.0, 1.0, h0 / ly, ratio)
    X = np.empty((ni + 1, nj + 1, nk + 1)); Y = np.empty_like(X); Z = np.empty_like(X)
    yw = 0.05 * np.exp(-((xs - 1.5) / 0.2) ** 2)
    for i in range(ni + 1):
        y = yw[i] + eta * (ly - yw[i])
        X[i, :, :] = xs[i]
        Y[i, :, :] = y[:, None]
    Z[:, :, :] = zs[None, None, :]
    rng = np.random.default_rng(seed)
    if jitter > 0:
        dx = lx / ni; dz = lz / nk
        dy = np.diff(Y, axis=1)
        hy = np.minimum(np.concatenate([dy[:, :1], dy], axis=1), np.concatenate([dy, dy[:, -1:]], axis=1))
        r = rng.uniform(-1.0, 1.0, size=(3,) + X.shape)
        s = (slice(1, -1),) * 3
        X[s] += jitter * dx * r[0][s]
        Y[s] += jitter * hy[s] * r[1][s]
        Z[s] += jitter * dz * r[2][s]
    # ghost layers: linear extrapolation along each index direction
    def ext(A, axis):
        lo = [A.take([0], axis) - (k + 1) * (A.take([1], axis) - A.take([0], axis)) for k in range(nfic)][::-1]
        hi = [A.take([-1], axis) + (k + 1) * (A.take([-1], axis) - A.take([-2], axis)) for k in range(nfic)]
        return np.concatenate(lo + [A] + hi, axis=axis)
    out = []
    for A in (X, Y, Z):
        for ax in range(3):
            A = ext(A, ax)
        out.append(np.asfortranarray(A).ravel(order="F"))
    return tuple(out)


def cell_centers_from_vertices(x, y, z, nv):
    """Average of the 8 vertices of each cell (cells including ghosts)."""
    X = x.reshape(nv, order="F"); Y = y.reshape(nv, order="F"); Z = z.reshape(nv, order="F")
    def c(A):
        return 0.125 * (A[:-1, :-1, :-1] + A[1:, :-1, :-1] + A[:-1, 1:, :-1] + A[1:, 1:, :-1]
                        + A[:-1, :-1, 1:] + A[1:, :-1, 1:] + A[:-1, 1:, 1:] + A[1:, 1:, 1:])
    return c(X).ravel(order="F"), c(Y).ravel(order="F"), c(Z).ravel(order="F")


def tgv_primitives(xc, yc, zc, state, U0=None):
    """Taylor-Green vortex primitive field (rho, u, v, w, T) at the given cell centres.

    u = U0 sin x cos y cos z, v = -U0 cos x sin y cos z, w = 0,
    p = p0 + rho0 U0^2/16 (cos 2x + cos 2y)(cos 2z + 2), T = T0, rho = p / (R T0)
    (the initialisation of Fast/FastS/test/explicitLocalTgv_t1.py:86-90)."""
    U0 = state["u"] if U0 is None else U0
    ro0, T0, p0 = state["ro"], state["T"], state["p"]
    R = state["cv"] * (state["gamma"] - 1.0)
    u = U0 * np.sin(xc) * np.cos(yc) * np.cos(zc)
    v = -U0 * np.cos(xc) * np.sin(yc) * np.cos(zc)
    w = np.zeros_like(u)
    p = p0 + ro0 * U0 * U0 / 16.0 * (np.cos(2 * xc) + np.cos(2 * yc)) * (np.cos(2 * zc) + 2.0)
    T = np.full_like(u, T0)
    ro = p / (R * T0)
    return np.concatenate([ro, u, v, w, T])
