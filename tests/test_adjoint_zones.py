"""Transposed residual Jacobian of several zones joined by injection transfers (the transpose of
the ghost fill of FastC/Com/setInterpTransfersFast.cpp:440-760, reverse scatter-add onto the
donor cells), and of a periodic zone joined to itself.

The forward map is F(interior states of all zones) = face balance of every zone after the
transfers have filled the ghost cells; the oracle's transpose is checked against central
differences of F, and the GPU entry point fm_residual_vjp_zones against the oracle bit for bit."""
import numpy as np
import pytest

import oracle
from fast_amd import cases, connect
from tests import helpers as H

SPLITS = {
    "tiled": ((3, 5), (4, 2), (4,)),      # 4 blocks, periodic; z joined to itself
    "single": ((6,), (5,), (4,)),         # one zone, periodic in all three directions
}


@pytest.fixture(scope="module")
def orc():
    return oracle.Oracle()


def make(split, extra=False):
    nx, ny, nz = SPLITS[split]
    num = H.numerics("ausmpred", "bdf")
    zones, pairs, W = cases.tiled_box(nx, ny, nz, num, flow="euler")
    racs = connect.build(zones, pairs)
    if extra:
        racs.append(overwrite_raccord(zones[0]))
    return zones, racs


def overwrite_raccord(z):
    """A raccord of zone 0 onto itself that fills both imin ghost layers of a window from the
    first interior cell (repeated donors), over ghosts an earlier raccord filled: its transpose
    takes the point-by-point path and the overwrite must cut the earlier raccord's share."""
    p = z.param_int
    J, K = np.meshgrid(np.arange(1, 4), np.arange(1, 3), indexing="ij")
    rcv = np.concatenate([connect._lin(p, 0, J, K).ravel(), connect._lin(p, -1, J, K).ravel()])
    dnr = np.concatenate([connect._lin(p, 1, J, K).ravel()] * 2)
    return connect.Raccord(0, 0, rcv, dnr)


def filled(zones, racs, states):
    lv = [s.copy() for s in states]
    connect.apply(zones, racs, lv)
    return lv


def interior_mask(z):
    m = np.zeros(5 * z.nd, dtype=bool)
    z.interior(m)[...] = True
    return m


@pytest.mark.parametrize("split,extra", [("tiled", False), ("single", False), ("tiled", True)])
def test_residual_vjp_zones_dot_product(orc, split, extra):
    """<lam, F'(W) v> by central differences equals <F'(W)^T lam, v> for directions v over the
    interior cells of every zone (error 1e-9 at h = 1e-5); every ghost cell gets exactly zero."""
    zones, racs = make(split, extra)
    rng = np.random.default_rng(7)
    for z in zones:
        orc.init_metric(z)
    base = filled(zones, racs, [z.ro for z in zones])
    lams = [rng.standard_normal(5 * z.nd) for z in zones]
    g = orc.residual_vjp_zones(zones, racs, lams, base)
    masks = [interior_mask(z) for z in zones]
    for z, gz, m in zip(zones, g, masks):
        assert not np.any(gz[~m]), z.name
    for _ in range(2):
        vs = []
        for z, m, b in zip(zones, masks, base):
            mag = np.abs(b.reshape(5, -1)).max(axis=1)
            v = (rng.uniform(-1, 1, (5, z.nd)) * mag[:, None]).ravel()
            v[~m] = 0.0
            vs.append(v)
        rhs = sum(np.dot(gz, v) for gz, v in zip(g, vs))
        h = 1e-5
        plus = filled(zones, racs, [b + h * v for b, v in zip(base, vs)])
        minus = filled(zones, racs, [b - h * v for b, v in zip(base, vs)])
        lhs = 0.0
        for z, lam, a, c in zip(zones, lams, plus, minus):
            jv = (orc.flux_balance(z, a) - orc.flux_balance(z, c)) / (2 * h)
            lhs += np.dot(z.interior(lam).ravel(), z.interior(jv).ravel())
        assert abs(lhs - rhs) / abs(rhs) < 1e-9, (lhs, rhs)


def test_residual_vjp_zones_is_the_single_zone_transpose_folded(orc):
    """Without raccords the product is each zone's own fm_residual_vjp; with them, the sum of
    the gradient over the receiver and donor cells of a raccord is kept (the fold moves it)."""
    zones, racs = make("tiled")
    for z in zones:
        orc.init_metric(z)
    base = filled(zones, racs, [z.ro for z in zones])
    rng = np.random.default_rng(3)
    lams = [rng.standard_normal(5 * z.nd) for z in zones]
    plain = [orc.residual_vjp(z, l, b) for z, l, b in zip(zones, lams, base)]
    assert all(np.array_equal(a, b) for a, b in zip(orc.residual_vjp_zones(zones, [], lams, base), plain))
    folded = orc.residual_vjp_zones(zones, racs, lams, base)
    assert np.isclose(sum(a.sum() for a in folded), sum(a.sum() for a in plain), rtol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("split,extra", [("tiled", False), ("single", False), ("tiled", True)])
def test_residual_vjp_zones_gpu(orc, split, extra):
    """fm_residual_vjp_zones equals the oracle bit for bit: the per-zone products and the
    transfers' transpose (point-parallel, and point by point for the repeated-donor raccord)."""
    from fast_amd import _lib
    zo, racs = make(split, extra)
    zg, _ = make(split, extra)
    for z in zo:
        orc.init_metric(z)
    base = filled(zo, racs, [z.ro for z in zo])
    ctx = _lib.Context(0)
    ids = [ctx.add_zone(z) for z in zg]
    ctx.set_transfers(racs)
    for i, b in zip(ids, base):
        ctx.set(i, _lib.FIELD_RO, b)
    rng = np.random.default_rng(11)
    lams = [rng.standard_normal(5 * z.nd) for z in zo]
    got = ctx.residual_vjp_zones(lams)
    want = orc.residual_vjp_zones(zo, racs, lams, base)
    for z, a, b in zip(zo, got, want):
        assert np.array_equal(a, b), z.name
    ctx.close()


@pytest.mark.gpu
def test_flux_balance_zones_autograd(orc):
    """torch.autograd through flux_balance_zones: the forward is the oracle's balance after the
    transfers, the gradient of sum(lam * R) the oracle's transposed product with the fold."""
    import torch
    from fast_amd import _lib, adjoint
    zo, racs = make("tiled")
    zg, _ = make("tiled")
    for z in zo:
        orc.init_metric(z)
    ctx = _lib.Context(0)
    for z in zg:
        ctx.add_zone(z)
    ctx.set_transfers(racs)
    Ws = [torch.tensor(z.ro, dtype=torch.float64, requires_grad=True) for z in zo]
    Rs = adjoint.flux_balance_zones(Ws, ctx)
    base = filled(zo, racs, [z.ro for z in zo])
    for z, R, b in zip(zo, Rs, base):
        assert np.array_equal(z.interior(R.detach().numpy()), z.interior(orc.flux_balance(z, b))), z.name
    rng = np.random.default_rng(5)
    lams = [rng.standard_normal(5 * z.nd) for z in zo]
    J = sum((R * torch.from_numpy(l)).sum() for R, l in zip(Rs, lams))
    gs = torch.autograd.grad(J, Ws)
    want = orc.residual_vjp_zones(zo, racs, [l * interior_mask(z) for l, z in zip(lams, zo)], base)
    for z, g, w in zip(zo, gs, want):
        assert np.allclose(g.numpy(), w, rtol=1e-12, atol=1e-12), z.name
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("split,extra", [("tiled", False), ("tiled", True)])
def test_transfers_overwrite_in_raccord_order(split, extra):
    """fm_transfer fills the ghost cells as the raccords applied in order do: with a raccord
    that overwrites ghosts of an earlier one (and repeats its donors), the later value wins."""
    from fast_amd import _lib
    zo, racs = make(split, extra)
    zg, _ = make(split, extra)
    want = filled(zo, racs, [z.ro for z in zo])
    ctx = _lib.Context(0)
    ids = [ctx.add_zone(z) for z in zg]
    ctx.set_transfers(racs)
    ctx.transfer(0)
    for i, z, w in zip(ids, zg, want):
        got = np.zeros_like(z.ro)
        ctx.get(i, _lib.FIELD_RO, got)
        assert np.array_equal(got, w), z.name
    ctx.close()
