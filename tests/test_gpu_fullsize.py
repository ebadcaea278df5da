"""The workloads at their own sizes, on the GPU against the oracle.

* the channel with a bump, 257 x 129 x 33 (3dfull, viscous walls, farfield, local time step at
  CFL 10): one implicit iteration bit for bit, ghost cells included;
* the refined channel 513 x 257 x 65: one implicit iteration bit for bit;
* the 64-zone set (4 x 4 x 4 abutting blocks, sizes from default_rng(1) in [16, 96], about 12 M
  cells) with the LU-SGS of all zones batched: one iteration bit for bit, every zone;
* the Taylor-Green vortex at 64^3 (Roe minmod, explicit RK3): all 100 time steps, the state bit
  for bit after every step and the kinetic energy sum of every step.

All of them run at the reference's default cache blocking (2048 / 2 / 7).  Each test prints its
case before the GPU work starts.
"""
import sys

import numpy as np
import pytest

from fast_amd import cases, connect, params as P
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from fast_amd import _lib
    return _lib


@pytest.fixture(scope="module")
def orc():
    import oracle
    return oracle.Oracle()


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def start_gpu(lib, zones, racs=(), options=()):
    ctx = lib.Context(0)
    for o, v in options:
        ctx.set_option(o, v)
    ids = [ctx.add_zone(z) for z in zones]
    if racs:
        ctx.set_transfers(racs)
        ctx.transfer(0)
    for i, z in zip(ids, zones):
        ctx.apply_bc(i, 0)
        t = np.zeros_like(z.ro)
        ctx.get(i, lib.FIELD_RO, t)
        ctx.set(i, lib.FIELD_RO_M1, t)
        ctx.set(i, lib.FIELD_RO_P1, t)
    return ctx, ids


def start_oracle(orc, zones, racs=()):
    for z in zones:
        orc.init_metric(z)
    if racs:
        connect.apply(zones, racs, [z.ro for z in zones])
    for z in zones:
        orc.apply_bc(z, z.ro)
        z.ro_m1[:] = z.ro
        z.ro_p1[:] = z.ro


def first_difference(a, b):
    d = np.flatnonzero(a != b)
    return "%d values differ, first at %d (%r vs %r)" % (d.size, d[0], a[d[0]], b[d[0]])


@pytest.mark.parametrize("dims", [(257, 129, 33), (513, 257, 65)], ids=["c1_257x129x33", "c3_513x257x65"])
def test_channel_one_iteration_full_size(lib, orc, dims):
    """One implicit iteration of the channel with a bump at full size (laminar AUSM+(P) o3,
    LU-SGS with the wall extension, local time step CFL 10): every level, ghost cells included,
    equals the oracle bit for bit."""
    say("channel %s: oracle" % (dims,))
    zo = cases.channel_zone(*dims)
    start_oracle(orc, [zo])
    orc.substep(zo, 1, first_it=0)
    say("channel %s: GPU" % (dims,))
    zg = cases.channel_zone(*dims)
    ctx, (zid,) = start_gpu(lib, [zg])
    ctx.substep(zid, 1, first_it=0)
    ctx.pull(zid, zg)
    ctx.close()
    assert zo.zone_type == P.ZONE_3DFULL
    for f in ("ro", "ro_m1", "ro_p1"):
        a, b = getattr(zo, f), getattr(zg, f)
        assert np.array_equal(a, b), f + ": " + first_difference(a, b)
    assert np.array_equal(zo.interior(zo.drodm), zg.interior(zg.drodm))


def test_zones64_one_iteration(lib, orc):
    """The 64-zone set (the bench's zones64 workload) for one implicit iteration with the LU-SGS
    of all zones batched and the transfers between them: every zone equals the oracle driven in
    gsdr3's order bit for bit."""
    from bench import zones64_case
    say("zones64: oracle")
    zo, racs, dims = zones64_case()
    start_oracle(orc, zo, racs)
    nsub = int(zo[0].param_int[P.NSSITER])
    ncell = sum(int(np.prod(z.dims)) for z in zo)
    assert len(zo) == 64 and ncell > 10_000_000
    # one step: sub-iteration 1 (RHS, LU-SGS, transfers, BCs); the last sub-iteration of a step
    # without verification is skipped (computePT.cpp:581-582)
    for it in range(1, nsub):
        for z in zo:
            orc.substep(z, it, first_it=0, stop=3)
        connect.apply(zo, racs, [z.ro_p1 for z in zo])
        for z in zo:
            orc.apply_bc(z, z.ro_p1)
    for z in zo:
        H.rotate(z)
    say("zones64: GPU (%d cells)" % ncell)
    zg, racs_g, _ = zones64_case()
    ctx, ids = start_gpu(lib, zg, racs_g, options=[(lib.OPT_ZONE_BATCH, 1)])
    ctx.compute(0, 1, nsub, nit_c=1, modulo_verif=0, first_it=0)
    ctx.sync()
    for i, z in zip(ids, zg):
        ctx.pull(i, z)
    ctx.close()
    for a, b in zip(zo, zg):
        for f in ("ro", "ro_m1"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), (a.name, f)


def test_c0_tgv64_100_steps_kinetic_energy(lib, orc):
    """The CPU-runnable case over its whole run: TGV 64^3, Roe with the minmod slope, laminar,
    explicit RK3 at dt = 0.5 h / max(|u| + c), 100 time steps.  After every step the GPU state
    equals the oracle's bit for bit (ghost cells included) and the kinetic energy sum
    1/2 rho |u|^2 vol (Fast/FastS/test/explicitLocalTgv_t1.py:86-90 defines the field) agrees to
    1e-13 (the two sums are reduced in different orders)."""
    num = H.numerics("roe_min", "explicit")

    def make():
        z = cases.tgv_zone(64, num)
        z.param_real[P.DTC] = cases.tgv_dt(z, cfl=0.5)
        return z
    zo, zg = make(), make()
    start_oracle(orc, [zo])
    ctx, (zid,) = start_gpu(lib, [zg])
    say("tgv64 RK3: 100 steps")
    ke_o, ke_g = [], []
    buf = np.zeros_like(zg.ro)
    for step in range(100):
        fi = 0 if step == 0 else 1
        for it in range(1, 4):
            orc.substep(zo, it, first_it=fi)
        H.rotate(zo)
        ctx.compute(step, 1, 3, nit_c=1, modulo_verif=0, first_it=fi)
        ctx.sync()
        e = orc.enstrophy(zo, zo.ro)
        g = ctx.tke(zid, 0)
        ke_o.append(e[0] / e[2])
        ke_g.append(g[0] / g[1])
        if step % 10 == 9 or step == 0:
            ctx.get(zid, lib.FIELD_RO, buf)
            assert np.array_equal(buf, zo.ro), "step %d: %s" % (step, first_difference(buf, zo.ro))
    ctx.pull(zid, zg)
    ctx.close()
    for f in ("ro", "ro_m1", "ro_p1"):
        assert np.array_equal(getattr(zo, f), getattr(zg, f)), f
    ke_o, ke_g = np.array(ke_o), np.array(ke_g)
    np.testing.assert_allclose(ke_g, ke_o, rtol=1e-13, atol=0)
    # the decaying vortex loses kinetic energy
    assert ke_o[-1] < ke_o[0]
