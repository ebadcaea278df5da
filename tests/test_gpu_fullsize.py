"""The workloads at their own sizes, on the GPU against the oracle.

* the channel with a bump, 257 x 129 x 33 (3dfull, viscous walls, farfield, local time step at
  CFL 10): one implicit iteration bit for bit, ghost cells included;
* the refined channel 513 x 257 x 65: one implicit iteration bit for bit;
* the 64-zone set (4 x 4 x 4 abutting blocks, sizes from default_rng(1) in [16, 96], about 12 M
  cells) with the LU-SGS of all zones batched: one iteration bit for bit, every zone;
* the Taylor-Green vortex at 64^3 (explicit RK3, Roe with the minmod and the kappa = 1/3
  slopes): all 100 time steps against the reference Fortran's own history, the state bit for bit
  after every step and the kinetic energy sum of every step;
* the channel 257 x 129 x 33 over 2000 iterations against the reference's residual history.

All of them run at the reference's default cache blocking (2048 / 2 / 7).  Each test prints its
case before the GPU work starts.
"""
import hashlib
import os
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


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


@pytest.mark.parametrize("scheme", ["roe_kap", "roe_min"])
def test_c0_tgv64_rk3_history_matches_reference(lib, scheme):
    """The CPU-runnable case over its whole run, against the reference Fortran's own history
    (tests/golden/tgv64_rk3_ke.npz, written by tests/golden/gen_histories.py from the compiled
    reference): TGV 64^3, laminar, explicit RK3 at dt = 0.5 h / max(|u| + c), 100 steps.  After
    every step the GPU state hashes to the reference's (bit for bit, ghost cells included) and
    the kinetic energy sum 1/2 rho |u|^2 vol agrees to the rounding of a sum of 64^3 positive
    terms (the reference sums under !$OMP SIMD, the GPU by blocks).  With the minmod slope the reference itself diverges (its state stops being
    finite at step 64); the GPU follows it step for step and diverges at the same step."""
    d = np.load(os.path.join(GOLDEN, "tgv64_rk3_ke.npz"), allow_pickle=False)
    ke_ref, sha_ref = d["ke_" + scheme], d["sha_" + scheme]
    finite = int(d["finite_steps_" + scheme])
    num = H.numerics(scheme, "explicit")
    z = cases.tgv_zone(64, num)
    z.param_real[P.DTC] = cases.tgv_dt(z, cfl=0.5)
    assert z.param_real[P.DTC] == float(d["dt"])
    ctx, (zid,) = start_gpu(lib, [z])
    say("tgv64 RK3 %s: %d steps" % (scheme, len(ke_ref)))
    buf = np.zeros_like(z.ro)
    for step in range(len(ke_ref)):
        ctx.compute(step, 1, 3, nit_c=1, modulo_verif=0, first_it=0 if step == 0 else 1)
        ctx.sync()
        ctx.get(zid, lib.FIELD_RO, buf)
        if step < finite:
            assert sha(buf) == str(sha_ref[step]), "state differs from the reference after step %d" % (step + 1)
            g = ctx.tke(zid, 0)
            # same cells, summed in another order: a sum of n positive terms carries at most
            # (n - 1) eps of relative rounding, 2.9e-11 for 64^3 cells
            np.testing.assert_allclose([g[0], g[1]], [ke_ref[step, 0], ke_ref[step, 2]], rtol=1e-11, atol=0)
        else:
            assert not np.isfinite(buf).all(), "the reference diverges at step %d" % (step + 1)
    ctx.close()
    if scheme == "roe_kap":
        assert finite == 100 and ke_ref[-1, 0] < ke_ref[0, 0]


def test_c1_channel_2000_iterations_history_matches_reference(lib):
    """The channel with a bump, 257 x 129 x 33, run 2000 implicit iterations from the free stream
    with every iteration verified, against the reference Fortran's own residual history
    (tests/golden/channel257_history.npz, written by tests/golden/gen_histories.py from the
    compiled reference, one thread): per iteration and equation the L-infinity norm is equal and
    the L2 norm agrees to 1e-12 (the sums of squares are reduced in another order), and after
    the last iteration both time levels hash to the reference's, bit for bit."""
    d = np.load(os.path.join(GOLDEN, "channel257_history.npz"), allow_pickle=False)
    nit, ncell, dtc = int(d["nit"]), int(d["ncell"]), float(d["dtc"])
    l2_ref = np.sqrt(d["s2"] / ncell) / dtc
    li_ref = d["mx"] / dtc
    z = cases.channel_zone(257, 129, 33)
    assert int(np.prod(z.dims)) == ncell and z.param_real[P.DTC] == dtc
    ctx, (zid,) = start_gpu(lib, [z])
    nsub = int(z.param_int[P.NSSITER])
    say("channel 257x129x33: %d verified iterations" % nit)
    l2 = np.zeros((nit, 5))
    li = np.zeros((nit, 5))
    for it in range(nit):
        ctx.compute(it, 1, nsub, nit_c=1, modulo_verif=1, first_it=0 if it == 0 else 1)
        n = ctx.residual_norms(zid)
        l2[it], li[it] = n[:5], n[5:]
        if it % 500 == 499:
            say("  iteration %d: L2(rho) %.6e (reference %.6e)" % (it + 1, l2[it, 0], l2_ref[it, 0]))
    for f, key in ((lib.FIELD_RO, "sha_ro"), (lib.FIELD_RO_M1, "sha_ro_m1")):
        buf = np.zeros_like(z.ro)
        ctx.get(zid, f, buf)
        assert sha(buf) == str(d[key]), key
    ctx.close()
    np.testing.assert_array_equal(li, li_ref)
    np.testing.assert_allclose(l2, l2_ref, rtol=1e-12, atol=0)
    assert np.isfinite(l2_ref).all()
