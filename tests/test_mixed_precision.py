"""Mixed precision (FM_OPT_LU_PRECISION = FM_LU_FP32): the LU-SGS sweep arithmetic in FP32,
the state, the Newton residual, the increments written back and the Newton update in FP64.

LU-SGS is the approximate factor of a defect correction: the Newton residual is formed in FP64
from the FP64 state, so the sub-iterations still drive it to the FP64 level; the FP32 sweeps only
perturb each increment at the FP32 rounding level."""
import numpy as np
import pytest

from fast_amd import cases, params as P
from tests import helpers as H

pytestmark = pytest.mark.gpu

MESHES = {
    "cart": lambda num: cases.tgv_zone((12, 10, 9), num, perturb=0.02, seed=5),
    "jittered": lambda num: H.jittered_box((10, 9, 8), num, bcs=H.WALL_J),
}


def ctx_for(z, fp32):
    from fast_amd import _lib
    ctx = _lib.Context(0)
    ctx.set_option(_lib.OPT_LU_PRECISION, _lib.LU_FP32 if fp32 else _lib.LU_FP64)
    zid = ctx.add_zone(z)
    ctx.apply_bc(zid, 0)
    t = np.zeros_like(z.ro)
    ctx.get(zid, _lib.FIELD_RO, t)
    ctx.set(zid, _lib.FIELD_RO_M1, t)
    ctx.set(zid, _lib.FIELD_RO_P1, t)
    return ctx, zid


@pytest.mark.parametrize("mesh", sorted(MESHES))
def test_fp32_sweeps_close_to_fp64(mesh):
    """One implicit sub-iteration: the new state differs from the FP64 sweeps' by at most 1e-5 of
    the increment (FP32 rounding through the wavefront sums), and is not identical to it."""
    from fast_amd import _lib
    num = H.numerics("ausmpred", "bdf")
    out = {}
    for fp32 in (False, True):
        z = MESHES[mesh](num)
        ctx, zid = ctx_for(z, fp32)
        w0 = np.zeros_like(z.ro); ctx.get(zid, _lib.FIELD_RO, w0)
        ctx.substep(zid, 1, first_it=0, stop=3)
        w1 = np.zeros_like(z.ro); ctx.get(zid, _lib.FIELD_RO_P1, w1)
        out[fp32] = (z.interior(w0), z.interior(w1))
        ctx.close()
    d64 = out[False][1] - out[False][0]
    diff = np.abs(out[True][1] - out[False][1])
    for e in range(5):
        scale = np.abs(d64[e]).max()
        assert diff[e].max() <= 1e-5 * scale, (e, diff[e].max(), scale)
    assert diff.max() > 0.


@pytest.mark.parametrize("mesh", sorted(MESHES))
def test_fp32_newton_converges_like_fp64(mesh):
    """Newton sub-iterations of one BDF step (first step, CFL about 0.5): the L2 norm of the Newton
    residual falls below 1e-10 of its first value with the FP32 sweeps as with the FP64 ones, in
    at most two more sub-iterations."""
    from fast_amd import _lib
    counts = {}
    for fp32 in (False, True):
        num = P.Numerics(scheme="ausmpred", temporal_scheme="implicit", ss_iteration=60, time_step=0.01)
        z = MESHES[mesh](num)
        z.param_real[P.DTC] = cases.tgv_dt(z, cfl=0.5)
        ctx, zid = ctx_for(z, fp32)
        first = None
        n = None
        for it in range(1, 61):
            ctx.substep(zid, it, first_it=0, lssiter_verif=1)
            r = ctx.residual_norms(zid)[:5]
            l2 = float(np.sqrt(np.sum(r ** 2)))
            first = l2 if first is None else first
            if l2 <= 1e-10 * first:
                n = it
                break
        ctx.close()
        assert n is not None, ("no convergence", mesh, fp32, l2 / first)
        counts[fp32] = n
    assert counts[True] <= counts[False] + 2, counts
