"""fm_wall_effort_noalloc, the wall-effort entry point that allocates nothing, against
fm_wall_effort: the same sums and per-face fields bit for bit, on two calls in a row (the second
on a smaller window, so the shared device array still holds the first call's values), and a
window above the array's size refused."""
import numpy as np
import pytest

from fast_amd import cases
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme", ["ausmpred", "roe_min", "senseur"])
def test_wall_effort_noalloc_equals_wall_effort(scheme):
    from fast_amd import _lib
    num = H.numerics(scheme, "bdf")
    z = cases.channel_zone(40, 24, 9, num)
    ctx = _lib.Context(0)
    zid = ctx.add_zone(z)
    ctx.apply_bc(zid, 0)
    ctx.substep(zid, 1, first_it=0)
    ctx.apply_bc(zid, 0)
    ni, nj, nk = z.dims
    pos = (0.3, -0.2, 0.1)
    windows = [(3, [1, ni, 1, 1, 1, nk]),            # the whole lower wall
               (4, [3, ni - 5, nj + 1, nj + 1, 2, nk - 1])]   # part of the upper wall: fewer faces
    for idir, w in windows:
        e_ref, f_ref = ctx.wall_effort(zid, idir, w, pos)
        e_na, f_na = ctx.wall_effort_noalloc(zid, idir, w, pos)
        assert np.array_equal(e_ref, e_na), (idir, e_ref - e_na)
        assert np.array_equal(f_ref, f_na), idir
        e_only = ctx.wall_effort_noalloc(zid, idir, w, pos, fields=False)
        assert np.array_equal(e_ref, e_only)
    ctx.close()


def test_wall_effort_noalloc_refuses_large_windows():
    from fast_amd import _lib
    num = H.numerics("ausmpred", "bdf")
    z = cases.tgv_zone((260, 260, 4), num)
    ctx = _lib.Context(0)
    zid = ctx.add_zone(z)
    with pytest.raises(_lib.FastMIError, match="at most 65536"):
        ctx.wall_effort_noalloc(zid, 6, [1, 260, 1, 260, 5, 5])
    e = ctx.wall_effort(zid, 6, [1, 260, 1, 260, 5, 5], fields=False)
    assert np.isfinite(e).all()
    ctx.close()
