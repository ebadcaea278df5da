"""Generate the convergence-history fixtures from the reference FastS kernels.

Runs in the build container only, where oracle/_ref/libfastref.so (the reference Fortran
kernels compiled from their own sources by ``make -C oracle ref``) exists, one thread.  The
fixtures hold histories, not states:

tgv64_rk3_ke.npz
    The Taylor-Green vortex at 64^3 (laminar, explicit RK3, dt = 0.5 h / max(|u| + c) of the
    initial field, the reference's default cache blocking), for the Roe scheme with the minmod
    slope (``roe_min``) and with the unlimited kappa = 1/3 slope (``roe_kap``): the kinetic
    energy sum 1/2 rho |u|^2 vol and the volume after every time step (post.for), for 100 steps
    or until the state stops being finite.  ``ke_<scheme>`` has one row per step
    (tke, enstrophy, volume), ``sha_<scheme>`` the SHA-256 of the state (level N, ghost cells
    included) after each step; ``finite_steps_<scheme>`` counts the finite steps.
channel257_history.npz
    The channel with a bump, 257 x 129 x 33 (laminar AUSM+(P) o3, LU-SGS, local time step at
    CFL 10), ``nit`` implicit iterations from the free stream, each verified: ``s2[it, e]`` and
    ``mx[it, e]`` are the sum of squares and the maximum of |R| of equation e over the interior
    (the cprdu3s1.for block values reduced over the blocks), and ``sha_ro`` / ``sha_ro_m1`` the
    SHA-256 of the time levels N and N-1 (ghost cells included) after the last iteration.

Usage: python tests/golden/gen_histories.py [tgv|channel] [nit]
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fast_amd import params as P, cases  # noqa: E402
from tests import helpers as H  # noqa: E402
import oracle  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def tgv64(ref, scheme, nsteps=100):
    num = H.numerics(scheme, "explicit")
    z = cases.tgv_zone(64, num)
    z.param_real[P.DTC] = cases.tgv_dt(z, cfl=0.5)
    ref.init_metric(z)
    ref.apply_bc(z, z.ro)
    z.ro_m1[:] = z.ro
    z.ro_p1[:] = z.ro
    rows, hashes, finite = [], [], 0
    for step in range(nsteps):
        for it in range(1, 4):
            ref.substep(z, it, first_it=0 if step == 0 else 1)
        H.rotate(z)
        e = ref.enstrophy(z, z.ro)
        rows.append(e)
        hashes.append(sha(z.ro))
        if not np.isfinite(z.ro).all():
            break
        finite += 1
    return np.array(rows), np.array(hashes), finite, z.param_real[P.DTC]


def channel(ref, nit):
    z = cases.channel_zone(257, 129, 33)
    ref.init_metric(z)
    ref.apply_bc(z, z.ro)
    z.ro_m1[:] = z.ro
    z.ro_p1[:] = z.ro
    nss = int(z.param_int[P.NSSITER])
    s2 = np.zeros((nit, 5))
    mx = np.zeros((nit, 5))
    rdm = np.zeros(z.rdm_size())
    t0 = time.time()
    for it in range(nit):
        fi = 0 if it == 0 else 1
        # sub-iteration 1: RHS, LU-SGS, BCs; the last sub-iteration of a verified step computes
        # the residual of the new state only (lexit_lu = 1)
        ref.substep(z, 1, first_it=fi)
        rdm[:] = 0.
        ref.substep(z, nss, first_it=fi, lssiter_verif=1, lexit_lu=1, rdm=rdm)
        r = rdm.reshape(-1, 2, 5, nss)
        s2[it] = r[:, 0, :, 0].sum(axis=0)
        mx[it] = r[:, 1, :, 0].max(axis=0)
        H.rotate(z)
        if it % 100 == 0:
            print("channel it %d  L2(rho) %.6e  %.0fs" % (it, np.sqrt(s2[it, 0]), time.time() - t0), flush=True)
    return dict(s2=s2, mx=mx, nit=np.int32(nit), ncell=np.int64(np.prod(z.dims)),
                dtc=np.float64(z.param_real[P.DTC]), sha_ro=np.array(sha(z.ro)), sha_ro_m1=np.array(sha(z.ro_m1)))


def main():
    if not oracle.Reference.available():
        raise SystemExit("oracle/_ref/libfastref.so not built (make -C oracle ref)")
    ref = oracle.Reference()
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("all", "tgv"):
        out = {}
        for scheme in ("roe_min", "roe_kap"):
            rows, hashes, finite, dt = tgv64(ref, scheme)
            out["ke_" + scheme] = rows
            out["sha_" + scheme] = hashes
            out["finite_steps_" + scheme] = np.int32(finite)
            out["dt"] = np.float64(dt)
            print("tgv64 %s: %d finite steps" % (scheme, finite), flush=True)
        np.savez_compressed(os.path.join(HERE, "tgv64_rk3_ke.npz"), **out)
    if what in ("all", "channel"):
        nit = int(sys.argv[2]) if len(sys.argv) > 2 else 2000
        np.savez_compressed(os.path.join(HERE, "channel257_history.npz"), **channel(ref, nit))
        print("wrote channel257_history.npz", flush=True)


if __name__ == "__main__":
    main()
