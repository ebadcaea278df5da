"""Test oracle for the FastS structured NS iteration.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of bench.py, and only as the checker.  The shipped package never imports
or loads anything from here.

Two backends with the same call contract (``substep``, ``init_metric``, ``apply_bc``):

* :class:`Oracle` -- ``liboracle.so``, the C restatement in ``fastoracle.c`` (always built);
* :class:`Reference` -- ``_ref/libfastref.so``, the reference FastS Fortran kernels compiled
  from their own sources by ``make ref`` plus the harness ``ref_driver.c``.  Only present where
  the reference sources were available at build time.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFLIB = os.path.join(HERE, "_ref", "libfastref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(ref=None):
    """Build liboracle.so and, when the reference sources are present, _ref/libfastref.so."""
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
    if ref is None:
        ref = os.path.isdir("/root/reference/Fast")
    if ref:
        subprocess.check_call(["make", "-s", "-j8", "-C", HERE, "ref"])


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


class _Backend:
    prefix = None

    def __init__(self, path):
        self.path = path
        self.lib = ctypes.CDLL(path)

    def substep(self, z, nitcfg, nitrun=1, first_it=0, lssiter_verif=0, lexit_lu=0, rdm=None,
                stop=0):
        cfl = np.array([0., 1e8, 0.])     # computePT.cpp:521 initial max, min, sum
        f = getattr(self.lib, self.prefix + "substep")
        f.restype = ctypes.c_int
        rc = f(_i(z.param_int), _d(z.param_real), ctypes.c_int(nitcfg), ctypes.c_int(nitrun),
               ctypes.c_int(first_it), ctypes.c_int(lssiter_verif), ctypes.c_int(lexit_lu),
               _d(z.x), _d(z.y), _d(z.z), _d(z.ro), _d(z.ro_m1), _d(z.ro_p1), _d(z.xmut),
               _d(z.ti), _d(z.tj), _d(z.tk), _d(z.vol), _d(z.drodm), _d(z.coe), _d(z.wig),
               _d(rdm), _d(cfl), ctypes.c_int(stop))
        if rc != 0:
            raise RuntimeError("%s substep failed (%d)" % (self.prefix, rc))
        return cfl

    def enstrophy(self, z, rop, grad=None):
        """(tke sum, enstrophy sum, volume sum) of a state (POST/post.for); the velocity
        gradients go to ``grad`` (9 x NDIMDX) when given."""
        out = np.zeros(3)
        f = getattr(self.lib, self.prefix + "enstrophy")
        f(_i(z.param_int), _d(z.param_real), _d(rop), _d(z.ti), _d(z.tj), _d(z.tk), _d(z.vol), _d(out),
          _d(grad))
        return out

    def effort(self, z, idir, ind, pos=(0., 0., 0.), rop=None, fields=True):
        """Wall efforts of the face window ``ind`` (face index in the normal direction idir,
        cells across it) of the state ``rop`` (bceffort.for): the 12 sums and, when ``fields``,
        the 21 per-face fields (21 x nfaces, i fastest)."""
        rop = z.ro if rop is None else rop
        ind = np.ascontiguousarray(ind, dtype=np.int32)
        nf = int((ind[1] - ind[0] + 1) * (ind[3] - ind[2] + 1) * (ind[5] - ind[4] + 1))
        out = np.zeros(12)
        flu = np.zeros(21 * nf) if fields else None
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        f = getattr(self.lib, self.prefix + "effort")
        f.restype = ctypes.c_int
        rc = f(_i(z.param_int), _d(z.param_real), ctypes.c_int(idir), _i(ind), _d(pos), _d(rop),
               _d(z.xmut), _d(z.wig), _d(z.x), _d(z.y), _d(z.z), _d(z.ti), _d(z.tj), _d(z.tk),
               _d(z.vol), _d(out), _d(flu))
        if rc != 0:
            raise RuntimeError("%s effort failed (%d)" % (self.prefix, rc))
        return (out, flu.reshape(21, nf)) if fields else out

    def dpj_dpw(self, z, idir, ind, cosa, sina, surfinv, rop=None, dcd=None, dcl=None):
        """Gradients of the pressure drag and lift coefficients with respect to the
        conservative state of the wall cells of the slip-wall window ``ind`` (face index in the
        normal direction idir), assigned into ``dcd``, ``dcl`` (5 x NDIMDX, zero elsewhere when
        created here).  Oracle: the 3-D statement; Reference: the reference's 2-D routine run
        plane by plane (zero w, face vectors in the (x, y) plane required)."""
        rop = z.ro if rop is None else rop
        ind = np.ascontiguousarray(ind, dtype=np.int32)
        dcd = np.zeros(5 * z.nd) if dcd is None else dcd
        dcl = np.zeros(5 * z.nd) if dcl is None else dcl
        f = getattr(self.lib, self.prefix + "dpj_dpw")
        f.restype = ctypes.c_int
        c = [ctypes.c_double(v) for v in (cosa, sina, surfinv)]
        if self.prefix == "fo_":
            rc = f(_i(z.param_int), _d(z.param_real), ctypes.c_int(idir), _i(ind), _d(rop), _d(z.ti),
                   _d(z.tj), _d(z.tk), c[0], c[1], c[2], _d(dcd), _d(dcl))
        else:
            rc = f(_i(z.param_int), _d(z.param_real), ctypes.c_int(idir), _i(ind), _d(rop), _d(z.wig),
                   _d(z.x), _d(z.y), _d(z.z), _d(z.ti), _d(z.tj), _d(z.tk), _d(z.vol), c[0], c[1],
                   c[2], _d(dcd), _d(dcl))
        if rc != 0:
            raise RuntimeError("%s dpj_dpw failed (%d)" % (self.prefix, rc))
        return dcd, dcl

    def flux_balance(self, z, rop=None):
        """The AUSM/Roe/SENSEUR face balance of the interior cells alone (5 x NDIMDX)."""
        rop = z.ro if rop is None else rop
        out = np.zeros(5 * z.nd)
        f = getattr(self.lib, self.prefix + "flux_balance")
        f.restype = ctypes.c_int
        rc = f(_i(z.param_int), _d(z.param_real), _d(rop), _d(z.wig), _d(z.ti), _d(z.tj), _d(z.tk),
               _d(z.vol), _d(z.xmut), _d(out))
        if rc != 0:
            raise RuntimeError("%s flux_balance failed (%d)" % (self.prefix, rc))
        return out

    def residual_vjp(self, z, lam, rop=None):
        """(dR/dW)^T lam for the AUSM+(P) o3 Euler face balance R, W the primitive state of
        every cell (5 x NDIMDX)."""
        rop = z.ro if rop is None else rop
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        out = np.zeros(5 * z.nd)
        f = getattr(self.lib, self.prefix + "residual_vjp")
        f.restype = ctypes.c_int
        rc = f(_i(z.param_int), _d(z.param_real), _d(rop), _d(z.ti), _d(z.tj), _d(z.tk), _d(lam), _d(out))
        if rc != 0:
            raise RuntimeError("%s residual_vjp failed (%d)" % (self.prefix, rc))
        return out

    def residual_vjp_zones(self, zones, racs, lams, rops=None):
        """(dR/dW)^T lam of every zone, then the transpose of the injection transfers
        rcv[rcvPts] = dnr[donorPts] (FastC/Com/setInterpTransfersFast.cpp:440-760), raccords last
        to first: the gradient on a receiver cell moves onto its donor cell.  The reference has
        no such routine; this is the transpose of its transfer, point by point."""
        rops = [z.ro for z in zones] if rops is None else rops
        outs = [self.residual_vjp(z, lam, rop) for z, lam, rop in zip(zones, lams, rops)]
        for r in reversed(racs):
            ndr, ndd = zones[r.zr].nd, zones[r.zd].nd
            for q in range(len(r.rcv) - 1, -1, -1):
                for e in range(5):
                    x = outs[r.zr][r.rcv[q] + e * ndr]
                    outs[r.zr][r.rcv[q] + e * ndr] = 0.0
                    outs[r.zd][r.dnr[q] + e * ndd] += x
        return outs

    def apply_bc(self, z, rop, nstep=1):
        f = getattr(self.lib, self.prefix + "apply_bc")
        rc = f(_i(z.param_int), _d(z.param_real), ctypes.c_int(nstep), _d(rop), _d(z.ti),
               _d(z.tj), _d(z.tk), _d(z.x), _d(z.y), _d(z.z), _d(z.xmut))
        if rc != 0:
            raise RuntimeError("%s apply_bc failed (%d)" % (self.prefix, rc))


class Oracle(_Backend):
    """C restatement (liboracle.so)."""
    prefix = "fo_"

    def __init__(self, path=LIB):
        if not os.path.exists(path):
            build(ref=False)
        super().__init__(path)

    def init_metric(self, z):
        rc = self.lib.fo_init_metric(_i(z.param_int), _d(z.x), _d(z.y), _d(z.z), _d(z.ti),
                                     _d(z.tj), _d(z.tk), _d(z.vol))
        if rc != 0:
            raise RuntimeError("fo_init_metric failed")

    def lu_sgs(self, z, ind, rotmp, rop, drodm, mjrnewton=1):
        ind = np.ascontiguousarray(ind, dtype=np.int32)
        rc = self.lib.fo_lu_sgs(_i(z.param_int), _d(z.param_real), _i(ind), _d(rotmp), _d(rop),
                                _d(drodm), _d(z.ti), _d(z.tj), _d(z.tk), _d(z.coe),
                                ctypes.c_int(mjrnewton))
        if rc != 0:
            raise RuntimeError("fo_lu_sgs failed")


class Reference(_Backend):
    """Reference Fortran kernels (oracle/_ref/libfastref.so)."""
    prefix = "ref_"

    @staticmethod
    def available(path=REFLIB):
        return os.path.exists(path)

    def __init__(self, path=REFLIB):
        super().__init__(path)

    def classify(self, z):
        """Run nature_geom_dom on a copy of the zone's param_int; return it."""
        p = z.param_int.copy()
        degen = np.zeros(int(p[42]), dtype=np.int32)
        self.lib.ref_nature_geom_dom(_i(p), _d(z.param_real), _d(z.x), _d(z.y), _d(z.z), _i(degen))
        return p, degen

    def init_metric(self, z):
        degen = np.zeros(int(z.param_int[42]), dtype=np.int32)
        rc = self.lib.ref_init_metric(_i(z.param_int), _d(z.x), _d(z.y), _d(z.z), _i(degen),
                                      _d(z.ti), _d(z.tj), _d(z.tk), _d(z.vol))
        if rc != 0:
            raise RuntimeError("ref_init_metric failed")
