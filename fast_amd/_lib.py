"""ctypes binding of libfastmi.so (the C ABI declared in include/fastmi.h).

The library is required: there is no CPU fallback.  ``load()`` raises if the built library is
missing, and every call that returns a non-zero status raises :class:`FastMIError` with the
library's message.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libfastmi.so")

FIELD_RO, FIELD_RO_M1, FIELD_RO_P1, FIELD_MUT, FIELD_DRODM, FIELD_COE, FIELD_METRIC, FIELD_WIG = range(8)

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_int64)

# exported symbols (checked by the CPU test suite against include/fastmi.h)
SIGNATURES = {
    "fm_last_error": (ctypes.c_char_p, []),
    "fm_ctx_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    "fm_ctx_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fm_zone_add": (ctypes.c_int, [ctypes.c_void_p, _ip, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp,
                                   _dp, _ip]),
    "fm_zone_metric": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fm_zone_set_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long]),
    "fm_zone_get_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long]),
    "fm_zone_set_param_real": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int]),
    "fm_zone_substep": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 7),
    "fm_apply_bc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "fm_compute": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 6),
    "fm_compute_timed": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 6 + [ctypes.POINTER(ctypes.c_float)]),
    "fm_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "fm_residual_norms": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp]),
    "fm_enstrophy_tke": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]),
    "fm_time_phase": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_float)]),
    "fm_ctx_set_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "fm_cfl": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp]),
    "fm_residual_history": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int, _ip]),
    "fm_transfers_set": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip, _ip, _lp, _lp, _lp]),
    "fm_transfer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fm_wall_effort": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip,
                                      _dp, _dp, _dp]),
    "fm_wall_effort_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              _ip, _dp, _dp, _dp]),
    "fm_dpj_dpw": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]),
    "fm_flux_balance": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]),
    "fm_residual_vjp": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "fm_residual_vjp_zones": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
}
PHASE_COE, PHASE_RESIDUAL, PHASE_LU, PHASE_BC, PHASE_VJP = range(5)
OPT_LU_PRECISION = 5
OPT_RES_KCOL_N, OPT_LU_FLOW_GRID = 7, 8
LU_FP64, LU_FP32 = 0, 1
OPT_LU_SCHEDULE = 1
LU_BLOCKS, LU_PIPELINE, LU_DATAFLOW, LU_STREAM = 0, 1, 2, 3
OPT_RESIDUAL_SCHEDULE = 2
RES_CELLS, RES_FACES, RES_KCOL, RES_AUTO = 0, 1, 2, 3
OPT_ZONE_BATCH = 3
OPT_GRAPH = 4
OPT_LU_GIVE_UP = 6


class FastMIError(RuntimeError):
    pass


def load(path=LIBPATH):
    """Load libfastmi.so (built by fast_amd.build); raise if it is not there."""
    global _lib
    if _lib is not None and path == LIBPATH:
        return _lib
    if not os.path.exists(path):
        raise FastMIError("libfastmi.so not built (run __graft_entry__.build() or python -m fast_amd.build)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path == LIBPATH:
        _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = _lib.fm_last_error().decode() if _lib is not None else ""
        raise FastMIError("fastmi error %d: %s" % (rc, msg))


def dptr(a):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("float64 contiguous array required")
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a.dtype != np.int32 or not a.flags.c_contiguous:
        raise TypeError("int32 contiguous array required")
    return a.ctypes.data_as(_ip)


class Context:
    """A device context with its zones."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.fm_ctx_create(ctypes.byref(h), device))
        self.h = h
        self.zones = []

    def set_option(self, option, value):
        check(self.lib.fm_ctx_set_option(self.h, option, value))

    def close(self):
        if self.h:
            self.lib.fm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_zone(self, hz):
        """Register a HostZone, compute its metric on the device and upload its time levels."""
        zid = ctypes.c_int()
        p = np.ascontiguousarray(hz.param_int, dtype=np.int32)
        r = np.ascontiguousarray(hz.param_real, dtype=np.float64)
        check(self.lib.fm_zone_add(self.h, iptr(p), p.size, dptr(r), r.size, dptr(hz.x),
                                   dptr(hz.y), dptr(hz.z), ctypes.byref(zid)))
        check(self.lib.fm_zone_metric(self.h, zid.value))
        for f, a in ((FIELD_RO, hz.ro), (FIELD_RO_M1, hz.ro_m1), (FIELD_RO_P1, hz.ro_p1)):
            self.set(zid.value, f, a)
        self.zones.append(hz)
        return zid.value

    def zone_nd(self, zone):
        """NDIMDX of a zone added through add_zone."""
        return int(self.zones[zone].nd)

    def set(self, zone, field, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(self.lib.fm_zone_set_field(self.h, zone, field, dptr(a), a.size))

    def get(self, zone, field, out):
        check(self.lib.fm_zone_get_field(self.h, zone, field, dptr(out), out.size))
        return out

    def set_param_real(self, zone, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        check(self.lib.fm_zone_set_param_real(self.h, zone, dptr(r), r.size))

    def substep(self, zone, nitcfg, nitrun=1, first_it=0, lssiter_verif=0, lexit_lu=0, stop=0):
        check(self.lib.fm_zone_substep(self.h, zone, nitcfg, nitrun, first_it, lssiter_verif,
                                       lexit_lu, stop))

    def apply_bc(self, zone, level=0, nstep=1):
        check(self.lib.fm_apply_bc(self.h, zone, level, nstep))

    def set_transfers(self, raccords):
        """Injection raccords: objects with zr, zd, rcv, dnr (fast_amd.connect.Raccord)."""
        n = len(raccords)
        no_r = np.array([r.zr for r in raccords], dtype=np.int32)
        no_d = np.array([r.zd for r in raccords], dtype=np.int32)
        nb = np.array([len(r.rcv) for r in raccords], dtype=np.int64)
        rcv = np.ascontiguousarray(np.concatenate([r.rcv for r in raccords]) if n else np.zeros(1), dtype=np.int64)
        dnr = np.ascontiguousarray(np.concatenate([r.dnr for r in raccords]) if n else np.zeros(1), dtype=np.int64)
        keep = (no_r, no_d, nb, rcv, dnr)
        check(self.lib.fm_transfers_set(self.h, n, no_r.ctypes.data_as(_ip), no_d.ctypes.data_as(_ip),
                                        nb.ctypes.data_as(_lp), rcv.ctypes.data_as(_lp),
                                        dnr.ctypes.data_as(_lp)))
        del keep

    def wall_effort(self, zone, idir, window, pos=(0., 0., 0.), level=0, fields=True):
        """(12 sums, 21 x nfaces fields) of a face window (bceffort.for)."""
        w = np.ascontiguousarray(window, dtype=np.int32)
        nf = int((w[1] - w[0] + 1) * (w[3] - w[2] + 1) * (w[5] - w[4] + 1))
        eff = np.zeros(12)
        fl = np.zeros(21 * nf) if fields else None
        p = np.ascontiguousarray(pos, dtype=np.float64)
        check(self.lib.fm_wall_effort(self.h, zone, level, idir, w.ctypes.data_as(_ip),
                                      p.ctypes.data_as(_dp), eff.ctypes.data_as(_dp),
                                      fl.ctypes.data_as(_dp) if fields else None))
        return (eff, fl.reshape(21, nf)) if fields else eff

    def wall_effort_noalloc(self, zone, idir, window, pos=(0., 0., 0.), level=0, fields=True):
        """wall_effort through fm_wall_effort_noalloc (no device allocation; at most 65536
        faces per window)."""
        w = np.ascontiguousarray(window, dtype=np.int32)
        nf = int((w[1] - w[0] + 1) * (w[3] - w[2] + 1) * (w[5] - w[4] + 1))
        eff = np.zeros(12)
        fl = np.zeros(21 * nf) if fields else None
        p = np.ascontiguousarray(pos, dtype=np.float64)
        check(self.lib.fm_wall_effort_noalloc(self.h, zone, level, idir, w.ctypes.data_as(_ip),
                                              p.ctypes.data_as(_dp), eff.ctypes.data_as(_dp),
                                              fl.ctypes.data_as(_dp) if fields else None))
        return (eff, fl.reshape(21, nf)) if fields else eff

    def dpj_dpw(self, zone, idir, window, cosa, sina, surfinv, dcd, dcl, level=0):
        """Assign the gradients of the pressure drag / lift coefficients of a slip-wall window
        into ``dcd``, ``dcl`` (5 x NDIMDX float64, in place) (compute_dpJ_dpW.cpp)."""
        w = np.ascontiguousarray(window, dtype=np.int32)
        for a in (dcd, dcl):
            if a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError("dcd, dcl must be contiguous float64 arrays")
        check(self.lib.fm_dpj_dpw(self.h, zone, level, idir, w.ctypes.data_as(_ip), float(cosa),
                                  float(sina), float(surfinv), dptr(dcd), dptr(dcl)))

    def flux_balance(self, zone, level=0):
        """The face balance of the interior cells alone (5 x NDIMDX)."""
        nd = self.zone_nd(zone)
        out = np.zeros(5 * nd)
        check(self.lib.fm_flux_balance(self.h, zone, level, dptr(out)))
        return out

    def residual_vjp(self, zone, lam, level=0):
        """(dR/dW)^T lam for the AUSM+(P) o3 Euler face balance (5 x NDIMDX)."""
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        out = np.zeros_like(lam)
        if lam.size != 5 * self.zone_nd(zone):
            raise ValueError("lam must hold 5 x NDIMDX values")
        check(self.lib.fm_residual_vjp(self.h, zone, level, dptr(lam), dptr(out)))
        return out

    def residual_vjp_zones(self, lams, level=0):
        """(dR/dW)^T lam of every zone with the transfers' transpose: lams[n] and the returned
        arrays are 5 x NDIMDX per zone, in the order the zones were added."""
        n = len(self.zones)
        if len(lams) != n:
            raise ValueError("one lam array per zone (%d)" % n)
        lams = [np.ascontiguousarray(a, dtype=np.float64) for a in lams]
        for z, a in enumerate(lams):
            if a.size != 5 * self.zone_nd(z):
                raise ValueError("lam of zone %d must hold 5 x NDIMDX values" % z)
        outs = [np.zeros_like(a) for a in lams]
        pl = (ctypes.c_void_p * n)(*[a.ctypes.data for a in lams])
        po = (ctypes.c_void_p * n)(*[a.ctypes.data for a in outs])
        check(self.lib.fm_residual_vjp_zones(self.h, level, pl, po))
        return outs

    def transfer(self, level=0):
        check(self.lib.fm_transfer(self.h, level))

    def compute(self, nitrun, nstep_deb, nstep_fin, nit_c=1, modulo_verif=0, first_it=1):
        check(self.lib.fm_compute(self.h, nitrun, nstep_deb, nstep_fin, nit_c, modulo_verif, first_it))

    def compute_timed(self, nitrun, nstep_deb, nstep_fin, nit_c=1, modulo_verif=0, first_it=1):
        ms = ctypes.c_float()
        check(self.lib.fm_compute_timed(self.h, nitrun, nstep_deb, nstep_fin, nit_c, modulo_verif,
                                        first_it, ctypes.byref(ms)))
        return ms.value

    def sync(self):
        check(self.lib.fm_sync(self.h))

    def residual_norms(self, zone):
        out = np.zeros(10)
        check(self.lib.fm_residual_norms(self.h, zone, dptr(out)))
        return out

    def residual_history(self, zone, nmax=64):
        """(nrec, 10) array: L2 (5) and Linf (5) of every sub-iteration of the last verified
        step, divided by DTC."""
        out = np.zeros(10 * nmax)
        n = ctypes.c_int()
        check(self.lib.fm_residual_history(self.h, zone, out.ctypes.data_as(_dp), nmax, ctypes.byref(n)))
        return out[:10 * min(n.value, nmax)].reshape(-1, 10)

    def cfl(self, zone):
        out = np.zeros(3)
        check(self.lib.fm_cfl(self.h, zone, dptr(out)))
        return out

    def tke(self, zone, level=0):
        """(sum 1/2 rho|u|^2 vol, sum vol, sum 1/2 rho|curl u|^2 vol) of a time level."""
        out = np.zeros(3)
        check(self.lib.fm_enstrophy_tke(self.h, zone, level, dptr(out)))
        return out

    def time_phase(self, zone, phase, reps=10):
        ms = ctypes.c_float()
        check(self.lib.fm_time_phase(self.h, zone, phase, reps, ctypes.byref(ms)))
        return ms.value

    def pull(self, zone, hz):
        """Copy every device field of a zone into its HostZone arrays."""
        for f, a in ((FIELD_RO, hz.ro), (FIELD_RO_M1, hz.ro_m1), (FIELD_RO_P1, hz.ro_p1),
                     (FIELD_MUT, hz.xmut), (FIELD_DRODM, hz.drodm), (FIELD_COE, hz.coe),
                     (FIELD_METRIC, hz.metric), (FIELD_WIG, hz.wig)):
            self.get(zone, f, a)
        return hz
