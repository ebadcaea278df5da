"""Host-side zone: descriptor vectors plus the numpy fields of one structured block.

The field layout is the reference's (``FastC._compact``, ``Fast/FastC/FastC/PyTree.py:2236-2277``):
one contiguous float64 buffer per time level holding the five primitive variables
(Density, VelocityX/Y/Z, Temperature) one after the other, each ``NDIMDX`` long, i fastest,
ghost layers included.
"""
import numpy as np
from . import params as P
from .metric import classify_zone, metric_sizes


class HostZone:
    """Numpy arrays of one zone.  ``ro``, ``ro_m1``, ``ro_p1`` are the N, N-1 and N+1 levels."""

    def __init__(self, name, ni, nj, nk, numerics, state, x, y, z, bcs=(), flow="lamin"):
        self.name = name
        self.numerics = numerics
        self.state = state
        p = P.build_param_int(ni, nj, nk, numerics, flow=flow)
        r = P.build_param_real(numerics, state, flow=flow)
        p, r = P.append_bcs(p, r, list(bcs))
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.z = np.ascontiguousarray(z, dtype=np.float64)
        if self.x.size != int(p[P.NDIMDX_XYZ]):
            raise ValueError("coordinates must include the ghost vertices")
        classify_zone(p, self.x, self.y, self.z)
        self.param_int = p
        self.param_real = r
        self._allocate()

    @classmethod
    def from_arrays(cls, param_int, param_real, x, y, z, name="zone"):
        """Rebuild a zone from stored descriptor vectors and coordinates (fixtures, restarts)."""
        self = cls.__new__(cls)
        self.name = name
        self.numerics = None
        self.state = None
        self.param_int = np.ascontiguousarray(param_int, dtype=np.int32).copy()
        self.param_real = np.ascontiguousarray(param_real, dtype=np.float64).copy()
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.z = np.ascontiguousarray(z, dtype=np.float64)
        self._allocate()
        return self

    def _allocate(self):
        p = self.param_int
        nd = int(p[P.NDIMDX])
        self.nd = nd
        self.ro = np.zeros(5 * nd)
        self.ro_m1 = np.zeros(5 * nd)
        self.ro_p1 = np.zeros(5 * nd)
        self.xmut = np.zeros(nd)
        self.drodm = np.zeros(5 * nd)
        self.coe = np.zeros(int(p[P.NEQ_COE]) * nd)
        self.wig = np.zeros(3 * nd)
        sti, stj, stk, svol = metric_sizes(p)
        self.metric = np.zeros(sti + stj + stk + svol)
        self.ti = self.metric[:sti]
        self.tj = self.metric[sti:sti + stj]
        self.tk = self.metric[sti + stj:sti + stj + stk]
        self.vol = self.metric[sti + stj + stk:]

    @property
    def dims(self):
        p = self.param_int
        return int(p[P.IJKV]), int(p[P.IJKV + 1]), int(p[P.IJKV + 2])

    @property
    def zone_type(self):
        return int(self.param_int[P.ITYPZONE])

    def interior(self, a, nvar=5):
        """View of the interior cells of a (nvar*NDIMDX) field as [nvar, nk, nj, ni]."""
        p = self.param_int
        ni, nj, nk = int(p[P.NIJK]), int(p[P.NIJK + 1]), int(p[P.NIJK + 2])
        nf, nfk = int(p[P.NIJK + 3]), int(p[P.NIJK + 4])
        v = a[: nvar * ni * nj * nk].reshape(nvar, nk, nj, ni)
        return v[:, nfk:nk - nfk, nf:nj - nf, nf:ni - nf]

    def rdm_size(self):
        p = self.param_int
        n = 1
        for i in range(3):
            n *= max(1, int(p[P.IJKV + i]) // int(p[P.SIZE_SSDOM + i]))
        return int(p[P.NSSITER]) * 5 * 2 * n

    def copy_state_from(self, other):
        for a in ("ro", "ro_m1", "ro_p1", "xmut", "drodm", "coe", "wig", "metric"):
            getattr(self, a)[:] = getattr(other, a)
