// Kernels of the FastS structured NS iteration for MI355X (gfx950).  One thread per cell
// (or per ghost cell, per plane cell); i is the fastest, coalesced index everywhere.
#include "fm_device.hpp"
#include "fm_kernels.hpp"

namespace fm {

// ---------------------------------------------------------------------------------------
// Metric (Fast/FastC/FastC/Metric/cp_tijk.for, cp_vol.for), curvilinear zone.  Faces of cells
// -1..n+2 plus the imax+1 / jmax+1 / kmax+1 faces; one thread per (i,j,k) of the extended box.
__global__ void k_metric_faces(KZ z) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1 - z.nf;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1 - z.nf;
    int k = blockIdx.z + 1 - z.nfk;
    int imax = z.ni + z.nf, jmax = z.nj + z.nf, kmax = z.nk + z.nfk;
    if (i > imax + 1 || j > jmax + 1 || k > kmax + 1) return;
    const double* x = z.x; const double* y = z.y; const double* zz = z.z;
    long inci = 1, incj = z.xi, inck = (long)z.xi * z.xj, ndm = z.ndm;
    long lt = MTR(z, i, j, k), l0 = XYZ(z, i, j, k);
    long l1 = l0 + incj, l2 = l0 + inck, l3 = l1 + inck, l4 = l0 + inci, l5 = l1 + inci, l6 = l2 + inci;
    bool ini = i <= imax, inj = j <= jmax, ink = k <= kmax;
    if ((inj && ink)) {  // i faces (including the imax+1 face)
        double a10 = (y[l1] - y[l0]) * (zz[l3] - zz[l0]) - (zz[l1] - zz[l0]) * (y[l3] - y[l0]);
        double b10 = (zz[l1] - zz[l0]) * (x[l3] - x[l0]) - (x[l1] - x[l0]) * (zz[l3] - zz[l0]);
        double c10 = (x[l1] - x[l0]) * (y[l3] - y[l0]) - (y[l1] - y[l0]) * (x[l3] - x[l0]);
        double a11 = (y[l2] - y[l3]) * (zz[l0] - zz[l2]) - (zz[l2] - zz[l3]) * (y[l0] - y[l2]);
        double b11 = (zz[l2] - zz[l3]) * (x[l0] - x[l2]) - (x[l2] - x[l3]) * (zz[l0] - zz[l2]);
        double c11 = (x[l2] - x[l3]) * (y[l0] - y[l2]) - (y[l2] - y[l3]) * (x[l0] - x[l2]);
        z.ti[lt] = .5 * (a10 + a11); z.ti[lt + ndm] = .5 * (b10 + b11); z.ti[lt + 2 * ndm] = .5 * (c10 + c11);
    }
    if ((ini && ink)) {  // j faces
        double a10 = (y[l0] - y[l4]) * (zz[l2] - zz[l0]) - (zz[l0] - zz[l4]) * (y[l2] - y[l0]);
        double b10 = (zz[l0] - zz[l4]) * (x[l2] - x[l0]) - (x[l0] - x[l4]) * (zz[l2] - zz[l0]);
        double c10 = (x[l0] - x[l4]) * (y[l2] - y[l0]) - (y[l0] - y[l4]) * (x[l2] - x[l0]);
        double a11 = (y[l6] - y[l2]) * (zz[l4] - zz[l6]) - (zz[l6] - zz[l2]) * (y[l4] - y[l6]);
        double b11 = (zz[l6] - zz[l2]) * (x[l4] - x[l6]) - (x[l6] - x[l2]) * (zz[l4] - zz[l6]);
        double c11 = (x[l6] - x[l2]) * (y[l4] - y[l6]) - (y[l6] - y[l2]) * (x[l4] - x[l6]);
        z.tj[lt] = .5 * (a10 + a11); z.tj[lt + ndm] = .5 * (b10 + b11); z.tj[lt + 2 * ndm] = .5 * (c10 + c11);
    }
    if ((ini && inj)) {  // k faces
        double a10 = (y[l4] - y[l0]) * (zz[l5] - zz[l4]) - (zz[l4] - zz[l0]) * (y[l5] - y[l4]);
        double b10 = (zz[l4] - zz[l0]) * (x[l5] - x[l4]) - (x[l4] - x[l0]) * (zz[l5] - zz[l4]);
        double c10 = (x[l4] - x[l0]) * (y[l5] - y[l4]) - (y[l4] - y[l0]) * (x[l5] - x[l4]);
        double a11 = (y[l1] - y[l5]) * (zz[l0] - zz[l1]) - (zz[l1] - zz[l5]) * (y[l0] - y[l1]);
        double b11 = (zz[l1] - zz[l5]) * (x[l0] - x[l1]) - (x[l1] - x[l5]) * (zz[l0] - zz[l1]);
        double c11 = (x[l1] - x[l5]) * (y[l0] - y[l1]) - (y[l1] - y[l5]) * (x[l0] - x[l1]);
        z.tk[lt] = .5 * (a10 + a11); z.tk[lt + ndm] = .5 * (b10 + b11); z.tk[lt + 2 * ndm] = .5 * (c10 + c11);
    }
}

__global__ void k_metric_vol(KZ z) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1 - z.nf;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1 - z.nf;
    int k = blockIdx.z + 1 - z.nfk;
    if (i > z.ni + z.nf || j > z.nj + z.nf || k > z.nk + z.nfk) return;
    const double* x = z.x; const double* y = z.y; const double* zz = z.z;
    long inci = 1, incj = z.xi, inck = (long)z.xi * z.xj, ndm = z.ndm;
    long lt = MTR(z, i, j, k), l0 = XYZ(z, i, j, k);
    long l1 = l0 + incj, l2 = l0 + inck, l3 = l1 + inck, l4 = l0 + inci, l5 = l1 + inci;
    long l6 = l2 + inci, l7 = l3 + inci;
    long lip = lt + z.mi, ljp = lt + z.mj, lkp = lt + z.mk;
    const double* ti = z.ti; const double* tj = z.tj; const double* tk = z.tk;
    double ax = .25 * ((x[l4] + x[l5] + x[l7] + x[l6]) * ti[lip] - (x[l0] + x[l1] + x[l3] + x[l2]) * ti[lt]
                       + (x[l1] + x[l3] + x[l5] + x[l7]) * tj[ljp] - (x[l0] + x[l2] + x[l4] + x[l6]) * tj[lt]
                       + (x[l2] + x[l3] + x[l6] + x[l7]) * tk[lkp] - (x[l0] + x[l1] + x[l4] + x[l5]) * tk[lt]);
    double ay = .25 * ((y[l4] + y[l5] + y[l7] + y[l6]) * ti[lip + ndm] - (y[l0] + y[l1] + y[l3] + y[l2]) * ti[lt + ndm]
                       + (y[l1] + y[l3] + y[l5] + y[l7]) * tj[ljp + ndm] - (y[l0] + y[l2] + y[l4] + y[l6]) * tj[lt + ndm]
                       + (y[l2] + y[l3] + y[l6] + y[l7]) * tk[lkp + ndm] - (y[l0] + y[l1] + y[l4] + y[l5]) * tk[lt + ndm]);
    double az = .25 * ((zz[l4] + zz[l5] + zz[l7] + zz[l6]) * ti[lip + 2 * ndm] - (zz[l0] + zz[l1] + zz[l3] + zz[l2]) * ti[lt + 2 * ndm]
                       + (zz[l1] + zz[l3] + zz[l5] + zz[l7]) * tj[ljp + 2 * ndm] - (zz[l0] + zz[l2] + zz[l4] + zz[l6]) * tj[lt + 2 * ndm]
                       + (zz[l2] + zz[l3] + zz[l6] + zz[l7]) * tk[lkp + 2 * ndm] - (zz[l0] + zz[l1] + zz[l4] + zz[l5]) * tk[lt + 2 * ndm]);
    z.vol[lt] = dmax((ax + ay + az) / 3., 1.e-30);
}

// Cartesian zone: one scalar per direction (cp_tijk/cp_vol 3dcart branch at cell (1,1,1))
__global__ void k_metric_cart(KZ z) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double* x = z.x; const double* y = z.y; const double* zz = z.z;
    long incj = z.xi, inck = (long)z.xi * z.xj;
    long l0 = XYZ(z, 1, 1, 1);
    long l1 = l0 + incj, l2 = l0 + inck, l3 = l1 + inck, l4 = l0 + 1, l5 = l1 + 1, l6 = l2 + 1, l7 = l3 + 1;
    double tix10 = (y[l1] - y[l0]) * (zz[l3] - zz[l0]) - (zz[l1] - zz[l0]) * (y[l3] - y[l0]);
    double tix11 = (y[l2] - y[l3]) * (zz[l0] - zz[l2]) - (zz[l2] - zz[l3]) * (y[l0] - y[l2]);
    double tjy10 = (zz[l0] - zz[l4]) * (x[l2] - x[l0]) - (x[l0] - x[l4]) * (zz[l2] - zz[l0]);
    double tjy11 = (zz[l6] - zz[l2]) * (x[l4] - x[l6]) - (x[l6] - x[l2]) * (zz[l4] - zz[l6]);
    double tkz10 = (x[l4] - x[l0]) * (y[l5] - y[l4]) - (y[l4] - y[l0]) * (x[l5] - x[l4]);
    double tkz11 = (x[l1] - x[l5]) * (y[l0] - y[l1]) - (y[l1] - y[l5]) * (x[l0] - x[l1]);
    double ti = .5 * (tix10 + tix11), tj = .5 * (tjy10 + tjy11), tk = .5 * (tkz10 + tkz11);
    z.ti[0] = ti; z.tj[0] = tj; z.tk[0] = tk;
    double ax = .25 * ((x[l4] + x[l5] + x[l7] + x[l6]) * ti - (x[l0] + x[l1] + x[l3] + x[l2]) * ti);
    double ay = .25 * ((y[l1] + y[l3] + y[l5] + y[l7]) * tj - (y[l0] + y[l2] + y[l4] + y[l6]) * tj);
    double az = .25 * ((zz[l2] + zz[l3] + zz[l6] + zz[l7]) * tk - (zz[l0] + zz[l1] + zz[l4] + zz[l5]) * tk);
    z.vol[0] = dmax((ax + ay + az) / 3., 1.e-30);
}

// ---------------------------------------------------------------------------------------
// Viscosity and time-step / LU coefficients over a box (invist.for; tstb3_global.for;
// tstb3c.for; tstb3_impli.for + dtloc.for).  mode: 0 explicit (mu on box, coe1 = DTC/vol on
// the interior), 1 implicit global, 2 implicit local.
template <bool CART>
__global__ void k_coe(KZ z, const double* __restrict__ rop, int first_it, int mode, Box b, int lamin) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + b.i0;
    int j = blockIdx.y * blockDim.y + threadIdx.y + b.j0;
    int k = blockIdx.z + b.k0;
    if (i > b.i1 || j > b.j1 || k > b.k1) return;
    long nd = z.nd, l = IDX(z, i, j, k);
    double T = rop[l + 4 * nd];
    double mu = 0.;
    if (lamin) {
        double cmus1 = z.cs, temp01 = 1. / z.temp0;
        double coesut = z.xmul0 * (1. + cmus1 * temp01) * sqrt(temp01);
        mu = coesut * sqrt(T) * T / (T + cmus1);
        z.xmut[l] = mu;
    }
    if (mode == 0) {
        if (i >= 1 && i <= z.ni && j >= 1 && j <= z.nj && k >= 1 && k <= z.nk)
            z.coe[l] = CART ? z.dtc / z.vol[0] : z.dtc / z.vol[MTR(z, i, j, k)];
        return;
    }
    bool local = mode == 2;
    double cut0x = local ? 1e-30 : 1e-20;
    double gamma = z.gam, rgp = z.cv * (gamma - 1.);
    double xid = local ? 1. : (first_it == 0 ? 1. : 1.5);
    double gam1 = gamma / z.prandt;
    if (z.iflow == 1) gam1 = local ? cut0x : 0.;
    double gam1_1 = 1. / gam1, gam2 = gamma * rgp;
    double rr = rop[l], uu = rop[l + nd], vv = rop[l + 2 * nd], ww = rop[l + 3 * nd];
    double c = sqrt(gam2 * T);
    double c1, c2, c3, c4, detj, tc2i, tc2j, tc2k;
    double vmu = gam1 * mu;
    if (CART) {
        double tcxi = z.ti[0], tcyj = z.tj[0], tczk = z.tk[0];
        tc2i = sqrt(tcxi * tcxi); tc2j = sqrt(tcyj * tcyj); tc2k = sqrt(tczk * tczk);
        detj = dmax(z.vol[0], cut0x);
        double xmvis = vmu / (detj * rr);
        double vm;
        vm = fabs(tcxi * uu) + c * tc2i; c2 = vm * 2. + xmvis * tc2i * tc2i * 4.;
        vm = fabs(tcyj * vv) + c * tc2j; c3 = vm * 2. + xmvis * tc2j * tc2j * 4.;
        vm = fabs(tczk * ww) + c * tc2k; c4 = vm * 2. + xmvis * tc2k * tc2k * 4.;
    } else {
        long lt = MTR(z, i, j, k), ndm = z.ndm;
        long li = lt + z.mi, lj = lt + z.mj, lk = lt + z.mk;
        const double* ti = z.ti; const double* tj = z.tj; const double* tk = z.tk;
        detj = dmax(z.vol[lt], cut0x);
        double xmvis = vmu / (detj * rr);
        double tcx, tcy, tcz, qn, vm;
        tcx = .5 * (ti[lt] + ti[li]); tcy = .5 * (ti[lt + ndm] + ti[li + ndm]); tcz = .5 * (ti[lt + 2 * ndm] + ti[li + 2 * ndm]);
        tc2i = .5 * (sqrt(ti[lt] * ti[lt] + ti[lt + ndm] * ti[lt + ndm] + ti[lt + 2 * ndm] * ti[lt + 2 * ndm])
                   + sqrt(ti[li] * ti[li] + ti[li + ndm] * ti[li + ndm] + ti[li + 2 * ndm] * ti[li + 2 * ndm]));
        qn = tcx * uu + tcy * vv + tcz * ww; vm = fabs(qn) + c * tc2i;
        c2 = vm * 2. + xmvis * tc2i * tc2i * 4.;
        tcx = .5 * (tj[lt] + tj[lj]); tcy = .5 * (tj[lt + ndm] + tj[lj + ndm]); tcz = .5 * (tj[lt + 2 * ndm] + tj[lj + 2 * ndm]);
        tc2j = .5 * (sqrt(tj[lt] * tj[lt] + tj[lt + ndm] * tj[lt + ndm] + tj[lt + 2 * ndm] * tj[lt + 2 * ndm])
                   + sqrt(tj[lj] * tj[lj] + tj[lj + ndm] * tj[lj + ndm] + tj[lj + 2 * ndm] * tj[lj + 2 * ndm]));
        qn = tcx * uu + tcy * vv + tcz * ww; vm = fabs(qn) + c * tc2j;
        c3 = vm * 2. + xmvis * tc2j * tc2j * 4.;
        tcx = .5 * (tk[lt] + tk[lk]); tcy = .5 * (tk[lt + ndm] + tk[lk + ndm]); tcz = .5 * (tk[lt + 2 * ndm] + tk[lk + 2 * ndm]);
        tc2k = .5 * (sqrt(tk[lt] * tk[lt] + tk[lt + ndm] * tk[lt + ndm] + tk[lt + 2 * ndm] * tk[lt + 2 * ndm])
                   + sqrt(tk[lk] * tk[lk] + tk[lk + ndm] * tk[lk + ndm] + tk[lk + 2 * ndm] * tk[lk + 2 * ndm]));
        qn = tcx * uu + tcy * vv + tcz * ww; vm = fabs(qn) + c * tc2k;
        c4 = vm * 2. + xmvis * tc2k * tc2k * 4.;
    }
    if (local) {
        double d = 1. / (tc2i * tc2i + tc2j * tc2j + tc2k * tc2k);
        double dtvis = 0.5 * detj * d * gam1_1 / z.xmut[l] * rr;
        double sp = sqrt(uu * uu + vv * vv + ww * ww) + c;
        double dtconv = sqrt(d) / sp;
        c1 = z.cfl * dmin(dtconv, dtvis);
    } else {
        c1 = CART ? z.dtc / z.vol[0] : z.dtc / z.vol[MTR(z, i, j, k)];
    }
    z.coe[l] = c1; z.coe[l + nd] = c2; z.coe[l + 2 * nd] = c3; z.coe[l + 3 * nd] = c4;
    z.coe[l + 4 * nd] = xid + c1 * (c2 + c3 + c4) * 0.5;
}

// SENSEUR sensor (SENSOR/wiggle2001.for): interior cells plus the max faces
__device__ __forceinline__ void wiggle_one(const KZ& z, const double* __restrict__ rop, long l, long inc, int comp) {
    long nd = z.nd, nm = l - inc, nm2 = l - 2 * inc, np = l + inc;
    double f[5];
#pragma unroll
    for (int e = 0; e < 5; e++) {
        const double* q = rop + e * nd;
        double qm = (q[nm] - q[nm2]) * (q[l] - q[nm]);
        double qp = (q[l] - q[nm]) * (q[np] - q[l]);
        f[e] = dsign(1., qm) + dsign(1., qp);
    }
    double test = dmin(dmin(dmin(dmin(f[0], f[1]), f[2]), f[3]), f[4]);
    z.wig[l + comp * nd] = (test < -1e-12) ? 1. : 0.;
}
__global__ void k_wiggle(KZ z, const double* __restrict__ rop) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > z.ni + 1 || j > z.nj + 1 || k > z.nk + 1) return;
    long l = IDX(z, i, j, k);
    bool ii = i <= z.ni, jj = j <= z.nj, kk = k <= z.nk;
    if (ii && jj) wiggle_one(z, rop, l, z.inck, 2);
    if (ii && kk) wiggle_one(z, rop, l, z.incj, 1);
    if (jj && kk) wiggle_one(z, rop, l, 1, 0);
}

// Update after the face balance (acc): core3as2.for (implicit, writes the Newton residual
// into drodm) or core3ark3.for (explicit, keeps drodm and writes the new stage state).
__device__ __forceinline__ void res_update(const KZ& z, const RhsArgs& a, const double* __restrict__ rop,
                                           long l, const double acc[5]) {
    long nd = z.nd;
    double cf = z.coe[l];
    if (a.implicit) {
        const double* A = rop; const double* B = a.rop_n; const double* C = a.rop_n1;
        double cv = z.cv, c1 = a.c1, c2 = a.c2, c3 = a.c3;
        double al = A[l], cl = C[l];
        if (a.nitcfg == 1) {
            double c4 = c1 + c2;
            z.drodm[l] = cf * acc[0] + c4 * al + c3 * cl;
#pragma unroll
            for (int m = 1; m < 4; m++)
                z.drodm[l + m * nd] = cf * acc[m] + c4 * A[l + m * nd] * al + c3 * C[l + m * nd] * cl;
            double roe = al * (cv * A[l + 4 * nd] + 0.5 * (A[l + nd] * A[l + nd] + A[l + 2 * nd] * A[l + 2 * nd] + A[l + 3 * nd] * A[l + 3 * nd]));
            double roe_n1 = cl * (cv * C[l + 4 * nd] + 0.5 * (C[l + nd] * C[l + nd] + C[l + 2 * nd] * C[l + 2 * nd] + C[l + 3 * nd] * C[l + 3 * nd]));
            z.drodm[l + 4 * nd] = cf * acc[4] + c4 * roe + c3 * roe_n1;
        } else {
            double bl = B[l];
            z.drodm[l] = cf * acc[0] + c1 * al + c2 * bl + c3 * cl;
#pragma unroll
            for (int m = 1; m < 4; m++)
                z.drodm[l + m * nd] = cf * acc[m] + c1 * A[l + m * nd] * al + c2 * B[l + m * nd] * bl + c3 * C[l + m * nd] * cl;
            double roe = al * (cv * A[l + 4 * nd] + 0.5 * (A[l + nd] * A[l + nd] + A[l + 2 * nd] * A[l + 2 * nd] + A[l + 3 * nd] * A[l + 3 * nd]));
            double roe_n = bl * (cv * B[l + 4 * nd] + 0.5 * (B[l + nd] * B[l + nd] + B[l + 2 * nd] * B[l + 2 * nd] + B[l + 3 * nd] * B[l + 3 * nd]));
            double roe_n1 = cl * (cv * C[l + 4 * nd] + 0.5 * (C[l + nd] * C[l + nd] + C[l + 2 * nd] * C[l + 2 * nd] + C[l + 3 * nd] * C[l + 3 * nd]));
            z.drodm[l + 4 * nd] = cf * acc[4] + c1 * roe + c2 * roe_n + c3 * roe_n1;
        }
    } else {
#pragma unroll
        for (int e = 0; e < 5; e++) z.drodm[l + e * nd] = acc[e];
        double cv = z.cv, cvinv = 1. / cv, cvinv2 = 0.5 * cvinv;
        double r_1 = a.coefW * cf;
        double t1 = acc[0] * r_1, t2 = acc[1] * r_1, t3 = acc[2] * r_1, t4 = acc[3] * r_1, t5 = acc[4] * r_1;
        const double* R = a.rk_in;
        double* O = a.rk_out;
        double ro_old = R[l], u_old = R[l + nd], v_old = R[l + 2 * nd], w_old = R[l + 3 * nd], t_old = R[l + 4 * nd];
        double r1 = ro_old + t1;
        O[l] = r1;
        r_1 = 1. / r1;
        double u1 = (ro_old * u_old + t2) * r_1, v1 = (ro_old * v_old + t3) * r_1, w1 = (ro_old * w_old + t4) * r_1;
        O[l + nd] = u1; O[l + 2 * nd] = v1; O[l + 3 * nd] = w1;
        double roe_old = ro_old * (cv * t_old + 0.5 * (u_old * u_old + v_old * v_old + w_old * w_old));
        O[l + 4 * nd] = (roe_old + t5) * r_1 * cvinv - cvinv2 * (u1 * u1 + v1 * v1 + w1 * w1);
    }
}

// shape_tab_mtr.for: which metric components a face BC reads
__device__ __forceinline__ void shape_tab(int neq_mtr, int idir, int& ic, int& jc, int& kc,
                                          double& cim, double& cjm, double& ckm) {
    if (neq_mtr == 3) { ic = 1; jc = 2; kc = 3; cim = 1.; cjm = 1.; ckm = 1.; }
    else if (neq_mtr == 2) { ic = 1; jc = 2; kc = 2; cim = 1.; cjm = 1.; ckm = 0.; }
    else if (neq_mtr == 1) {
        ic = 1; jc = 1; kc = 1;
        cim = idir <= 2 ? 1. : 0.; cjm = (idir == 3 || idir == 4) ? 1. : 0.; ckm = idir >= 5 ? 1. : 0.;
    } else { ic = 1; jc = 1; kc = 1; cim = 0.; cjm = 0.; ckm = 0.; }
}

// ---------------------------------------------------------------------------------------
// correct_flux.for:44-260 for Roe (lapply = 4), applied to the balance of a cell next to a wall
// face, BC by BC: the Euler flux of the face is taken back out (template_correction.for,
// assemble_drodm_corrflu.for: df = flu*sens + flux(li)*c28 with c28 = 0), then the wall flux
// is put in (bflwall.for, bflwallslip.for slip walls; assemble_drodm_corr.for).
template <int SLOPE, bool CART>
__device__ void wall_correction(const FluxConst& f, const KZ& z, const double* __restrict__ rop,
                                const RhsArgs& a, int i, int j, int k, double acc[5]) {
    for (int n = 0; n < a.ncorr; n++) {
        const WallCorr& w = a.corr[n];
        const int d = (w.idir - 1) / 2;
        const bool even = (w.idir % 2) == 0;
        int c[3] = {i, j, k};
        const int fc = w.c[2 * d];
        if (c[d] != (even ? fc - 1 : fc)) continue;
        bool in = true;
        for (int t = 0; t < 3; t++)
            if (t != d) in = in && c[t] >= w.c[2 * t] && c[t] <= w.c[2 * t + 1];
        if (!in) continue;
        c[d] = fc;
        double flu[5];
        if (d == 0) face_flux<5, SLOPE, false, CART, 0>(f, z, rop, c[0], c[1], c[2], flu);
        else if (d == 1) face_flux<5, SLOPE, false, CART, 1>(f, z, rop, c[0], c[1], c[2], flu);
        else face_flux<5, SLOPE, false, CART, 2>(f, z, rop, c[0], c[1], c[2], flu);
        const double sens = even ? 1. : -1.;
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] = acc[e] + (flu[e] * sens + 0.);
        // wall flux from the state averaged across the face
        const long nd = z.nd, ndm = z.ndm;
        const long inc = d == 0 ? 1 : d == 1 ? z.incj : z.inck;
        const long l = IDX(z, c[0], c[1], c[2]), iadrf = l - inc;
        const long lt = MTR(z, c[0], c[1], c[2]);
        const double* tijk = d == 0 ? z.ti : d == 1 ? z.tj : z.tk;
        int ic, jc, kc;
        double cim, cjm, ckm;
        shape_tab(w.neq_mtr, w.idir, ic, jc, kc, cim, cjm, ckm);
        const double tcx = tijk[lt + ndm * (ic - 1)] * cim;
        const double tcy = tijk[lt + ndm * (jc - 1)] * cjm;
        const double tcz = tijk[lt + ndm * (kc - 1)] * ckm;
        const double qen = (0. * tcx + 0. * tcy + 0. * tcz * a.ck_vent) * a.c_ale;
        const double rgp = z.cv * (z.gam - 1.);
        const double r = 0.5 * (rop[l] + rop[iadrf]);
        const double u = 0.5 * (rop[l + nd] + rop[iadrf + nd]);
        const double v = 0.5 * (rop[l + 2 * nd] + rop[iadrf + 2 * nd]);
        const double ww = 0.5 * (rop[l + 3 * nd] + rop[iadrf + 3 * nd]);
        const double u_int = tcx * u + tcy * v + tcz * ww - qen;
        const double pw = 0.5 * (rop[l + 4 * nd] * rop[l] + rop[iadrf + 4 * nd] * rop[iadrf]) * rgp;
        if (!w.slip) {
            flu[0] = 0.;
            flu[1] = tcx * pw + u_int * u * r;
            flu[2] = tcy * pw + u_int * v * r;
            flu[3] = tcz * pw + u_int * ww * r;
            flu[4] = pw * qen;
        } else {
            const double gam2 = z.gam / (z.gam - 1.);
            const double pi = rop[iadrf + 4 * nd] * rop[iadrf] * rgp;
            const double h = gam2 * pi + r * (u * u + v * v + ww * ww) * 0.5;
            flu[0] = u_int * r;
            flu[1] = tcx * pw + u_int * r * u;
            flu[2] = tcy * pw + u_int * r * v;
            flu[3] = tcz * pw + u_int * r * ww;
            flu[4] = u_int * h + pw * qen;
        }
        const double sw = even ? -1. : 1.;
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] = acc[e] + flu[e] * sw;
    }
}

// ---------------------------------------------------------------------------------------
// Residual: fused init_rhs + 6 face fluxes + balance + update, one thread per interior cell.
// Accumulation order per cell as the reference sweep (template_FluxAndBalance.for:192-271):
//   init, +Fk(k), +Fj(j), +Fi(i), -Fi(i+1), -Fj(j+1), -Fk(k+1).
// The reference runs that sweep per cache block (navier_stokes_struct.for LOOP_CACHE: k blocks
// outermost, then j; one block in i) and the face between two blocks belongs to the later
// block, which also follows the Roe wall corrections of the earlier one (correct_flux runs per
// block).  So for a cell on the last row of a j block that is not the last (jn) or on the last
// plane of such a k block (kn), the max faces come in the order
//   [-Fj(j+1) unless jn] [-Fk(k+1) unless kn] [wall corrections] [-Fj(j+1) if jn] [-Fk(k+1) if kn]
// which is the single-block order when neither holds.
__device__ __forceinline__ bool cb_next(int c, int size, int nb) { return c % size == 0 && c / size < nb; }
// Update: core3as2.for (implicit, writes the Newton residual into drodm) or core3ark3.for
// (explicit, keeps drodm and writes the new stage state).
// FM_RES_WAVES: waves per SIMD the register allocation of k_residual aims at (2: no spills)
#ifndef FM_RES_WAVES
#define FM_RES_WAVES 2
#endif
template <int KFLU, int SLOPE, bool LAMIN, bool CART>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FM_RES_WAVES))) k_residual(KZ z, RhsArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > z.ni || j > z.nj || k > z.nk) return;
    const double* __restrict__ rop = a.rop;
    long nd = z.nd, l = IDX(z, i, j, k);
    FluxConst f = flux_const(z);
    double acc[5], flu[5];
    if (a.zero_init) {
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] = 0.;
    } else {
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] = z.drodm[l + e * nd] * a.coefH;
    }
    face_flux<KFLU, SLOPE, LAMIN, CART, 2>(f, z, rop, i, j, k, flu);
#pragma unroll
    for (int e = 0; e < 5; e++) acc[e] = acc[e] + flu[e];
    face_flux<KFLU, SLOPE, LAMIN, CART, 1>(f, z, rop, i, j, k, flu);
#pragma unroll
    for (int e = 0; e < 5; e++) acc[e] = acc[e] + flu[e];
    face_flux<KFLU, SLOPE, LAMIN, CART, 0>(f, z, rop, i, j, k, flu);
#pragma unroll
    for (int e = 0; e < 5; e++) acc[e] = acc[e] + flu[e];
    face_flux<KFLU, SLOPE, LAMIN, CART, 0>(f, z, rop, i + 1, j, k, flu);
#pragma unroll
    for (int e = 0; e < 5; e++) acc[e] = acc[e] - flu[e];
    const bool jn = cb_next(j, z.cbj, z.nbj), kn = cb_next(k, z.cbk, z.nbk);
    double fj1[5];
    face_flux<KFLU, SLOPE, LAMIN, CART, 1>(f, z, rop, i, j + 1, k, fj1);
    if (!jn) {
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] = acc[e] - fj1[e];
    }
    face_flux<KFLU, SLOPE, LAMIN, CART, 2>(f, z, rop, i, j, k + 1, flu);
    if (!kn) {
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] = acc[e] - flu[e];
    }
    if (KFLU == 5 && a.ncorr > 0) wall_correction<SLOPE, CART>(f, z, rop, a, i, j, k, acc);
    if (jn) {
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] = acc[e] - fj1[e];
    }
    if (kn) {
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] = acc[e] - flu[e];
    }

    res_update(z, a, rop, l, acc);
}

// Residual, a column of a.kcol_n cells up k per thread: the k face between two cells of the
// column is evaluated once (as the top face of the lower cell and the bottom face of the upper
// one), and the planes the k stencils share stay in the cache of the block.  Per cell the sum
// order is the reference's, so the result is the one of k_residual.
template <int KFLU, int SLOPE, bool LAMIN, bool CART>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FM_RES_WAVES))) k_residual_kcol(KZ z, RhsArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int kn_col = a.kcol_n > 0 ? a.kcol_n : RES_KCOL_N;
    int k0 = blockIdx.z * kn_col + 1;
    if (i > z.ni || j > z.nj || k0 > z.nk) return;
    const double* __restrict__ rop = a.rop;
    long nd = z.nd;
    FluxConst f = flux_const(z);
    double fk[5], flu[5], acc[5];
    face_flux<KFLU, SLOPE, LAMIN, CART, 2>(f, z, rop, i, j, k0, fk);
    const int k1 = k0 + kn_col - 1 < z.nk ? k0 + kn_col - 1 : z.nk;
    for (int k = k0; k <= k1; k++) {
        long l = IDX(z, i, j, k);
        if (a.zero_init) {
#pragma unroll
            for (int e = 0; e < 5; e++) acc[e] = 0.;
        } else {
#pragma unroll
            for (int e = 0; e < 5; e++) acc[e] = z.drodm[l + e * nd] * a.coefH;
        }
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] = acc[e] + fk[e];
        face_flux<KFLU, SLOPE, LAMIN, CART, 1>(f, z, rop, i, j, k, flu);
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] = acc[e] + flu[e];
        face_flux<KFLU, SLOPE, LAMIN, CART, 0>(f, z, rop, i, j, k, flu);
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] = acc[e] + flu[e];
        face_flux<KFLU, SLOPE, LAMIN, CART, 0>(f, z, rop, i + 1, j, k, flu);
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] = acc[e] - flu[e];
        const bool jn = cb_next(j, z.cbj, z.nbj), kn = cb_next(k, z.cbk, z.nbk);
        face_flux<KFLU, SLOPE, LAMIN, CART, 1>(f, z, rop, i, j + 1, k, flu);
        if (!jn) {
#pragma unroll
            for (int e = 0; e < 5; e++) acc[e] = acc[e] - flu[e];
        }
        face_flux<KFLU, SLOPE, LAMIN, CART, 2>(f, z, rop, i, j, k + 1, fk);
        if (!kn) {
#pragma unroll
            for (int e = 0; e < 5; e++) acc[e] = acc[e] - fk[e];
        }
        if (KFLU == 5 && a.ncorr > 0) wall_correction<SLOPE, CART>(f, z, rop, a, i, j, k, acc);
        if (jn) {
#pragma unroll
            for (int e = 0; e < 5; e++) acc[e] = acc[e] - flu[e];
        }
        if (kn) {
#pragma unroll
            for (int e = 0; e < 5; e++) acc[e] = acc[e] - fk[e];
        }
        res_update(z, a, rop, l, acc);
    }
}

// Residual with every face flux evaluated once.  Block: 64 i-lanes x 5 j-rows of threads
// marching RES_KC planes in k.  Lane tx of rows 0..3 evaluates the i-face at i0 + tx (its cell
// i0 + tx takes face i + 1 from lane tx + 1, so a block covers 63 cells); row r evaluates the
// j-face at j0 + r (row 4 only provides the face above the last cell row); each cell carries
// the k-face under it from the previous plane.  Per cell the faces are summed in the reference
// order (template_FluxAndBalance.for:192-271), so the result is the one of k_residual.
constexpr int RES_KC = 32;

template <int KFLU, int SLOPE, bool LAMIN, bool CART>
__global__ void __launch_bounds__(320) k_residual_faces(KZ z, RhsArgs a) {
    __shared__ double fj[2][5][5][64];        // [plane parity][row][eq][lane]
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = blockIdx.x * 63 + tx + 1;
    const int j = blockIdx.y * 4 + ty + 1;
    const int k0 = blockIdx.z * RES_KC + 1;
    const int k1 = k0 + RES_KC - 1 < z.nk ? k0 + RES_KC - 1 : z.nk;
    const bool cell = tx < 63 && ty < 4 && i <= z.ni && j <= z.nj;
    const bool iface = ty < 4 && j <= z.nj && i <= z.ni + 1;
    const bool jface = tx < 63 && i <= z.ni && j <= z.nj + 1;
    const double* __restrict__ rop = a.rop;
    const long nd = z.nd;
    FluxConst f = flux_const(z);
    double fk[5] = {0., 0., 0., 0., 0.};
    if (cell) face_flux<KFLU, SLOPE, LAMIN, CART, 2>(f, z, rop, i, j, k0, fk);
    for (int k = k0; k <= k1; k++) {
        const int par = k & 1;
        double fl[5];
        if (jface) {
            face_flux<KFLU, SLOPE, LAMIN, CART, 1>(f, z, rop, i, j, k, fl);
#pragma unroll
            for (int e = 0; e < 5; e++) fj[par][ty][e][tx] = fl[e];
        }
        double fi[5] = {0., 0., 0., 0., 0.};
        if (iface) face_flux<KFLU, SLOPE, LAMIN, CART, 0>(f, z, rop, i, j, k, fi);
        double fi1[5];
#pragma unroll
        for (int e = 0; e < 5; e++) fi1[e] = __shfl_down(fi[e], 1, 64);
        double fk1[5];
        if (cell) face_flux<KFLU, SLOPE, LAMIN, CART, 2>(f, z, rop, i, j, k + 1, fk1);
        __syncthreads();
        if (cell) {
            const long l = IDX(z, i, j, k);
            double acc[5];
            if (a.zero_init) {
#pragma unroll
                for (int e = 0; e < 5; e++) acc[e] = 0.;
            } else {
#pragma unroll
                for (int e = 0; e < 5; e++) acc[e] = z.drodm[l + e * nd] * a.coefH;
            }
#pragma unroll
            for (int e = 0; e < 5; e++) acc[e] = acc[e] + fk[e];
#pragma unroll
            for (int e = 0; e < 5; e++) acc[e] = acc[e] + fj[par][ty][e][tx];
#pragma unroll
            for (int e = 0; e < 5; e++) acc[e] = acc[e] + fi[e];
#pragma unroll
            for (int e = 0; e < 5; e++) acc[e] = acc[e] - fi1[e];
            const bool jn = cb_next(j, z.cbj, z.nbj), kn = cb_next(k, z.cbk, z.nbk);
            if (!jn) {
#pragma unroll
                for (int e = 0; e < 5; e++) acc[e] = acc[e] - fj[par][ty + 1][e][tx];
            }
            if (!kn) {
#pragma unroll
                for (int e = 0; e < 5; e++) acc[e] = acc[e] - fk1[e];
            }
            if (jn) {
#pragma unroll
                for (int e = 0; e < 5; e++) acc[e] = acc[e] - fj[par][ty + 1][e][tx];
            }
            if (kn) {
#pragma unroll
                for (int e = 0; e < 5; e++) acc[e] = acc[e] - fk1[e];
            }
            res_update(z, a, rop, l, acc);
#pragma unroll
            for (int e = 0; e < 5; e++) fk[e] = fk1[e];
        }
    }
}

// ---------------------------------------------------------------------------------------
// Ghost cells.  Face BC over a window (BCExtrapolate.for, BCPeriodic.for,
// BCWallViscous_{i,j,k}.for), then correct_coins.for edges and corners.
// metric index of the boundary face of the ghost column through (c0, c1, c2)
__device__ __forceinline__ long bc_face_mtr(const KZ& z, const BcArgs& a, int c0, int c1, int c2) {
    int d = (a.idir - 1) / 2;
    bool up = (a.idir % 2) == 0;
    int c[3] = {c0, c1, c2};
    c[d] = up ? a.w[2 * d] : a.w[2 * d + 1] + 1;
    return MTR(z, c[0], c[1], c[2]);
}

__global__ void k_bc_face(KZ z, double* __restrict__ rop, BcArgs a) {
    int n0 = a.w[1] - a.w[0] + 1, n1 = a.w[3] - a.w[2] + 1;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int k = a.w[4] + blockIdx.y;
    if (t >= n0 * n1 || k > a.w[5]) return;
    int i = a.w[0] + t % n0, j = a.w[2] + t / n0;
    long nd = z.nd, l = IDX(z, i, j, k);
    int d = (a.idir - 1) / 2;
    bool up = (a.idir % 2) == 0;
    int c[3] = {i, j, k};
    int n = d == 0 ? z.ni : d == 1 ? z.nj : z.nk;
    if (a.kind == 0) c[d] = up ? a.w[2 * d] - 1 : a.w[2 * d + 1] + 1;
    else if (a.kind == 2) c[d] = up ? c[d] - n : c[d] + n;
    else c[d] = up ? 2 * a.w[2 * d] - 1 - c[d] : 2 * a.w[2 * d + 1] + 1 - c[d];
    long l1 = IDX(z, c[0], c[1], c[2]);
    double c1 = a.lrhs ? 0. : 1.;
    if (a.kind == 0) {
#pragma unroll
        for (int e = 0; e < 5; e++) rop[l + e * nd] = dmax(-1.e30, rop[l1 + e * nd]) * c1;
    } else if (a.kind == 2) {
#pragma unroll
        for (int e = 0; e < 5; e++) rop[l + e * nd] = rop[l1 + e * nd] * c1;
    } else if (a.kind == 1) {
        double u = rop[l1 + nd], v = rop[l1 + 2 * nd], w = rop[l1 + 3 * nd];
        rop[l] = rop[l1];
        rop[l + nd] = 0. - u; rop[l + 2 * nd] = 0. - v; rop[l + 3 * nd] = 0. - w;
        rop[l + 4 * nd] = rop[l1 + 4 * nd];
    } else {
        // BCWallInviscid.for: mirror state with the normal velocity reversed (no mesh motion,
        // so the entrainment terms are +0)
        int ic, jc, kc; double cim, cjm, ckm;
        shape_tab(a.neq_mtr, a.idir, ic, jc, kc, cim, cjm, ckm);
        long lm = bc_face_mtr(z, a, i, j, k), ndm = z.ndm;
        double tcx = a.tijk[lm + (ic - 1) * ndm] * cim, tcy = a.tijk[lm + (jc - 1) * ndm] * cjm;
        double tcz = a.tijk[lm + (kc - 1) * ndm] * ckm;
        double surf = dmax(sqrt(tcx * tcx + tcy * tcy + tcz * tcz), 1e-30);
        double s_1 = 1. / surf;
        tcx = tcx * s_1; tcy = tcy * s_1; tcz = tcz * s_1;
        double u = rop[l1 + nd], v = rop[l1 + 2 * nd], w = rop[l1 + 3 * nd];
        double qn = (u * tcx + v * tcy + w * tcz);
        double ut = u - qn * tcx, vt = v - qn * tcy, wt = w - qn * tcz;
        double ua = 2. * ut - u, va = 2. * vt - v, wa = 2. * wt - w;
        rop[l] = rop[l1];
        rop[l + nd] = 0. + ua; rop[l + 2 * nd] = 0. + va; rop[l + 3 * nd] = 0. + wa;
        rop[l + 4 * nd] = rop[l1 + 4 * nd];
    }
}

// Characteristic ghost state (FastS BC/non_reflection.for): exterior state (roe..ete),
// interior cell l1, boundary normal (tcx, tcy, tcz) pointing out of the zone.
struct NonRefl { double svar1, svar2, svar3, svar4, qvar5, r, u, v, w, qn, sn; };

__device__ __forceinline__ NonRefl non_reflection(const double* __restrict__ rop, long nd, long l1,
                                                  double roe, double rue, double rve, double rwe,
                                                  double ete, double tcx, double tcy, double tcz,
                                                  double gam, double gam1, double gam1_1, double cvinf) {
    NonRefl o;
    double sn = sqrt(tcx * tcx + tcy * tcy + tcz * tcz);
    double roi = rop[l1];
    double rui = rop[l1 + nd] * roi, rvi = rop[l1 + 2 * nd] * roi, rwi = rop[l1 + 3 * nd] * roi;
    double eti = rop[l1 + 4 * nd] * cvinf * roi + .5 * (rui * rui + rvi * rvi + rwi * rwi) / roi;
    double qref1 = .5 * (roe + roi), qref2 = .5 * (rue + rui), qref3 = .5 * (rve + rvi);
    double qref4 = .5 * (rwe + rwi), qref5 = .5 * (ete + eti);
    double r = qref1, ri = 1. / r;
    double u = qref2 * ri, v = qref3 * ri, w = qref4 * ri;
    double p = gam1 * (qref5 - .5 * r * (u * u + v * v + w * w));
    double c = sqrt(gam * p * ri), ci = 1. / c, s_1 = 1. / sn;
    double qn = tcx * u + tcy * v + tcz * w - 0.;
    double tnx, tny;
    if (sn == 0.) { tnx = 1.; tny = 0.; } else { tnx = tcx / sn; tny = tcy / sn; }
    double eig4 = qn + sn * c, eig5 = qn - sn * c;
    double q[2][5] = {{roe, rue, rve, rwe, ete}, {roi, rui, rvi, rwi, eti}};
    double ch[2][5];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        double q1 = q[s][0], q2 = q[s][1], q3 = q[s][2], q4 = q[s][3], q5 = q[s][4];
        double s1 = q1, s2 = (q2 - u * q1) * ri, s3 = (q3 - v * q1) * ri, s4 = (q4 - w * q1) * ri;
        double s5 = gam1 * (.5 * (u * u + v * v + w * w) * q1 - u * q2 - v * q3 - w * q4 + q5);
        ch[s][0] = s1 - ci * ci * s5;
        ch[s][1] = tcz * s_1 * (tnx * s2 + tny * s3) - s4;
        ch[s][2] = -tny * s2 + tnx * s3;
        ch[s][3] = .5 * s_1 * (tcx * s2 + tcy * s3 + tcz * s4) + .5 * ri * ci * s5;
        ch[s][4] = -.5 * s_1 * (tcx * s2 + tcy * s3 + tcz * s4) + .5 * ri * ci * s5;
    }
    // characteristic variables from outside where the eigenvalue is negative
    double rv1 = qn < 0. ? ch[0][0] : ch[1][0], rv2 = qn < 0. ? ch[0][1] : ch[1][1];
    double rv3 = qn < 0. ? ch[0][2] : ch[1][2];
    double rv4 = eig4 < 0. ? ch[0][3] : ch[1][3], rv5 = eig5 < 0. ? ch[0][4] : ch[1][4];
    o.svar1 = rv1 + r * ci * (rv4 + rv5);
    o.svar2 = tcz * s_1 * tnx * rv2 - tny * rv3 + tcx * s_1 * (rv4 - rv5);
    o.svar3 = tcz * s_1 * tny * rv2 + tnx * rv3 + tcy * s_1 * (rv4 - rv5);
    o.svar4 = -rv2 + tcz * s_1 * (rv4 - rv5);
    double svar5 = r * c * (rv4 + rv5);
    o.qvar5 = .5 * (u * u + v * v + w * w) * o.svar1 + r * (u * o.svar2 + v * o.svar3 + w * o.svar4)
              + gam1_1 * svar5;
    o.r = r; o.u = u; o.v = v; o.w = w; o.qn = qn; o.sn = sn;
    return o;
}

// Face BCs computed one ghost column at a time: the first rank from the adjacent interior
// cell (bvbs_farfield / bvbs_inflow: BCFarfield.for; bvbs_outflow: BCOutflow.for;
// bvbs_outpres: BCOutpres + linearization_state.for), then the outer ranks in the reference
// loop order (BC_nextrank.for for farfield / inflow, BC_extrap.for otherwise).
__global__ void k_bc_column(KZ z, double* __restrict__ rop, BcArgs a) {
    const int d = (a.idir - 1) / 2;
    const bool up = (a.idir % 2) == 0;
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    const int na = a.w[2 * ta + 1] - a.w[2 * ta] + 1, nb = a.w[2 * tb + 1] - a.w[2 * tb] + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na * nb) return;
    const int ca = a.w[2 * ta] + t % na, cb = a.w[2 * tb] + t / na;
    const long nd = z.nd, ndm = z.ndm;
    int ic, jc, kc; double cim, cjm, ckm;
    shape_tab(a.neq_mtr, a.idir, ic, jc, kc, cim, cjm, ckm);
    const double snorm = up ? 1. : -1.;
    const double gam = z.gam, gam1 = z.gam - 1., gam1_1 = 1. / gam1;
    const double rgp = z.cv * gam1, cvinv = 1. / z.cv;
    const double c4 = 5. / 6., c5 = 2. / 6., c6 = -1. / 6.;
    const double c0 = 1. / c6, c1 = -(c4 + c5) * c0, c2 = -c6 * c0, c3 = (2. - c5 - c4) * c0;
    const long inc = d == 0 ? 1 : d == 1 ? z.incj : z.inck;
    const long sinc = up ? -inc : inc;
    const int g0 = up ? a.w[2 * d] : a.w[2 * d + 1];
    int c[3];
    c[ta] = ca; c[tb] = cb; c[d] = g0;
    const long lm = bc_face_mtr(z, a, c[0], c[1], c[2]);
    const double tcx = a.tijk[lm + (ic - 1) * ndm] * cim * snorm;
    const double tcy = a.tijk[lm + (jc - 1) * ndm] * cjm * snorm;
    const double tcz = a.tijk[lm + (kc - 1) * ndm] * ckm * snorm;
    const long l = IDX(z, c[0], c[1], c[2]);
    const long l1 = l + sinc;
    if (a.kind == 4) {
        NonRefl o = non_reflection(rop, nd, l1, a.state[0], a.state[1], a.state[2], a.state[3],
                                   a.state[4], tcx, tcy, tcz, gam, gam1, gam1_1, z.cv);
        double roinv = 1. / o.svar1;
        double u = o.u + o.r * o.svar2 * roinv, v = o.v + o.r * o.svar3 * roinv, w = o.w + o.r * o.svar4 * roinv;
        rop[l] = o.svar1; rop[l + nd] = u; rop[l + 2 * nd] = v; rop[l + 3 * nd] = w;
        rop[l + 4 * nd] = (o.qvar5 * roinv - .5 * (u * u + v * v + w * w)) * cvinv;
    } else if (a.kind == 5) {
        double roe = rop[l1], rue = rop[l1 + nd] * roe, rve = rop[l1 + 2 * nd] * roe, rwe = rop[l1 + 3 * nd] * roe;
        double roe_inv = 1. / a.state[0];
        double ete = a.state[4] - 0.5 * (a.state[1] * a.state[1] + a.state[2] * a.state[2] + a.state[3] * a.state[3]) * roe_inv
                     + 0.5 * (rue * rop[l1 + nd] + rve * rop[l1 + 2 * nd] + rwe * rop[l1 + 3 * nd]);
        NonRefl o = non_reflection(rop, nd, l1, roe, rue, rve, rwe, ete, tcx, tcy, tcz, gam, gam1, gam1_1, z.cv);
        rop[l] = rop[l1];
        double roinv = 1. / rop[l];
        double u2, u3, u4;
        if (o.qn < 0) {
            u2 = rop[l1 + nd] - o.qn / (o.sn * o.sn) * tcx;
            u3 = rop[l1 + 2 * nd] - o.qn / (o.sn * o.sn) * tcy;
            u4 = rop[l1 + 3 * nd] - o.qn / (o.sn * o.sn) * tcz;
        } else {
            u2 = rop[l1 + nd]; u3 = rop[l1 + 2 * nd]; u4 = rop[l1 + 3 * nd];
        }
        rop[l + nd] = u2; rop[l + 2 * nd] = u3; rop[l + 3 * nd] = u4;
        rop[l + 4 * nd] = (o.qvar5 * roinv - .5 * (u2 * u2 + u3 * u3 + u4 * u4)) * cvinv;
    } else {
        long li = (long)(ca - a.inc_bc[1]) + (long)(cb - a.inc_bc[2]) * a.inc_bc[0];
        double sn = sqrt(tcx * tcx + tcy * tcy + tcz * tcz);
        double s_1 = 1. / sn, tnx, tny, tnz;
        if (sn == 0.) { tnx = 1.; tny = 0.; tnz = 0.; }
        else { tnx = tcx * s_1; tny = tcy * s_1; tnz = tcz * s_1; }
        double roi = rop[l1], ui = rop[l1 + nd], vi = rop[l1 + 2 * nd], wi = rop[l1 + 3 * nd];
        double Ti = rop[l1 + 4 * nd], pi = roi * rgp * Ti;
        double ro0 = roi, p0 = ro0 * rgp * Ti;
        double roc0 = sqrt(ro0 * gam * p0);
        double pe = a.pext[li];
        double dqn = (pe - pi) / roc0;
        double rog = roi + ro0 * ro0 * dqn / roc0;
        rop[l] = rog;
        rop[l + nd] = ui - dqn * tnx; rop[l + 2 * nd] = vi - dqn * tny; rop[l + 3 * nd] = wi - dqn * tnz;
        rop[l + 4 * nd] = pe / (rog * rgp);
    }
    // outer ranks in the reference loop order (ascending index)
    const int glo = up ? g0 + 1 : a.w[2 * d], ghi = up ? a.w[2 * d + 1] : g0 - 1;
    for (int g = glo; g <= ghi; g++) {
        c[d] = g;
        const long lg = IDX(z, c[0], c[1], c[2]);
        const long q0 = lg + sinc, q1 = lg + 2 * sinc, q2 = lg + 3 * sinc;
        if (a.kind == 4) {
            rop[lg] = rop[q0] * c3 + rop[q1] * c1 + rop[q2] * c2;
            if (fabs(rop[lg] / rop[q0] - 1.) >= 0.7) {
#pragma unroll
                for (int e = 0; e < 5; e++) rop[lg + e * nd] = rop[q0 + e * nd];
            } else {
#pragma unroll
                for (int e = 1; e < 5; e++)
                    rop[lg + e * nd] = rop[q0 + e * nd] * c3 + rop[q1 + e * nd] * c1 + rop[q2 + e * nd] * c2;
            }
        } else {
#pragma unroll
            for (int e = 0; e < 5; e++) rop[lg + e * nd] = rop[l + e * nd];
        }
    }
}

// Edges: 12 bands of nf x nf ghosts along the zone edges; mean of the two face ghosts (which no
// edge or corner writes).  Thread t of band 0..11.
__device__ __forceinline__ void coins_edge(const KZ& z, double* __restrict__ rop, const Box& L, int band, int t) {
    int nf = z.nf, nfk = z.nfk;
    long nd = z.nd;
    int a, b, len;
    // bands 0-3: i-faces x {jmin,jmax} (along k) ; 4-7: i-faces x {kmin,kmax} (along j);
    // 8-11: j-faces x {kmin,kmax} (along i)
    int face = band / 2;   // 0,1: imin/imax with j-band; 2,3 imin/imax with k-band ...
    (void)face;
    int ic, is, jc, js, kc, ks;
    if (band < 4) {
        int iside = band & 1, jside = band >> 1;
        if (!(iside ? L.i1 == z.ni : L.i0 == 1) || !(jside ? L.j1 == z.nj : L.j0 == 1)) return;
        len = L.k1 - L.k0 + 1;
        if (t >= len * nf * nf) return;
        int k = L.k0 + t / (nf * nf); int r = t % (nf * nf); int j = 1 + r / nf; int i = 1 + r % nf;
        ic = iside ? L.i1 : L.i0; is = iside ? 1 : -1; jc = jside ? L.j1 : L.j0; js = jside ? 1 : -1;
        int i1 = i + 1 < nf ? i + 1 : nf, j1 = j + 1 < nf ? j + 1 : nf;
        long l1 = IDX(z, ic + i * is, jc + j * js, k), l11 = IDX(z, ic + i1 * is, jc, k), l12 = IDX(z, ic, jc + j1 * js, k);
#pragma unroll
        for (int e = 0; e < 5; e++) rop[l1 + e * nd] = (rop[l11 + e * nd] + rop[l12 + e * nd]) * 0.5;
    } else if (band < 8) {
        int iside = band & 1, kside = (band - 4) >> 1;
        if (!(iside ? L.i1 == z.ni : L.i0 == 1) || !(kside ? L.k1 == z.nk : L.k0 == 1)) return;
        len = L.j1 - L.j0 + 1;
        if (t >= len * nf * nfk) return;
        int j = L.j0 + t / (nf * nfk); int r = t % (nf * nfk); int k = 1 + r / nf; int i = 1 + r % nf;
        ic = iside ? L.i1 : L.i0; is = iside ? 1 : -1; kc = kside ? L.k1 : L.k0; ks = kside ? 1 : -1;
        int k1 = k + 1 < nfk ? k + 1 : nfk, i1 = i + 1 < nf ? i + 1 : nf;
        long l1 = IDX(z, ic + i * is, j, kc + k * ks), l11 = IDX(z, ic, j, kc + k1 * ks), l12 = IDX(z, ic + i1 * is, j, kc);
#pragma unroll
        for (int e = 0; e < 5; e++) rop[l1 + e * nd] = (rop[l11 + e * nd] + rop[l12 + e * nd]) * 0.5;
    } else {
        int jside = band & 1, kside = (band - 8) >> 1;
        if (!(jside ? L.j1 == z.nj : L.j0 == 1) || !(kside ? L.k1 == z.nk : L.k0 == 1)) return;
        len = L.i1 - L.i0 + 1;
        if (t >= len * nf * nfk) return;
        int i = L.i0 + t % len; int r = t / len; int j = 1 + r % nf; int k = 1 + r / nf;
        jc = jside ? L.j1 : L.j0; js = jside ? 1 : -1; kc = kside ? L.k1 : L.k0; ks = kside ? 1 : -1;
        int k1 = k + 1 < nfk ? k + 1 : nfk, j1 = j + 1 < nf ? j + 1 : nf;
        long l1 = IDX(z, i, jc + j * js, kc + k * ks), l11 = IDX(z, i, jc, kc + k1 * ks), l12 = IDX(z, i, jc + j1 * js, kc);
#pragma unroll
        for (int e = 0; e < 5; e++) rop[l1 + e * nd] = (rop[l11 + e * nd] + rop[l12 + e * nd]) * 0.5;
    }
    (void)a; (void)b;
}
__global__ void k_coins_edges(KZ z, double* __restrict__ rop, Box L) {
    coins_edge(z, rop, L, blockIdx.y, blockIdx.x * blockDim.x + threadIdx.x);
}

// Corners: nf x nf x nfk ghosts, the mean of three face ghosts.  Thread t of corner 0..7 (bit0
// i, bit1 j, bit2 k).
__device__ __forceinline__ void coins_corner(const KZ& z, double* __restrict__ rop, const Box& L, int corner, int t) {
    int nf = z.nf, nfk = z.nfk;
    if (t >= nf * nf * nfk) return;
    int ic = corner & 1, jc = (corner >> 1) & 1, kc = (corner >> 2) & 1;
    int i0 = ic ? L.i1 : L.i0, j0 = jc ? L.j1 : L.j0, k0 = kc ? L.k1 : L.k0;
    if (!((ic ? i0 == z.ni : i0 == 1) && (jc ? j0 == z.nj : j0 == 1) && (kc ? k0 == z.nk : k0 == 1))) return;
    int si = ic ? 1 : -1, sj = jc ? 1 : -1, sk = kc ? 1 : -1;
    int i = 1 + t % nf, j = 1 + (t / nf) % nf, k = 1 + t / (nf * nf);
    int i1 = i + 1 < nf ? i + 1 : nf, j1 = j + 1 < nf ? j + 1 : nf, k1 = k + 1 < nfk ? k + 1 : nfk;
    long nd = z.nd;
    long l = IDX(z, i0 + si * i, j0 + sj * j, k0 + sk * k);
    long l11 = IDX(z, i0, j0, k0 + sk * k1), l12 = IDX(z, i0 + si * i1, j0, k0), l13 = IDX(z, i0, j0 + sj * j1, k0);
    double c3 = 1. / 3.;
#pragma unroll
    for (int e = 0; e < 5; e++) rop[l + e * nd] = (rop[l11 + e * nd] + rop[l12 + e * nd] + rop[l13 + e * nd]) * c3;
}
__global__ void k_coins_corners(KZ z, double* __restrict__ rop, Box L) {
    coins_corner(z, rop, L, blockIdx.x, threadIdx.x);
}

// Edges and corners of several zones in one launch: workgroup (x, y, n) covers threads
// 256 x .. 256 x + 255 of band y (0..11) of zone n, or with y = 12 corner x (x < 8).  Edges and
// corners read face ghosts only and write disjoint cells, so one launch gives what the two
// kernels give in turn.
__global__ __launch_bounds__(256) void k_coins_list(const CoinsZone* __restrict__ zones) {
    const CoinsZone& c = zones[blockIdx.z];
    if (blockIdx.y < 12) coins_edge(c.z, c.rop, c.L, blockIdx.y, blockIdx.x * blockDim.x + threadIdx.x);
    else if (blockIdx.x < 8) coins_corner(c.z, c.rop, c.L, blockIdx.x, threadIdx.x);
}

// ---------------------------------------------------------------------------------------
// Residual norms (cprdu3s1.for restated as a reduction): per block sum of squares and max
// of |drodm| per equation over the interior, then a second pass over the block results.
__global__ void k_norm_partial(KZ z, const double* __restrict__ dr, double* __restrict__ part) {
    __shared__ double s2[5][256];
    __shared__ double mx[5][256];
    int tid = threadIdx.x;
    long ncell = (long)z.ni * z.nj * z.nk;
    double a2[5] = {0, 0, 0, 0, 0}, am[5] = {0, 0, 0, 0, 0};
    for (long c = blockIdx.x * (long)blockDim.x + tid; c < ncell; c += (long)gridDim.x * blockDim.x) {
        int i = 1 + (int)(c % z.ni); long r = c / z.ni; int j = 1 + (int)(r % z.nj); int k = 1 + (int)(r / z.nj);
        long l = IDX(z, i, j, k);
#pragma unroll
        for (int e = 0; e < 5; e++) {
            double v = dr[l + e * z.nd];
            a2[e] += v * v;
            double av = fabs(v);
            if (av > am[e]) am[e] = av;
        }
    }
#pragma unroll
    for (int e = 0; e < 5; e++) { s2[e][tid] = a2[e]; mx[e][tid] = am[e]; }
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w) {
#pragma unroll
            for (int e = 0; e < 5; e++) {
                s2[e][tid] += s2[e][tid + w];
                if (mx[e][tid + w] > mx[e][tid]) mx[e][tid] = mx[e][tid + w];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
#pragma unroll
        for (int e = 0; e < 5; e++) { part[blockIdx.x * 10 + e] = s2[e][0]; part[blockIdx.x * 10 + 5 + e] = mx[e][0]; }
    }
}

// cpcfl3.for without mesh motion: local CFL dt / min(viscous, convective time) per interior
// cell; partial max / min / sum per block (the sum order differs from the reference loop).
__global__ void k_cfl_partial(KZ z, const double* __restrict__ rop, double* __restrict__ part) {
    __shared__ double smax[256], smin[256], ssum[256];
    int tid = threadIdx.x;
    long ncell = (long)z.ni * z.nj * z.nk;
    double rgp = z.cv * (z.gam - 1.), gam1 = z.gam / z.prandt, gam2 = z.gam * rgp, dt = z.dtc;
    bool lamin = z.iflow >= 2, cart = z.typ == 2;
    double surf_c = 0.;
    if (cart) {
        double ddi = 2. * sqrt(z.ti[0] * z.ti[0]), ddj = 2. * sqrt(z.tj[0] * z.tj[0]), ddk = 2. * sqrt(z.tk[0] * z.tk[0]);
        surf_c = 1. / dmax(dmax(ddi, ddj), ddk);
    }
    double cmax = 0., cmin = 1e8, csum = 0.;
    long ndm = z.ndm;
    for (long cidx = blockIdx.x * (long)blockDim.x + tid; cidx < ncell; cidx += (long)gridDim.x * blockDim.x) {
        int i = 1 + (int)(cidx % z.ni); long r = cidx / z.ni; int j = 1 + (int)(r % z.nj); int k = 1 + (int)(r / z.nj);
        long l = IDX(z, i, j, k);
        double surf, vo;
        if (cart) { surf = surf_c; vo = z.vol[0]; }
        else {
            const double* ti = z.ti; const double* tj = z.tj; const double* tk = z.tk;
            long lt = MTR(z, i, j, k), li = lt + z.mi, lj = lt + z.mj, lk = lt + z.mk;
            double ddi = sqrt(ti[li] * ti[li] + ti[li + ndm] * ti[li + ndm] + ti[li + 2 * ndm] * ti[li + 2 * ndm]);
            ddi = ddi + sqrt(ti[lt] * ti[lt] + ti[lt + ndm] * ti[lt + ndm] + ti[lt + 2 * ndm] * ti[lt + 2 * ndm]);
            double ddj = sqrt(tj[lj] * tj[lj] + tj[lj + ndm] * tj[lj + ndm] + tj[lj + 2 * ndm] * tj[lj + 2 * ndm]);
            ddj = ddj + sqrt(tj[lt] * tj[lt] + tj[lt + ndm] * tj[lt + ndm] + tj[lt + 2 * ndm] * tj[lt + 2 * ndm]);
            double ddk = sqrt(tk[lk] * tk[lk] + tk[lk + ndm] * tk[lk + ndm] + tk[lk + 2 * ndm] * tk[lk + 2 * ndm]);
            ddk = ddk + sqrt(tk[lt] * tk[lt] + tk[lt + ndm] * tk[lt + ndm] + tk[lt + 2 * ndm] * tk[lt + 2 * ndm]);
            surf = 1. / dmax(dmax(ddi, ddj), ddk);
            vo = z.vol[lt];
        }
        long nd = z.nd;
        double u = rop[l + nd], v = rop[l + 2 * nd], w = rop[l + 3 * nd], rr = rop[l];
        double c = sqrt(gam2 * rop[l + 4 * nd]);
        double sp = sqrt(u * u + v * v + w * w);
        double xinvspc = 1. / (sp + c);
        double conv = 2. * vo * surf * xinvspc;
        double cfloc;
        if (lamin) {
            double xinvkt = gam1 / z.xmut[l];
            double vis = 2. * rr * vo * vo * surf * surf * xinvkt;
            cfloc = dt / dmin(vis, conv);
        } else cfloc = dt / conv;
        cmax = dmax(cmax, cfloc); cmin = dmin(cmin, cfloc); csum = csum + cfloc;
    }
    smax[tid] = cmax; smin[tid] = cmin; ssum[tid] = csum;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w) {
            smax[tid] = dmax(smax[tid], smax[tid + w]);
            smin[tid] = dmin(smin[tid], smin[tid + w]);
            ssum[tid] += ssum[tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) { part[blockIdx.x * 3] = smax[0]; part[blockIdx.x * 3 + 1] = smin[0]; part[blockIdx.x * 3 + 2] = ssum[0]; }
}

__global__ void k_cfl_final(const double* __restrict__ part, int nblk, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double m = 0., n = 1e8, s = 0.;
    for (int b = 0; b < nblk; b++) { m = dmax(m, part[3 * b]); n = dmin(n, part[3 * b + 1]); s += part[3 * b + 2]; }
    out[0] = m; out[1] = n; out[2] = s;
}

__global__ void k_norm_final(const double* __restrict__ part, int nblk, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double s[5] = {0, 0, 0, 0, 0}, m[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < nblk; b++)
        for (int e = 0; e < 5; e++) {
            s[e] += part[b * 10 + e];
            if (part[b * 10 + 5 + e] > m[e]) m[e] = part[b * 10 + 5 + e];
        }
    for (int e = 0; e < 5; e++) { out[e] = s[e]; out[5 + e] = m[e]; }
}

// Kinetic energy and volume sums over the interior (POST/computePT_enstrophy.cpp, tke part)
// Kinetic energy and enstrophy sums (POST/post.for): Green-Gauss velocity gradients from
// 4-point face values (grad_{i,j,k}_3d{cart,full}.for, c1 = 7/12, c2 = 1/12) accumulated per
// component in the reference loop order, divided by the volume (cp_gradu.for), then
// 1/2 rho |u|^2 vol, 1/2 rho |curl u|^2 vol and vol summed (cptaylor.for; the sums are reduced
// in another order than the reference loop).  On 3dfull zones the i-face above the last cell
// carries the metric index of the face below it, as in cp_gradu_3dfull.for, and the max faces
// follow the cache-block order of the residual.
__device__ __forceinline__ double post_face(const double* __restrict__ q, long l, long inc) {
    const double c1 = 7. / 12, c2 = 1. / 12;
    return (q[l] + q[l - inc]) * c1 - (q[l + inc] + q[l - 2 * inc]) * c2;
}

__global__ void k_post_partial(KZ z, const double* __restrict__ ro, double* __restrict__ part) {
    __shared__ double sk[256], se[256], sv[256];
    int tid = threadIdx.x;
    long ncell = (long)z.ni * z.nj * z.nk;
    double ak = 0., ae = 0., av = 0.;
    const long nd = z.nd, ndm = z.ndm;
    const long inc[3] = {1, z.incj, z.inck};
    const long minc[3] = {z.mi, z.mj, z.mk};
    const double* tm[3] = {z.ti, z.tj, z.tk};
    const bool cart = z.typ == 2;
    for (long c = blockIdx.x * (long)blockDim.x + tid; c < ncell; c += (long)gridDim.x * blockDim.x) {
        int i = 1 + (int)(c % z.ni); long r = c / z.ni; int j = 1 + (int)(r % z.nj); int k = 1 + (int)(r / z.nj);
        long l = IDX(z, i, j, k);
        long lt = cart ? 0 : MTR(z, i, j, k);
        double vo = cart ? z.vol[0] : z.vol[lt];
        double volinv = 1. / vo;
        // post.for runs per cache block: the max face from the next j block comes after the
        // one from the same block in k (as in the residual)
        const bool jn = cb_next(j, z.cbj, z.nbj), kn = cb_next(k, z.cbk, z.nbk);
        double g[3][3];
#pragma unroll
        for (int v = 0; v < 3; v++) {
            const double* q = ro + (1 + v) * nd;
            double flo[3], fhi[3];
#pragma unroll
            for (int d = 0; d < 3; d++) { flo[d] = post_face(q, l, inc[d]); fhi[d] = post_face(q, l + inc[d], inc[d]); }
#pragma unroll
            for (int x = 0; x < 3; x++) {
                double a;
                if (cart) {
                    double t = tm[x][0];
                    a = (0. - flo[x] * t) + fhi[x] * t;
                } else {
                    long lhi = (i == z.ni) ? lt : lt + minc[0];
                    a = 0.;
                    a = a - flo[2] * tm[2][lt + x * ndm];
                    a = a - flo[1] * tm[1][lt + x * ndm];
                    a = a - flo[0] * tm[0][lt + x * ndm];
                    a = a + fhi[0] * tm[0][lhi + x * ndm];
                    if (jn && !kn) {
                        a = a + fhi[2] * tm[2][lt + minc[2] + x * ndm];
                        a = a + fhi[1] * tm[1][lt + minc[1] + x * ndm];
                    } else {
                        a = a + fhi[1] * tm[1][lt + minc[1] + x * ndm];
                        a = a + fhi[2] * tm[2][lt + minc[2] + x * ndm];
                    }
                }
                g[v][x] = a * volinv;
            }
        }
        double rho = ro[l], u = ro[l + nd], v = ro[l + 2 * nd], w = ro[l + 3 * nd];
        ak += 0.5 * rho * (u * u + v * v + w * w) * vo;
        double w1 = (g[2][1] - g[1][2]), w2 = (g[0][2] - g[2][0]), w3 = (g[1][0] - g[0][1]);
        ae += 0.5 * rho * (w1 * w1 + w2 * w2 + w3 * w3) * vo;
        av += vo;
    }
    sk[tid] = ak; se[tid] = ae; sv[tid] = av;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w) { sk[tid] += sk[tid + w]; se[tid] += se[tid + w]; sv[tid] += sv[tid + w]; }
        __syncthreads();
    }
    if (tid == 0) { part[blockIdx.x * 3] = sk[0]; part[blockIdx.x * 3 + 1] = se[0]; part[blockIdx.x * 3 + 2] = sv[0]; }
}

// ---------------------------------------------------------------------------------------
// Launchers
static dim3 blk2(64, 4);

void launch_metric(const KZ& z, hipStream_t st) {
    if (z.typ == 2) { hipLaunchKernelGGL(k_metric_cart, dim3(1), dim3(64), 0, st, z); return; }
    int ni = z.ni + 2 * z.nf + 1, nj = z.nj + 2 * z.nf + 1, nk = z.nk + 2 * z.nfk + 1;
    dim3 g((ni + 63) / 64, (nj + 3) / 4, nk);
    hipLaunchKernelGGL(k_metric_faces, g, blk2, 0, st, z);
    hipLaunchKernelGGL(k_metric_vol, g, blk2, 0, st, z);
}

void launch_coe(const KZ& z, const double* rop, int first_it, int mode, Box b, int lamin, hipStream_t st) {
    dim3 g((b.i1 - b.i0 + 1 + 63) / 64, (b.j1 - b.j0 + 1 + 3) / 4, b.k1 - b.k0 + 1);
    if (z.typ == 2) hipLaunchKernelGGL(k_coe<true>, g, blk2, 0, st, z, rop, first_it, mode, b, lamin);
    else hipLaunchKernelGGL(k_coe<false>, g, blk2, 0, st, z, rop, first_it, mode, b, lamin);
}

void launch_wiggle(const KZ& z, const double* rop, hipStream_t st) {
    dim3 g((z.ni + 1 + 63) / 64, (z.nj + 1 + 3) / 4, z.nk + 1);
    hipLaunchKernelGGL(k_wiggle, g, blk2, 0, st, z, rop);
}

template <int KFLU, int SLOPE>
static void res_lc(const KZ& z, const RhsArgs& a, bool lamin, dim3 g, hipStream_t st) {
    if (lamin) {
        if (z.typ == 2) hipLaunchKernelGGL((k_residual<KFLU, SLOPE, true, true>), g, blk2, 0, st, z, a);
        else hipLaunchKernelGGL((k_residual<KFLU, SLOPE, true, false>), g, blk2, 0, st, z, a);
    } else {
        if (z.typ == 2) hipLaunchKernelGGL((k_residual<KFLU, SLOPE, false, true>), g, blk2, 0, st, z, a);
        else hipLaunchKernelGGL((k_residual<KFLU, SLOPE, false, false>), g, blk2, 0, st, z, a);
    }
}

template <int KFLU, int SLOPE>
static void res_kcol(const KZ& z, const RhsArgs& a, bool lamin, hipStream_t st) {
    const int n = a.kcol_n > 0 ? a.kcol_n : RES_KCOL_N;
    dim3 g((z.ni + 63) / 64, (z.nj + 3) / 4, (z.nk + n - 1) / n);
    if (lamin) {
        if (z.typ == 2) hipLaunchKernelGGL((k_residual_kcol<KFLU, SLOPE, true, true>), g, blk2, 0, st, z, a);
        else hipLaunchKernelGGL((k_residual_kcol<KFLU, SLOPE, true, false>), g, blk2, 0, st, z, a);
    } else {
        if (z.typ == 2) hipLaunchKernelGGL((k_residual_kcol<KFLU, SLOPE, false, true>), g, blk2, 0, st, z, a);
        else hipLaunchKernelGGL((k_residual_kcol<KFLU, SLOPE, false, false>), g, blk2, 0, st, z, a);
    }
}

template <int KFLU, int SLOPE>
static void res_faces(const KZ& z, const RhsArgs& a, bool lamin, hipStream_t st) {
    dim3 g((z.ni + 62) / 63, (z.nj + 3) / 4, (z.nk + RES_KC - 1) / RES_KC), b(64, 5);
    if (lamin) {
        if (z.typ == 2) hipLaunchKernelGGL((k_residual_faces<KFLU, SLOPE, true, true>), g, b, 0, st, z, a);
        else hipLaunchKernelGGL((k_residual_faces<KFLU, SLOPE, true, false>), g, b, 0, st, z, a);
    } else {
        if (z.typ == 2) hipLaunchKernelGGL((k_residual_faces<KFLU, SLOPE, false, true>), g, b, 0, st, z, a);
        else hipLaunchKernelGGL((k_residual_faces<KFLU, SLOPE, false, false>), g, b, 0, st, z, a);
    }
}

int launch_residual(const KZ& z, const RhsArgs& a, hipStream_t st) {
    bool lamin = z.iflow == 2;
    if (a.schedule == RES_KCOL) {
        if (z.kflu == 5) {
            if (z.slope == 1) res_kcol<5, 1>(z, a, lamin, st);
            else if (z.slope == 2) res_kcol<5, 2>(z, a, lamin, st);
            else if (z.slope == 3) res_kcol<5, 3>(z, a, lamin, st);
            else return -1;
        } else if (z.kflu == 1 && z.slope == 2) res_kcol<1, 2>(z, a, lamin, st);
        else if (z.kflu == 2 && z.slope == 2) res_kcol<2, 2>(z, a, lamin, st);
        else return -1;
        return 0;
    }
    if (a.schedule == RES_FACES && a.ncorr == 0) {
        if (z.kflu == 5) {
            if (z.slope == 1) res_faces<5, 1>(z, a, lamin, st);
            else if (z.slope == 2) res_faces<5, 2>(z, a, lamin, st);
            else if (z.slope == 3) res_faces<5, 3>(z, a, lamin, st);
            else return -1;
        } else if (z.kflu == 1 && z.slope == 2) res_faces<1, 2>(z, a, lamin, st);
        else if (z.kflu == 2 && z.slope == 2) res_faces<2, 2>(z, a, lamin, st);
        else return -1;
        return 0;
    }
    dim3 g((z.ni + 63) / 64, (z.nj + 3) / 4, z.nk);
    if (z.kflu == 5) {
        if (z.slope == 1) res_lc<5, 1>(z, a, lamin, g, st);
        else if (z.slope == 2) res_lc<5, 2>(z, a, lamin, g, st);
        else if (z.slope == 3) res_lc<5, 3>(z, a, lamin, g, st);
        else return -1;
    } else if (z.kflu == 1 && z.slope == 2) res_lc<1, 2>(z, a, lamin, g, st);
    else if (z.kflu == 2 && z.slope == 2) res_lc<2, 2>(z, a, lamin, g, st);
    else return -1;
    return 0;
}

void launch_bc_face(const KZ& z, double* rop, const BcArgs& a, hipStream_t st) {
    int n0 = a.w[1] - a.w[0] + 1, n1 = a.w[3] - a.w[2] + 1, n2 = a.w[5] - a.w[4] + 1;
    if (n0 <= 0 || n1 <= 0 || n2 <= 0) return;
    if (a.kind >= 4) {
        int d = (a.idir - 1) / 2;
        int ncol = (d == 0 ? n1 * n2 : d == 1 ? n0 * n2 : n0 * n1);
        hipLaunchKernelGGL(k_bc_column, dim3((ncol + 127) / 128), dim3(128), 0, st, z, rop, a);
        return;
    }
    dim3 g((n0 * n1 + 255) / 256, n2);
    hipLaunchKernelGGL(k_bc_face, g, dim3(256), 0, st, z, rop, a);
}

void launch_coins(const KZ& z, double* rop, Box L, hipStream_t st) {
    int mx = L.i1 - L.i0 + 1;
    if (L.j1 - L.j0 + 1 > mx) mx = L.j1 - L.j0 + 1;
    if (L.k1 - L.k0 + 1 > mx) mx = L.k1 - L.k0 + 1;
    int nt = mx * z.nf * (z.nf > z.nfk ? z.nf : z.nfk);
    hipLaunchKernelGGL(k_coins_edges, dim3((nt + 255) / 256, 12), dim3(256), 0, st, z, rop, L);
    hipLaunchKernelGGL(k_coins_corners, dim3(8), dim3(64), 0, st, z, rop, L);
}

int coins_blocks(const KZ& z, const Box& L) {
    int mx = L.i1 - L.i0 + 1;
    if (L.j1 - L.j0 + 1 > mx) mx = L.j1 - L.j0 + 1;
    if (L.k1 - L.k0 + 1 > mx) mx = L.k1 - L.k0 + 1;
    const int nt = mx * z.nf * (z.nf > z.nfk ? z.nf : z.nfk);
    const int nb = (nt + 255) / 256;
    return nb > 8 ? nb : 8;
}

void launch_coins_list(const CoinsZone* zones, int nz, int nblk, hipStream_t st) {
    if (nz <= 0) return;
    hipLaunchKernelGGL(k_coins_list, dim3((unsigned)nblk, 13, (unsigned)nz), dim3(256), 0, st, zones);
}

void launch_norms(const KZ& z, const double* dr, double* part, int nblk, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_norm_partial, dim3(nblk), dim3(256), 0, st, z, dr, part);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(64), 0, st, part, nblk, out);
}

void launch_cfl(const KZ& z, const double* rop, double* part, int nblk, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_cfl_partial, dim3(nblk), dim3(256), 0, st, z, rop, part);
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(64), 0, st, part, nblk, out);
}

void launch_tke(const KZ& z, const double* ro, double* part, int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_post_partial, dim3(nblk), dim3(256), 0, st, z, ro, part);
}

}  // namespace fm

// ---------------------------------------------------------------------------------------
// Injection transfers.  Receivers are ghost cells and donors interior cells, so one launch
// moves every raccord of every zone pair at once: one thread per point, 5 variables each.
namespace fm {

__global__ void k_transfer(const TransferPoint* __restrict__ pts, long npts, TransferArgs a) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts) return;
    TransferPoint q = pts[t];
    const double* __restrict__ d = a.f[q.zd];
    double* __restrict__ r = a.f[q.zr];
    long ndd = a.nd[q.zd], ndr = a.nd[q.zr];
#pragma unroll
    for (int v = 0; v < 5; v++) r[q.rcv + v * ndr] = d[q.dnr + v * ndd];
}

// Transpose of one raccord's injection: the gradient held by each receiver cell moves onto its
// donor cell (g_d[dnr] += g_r[rcv], then g_r[rcv] = 0).  Point-parallel when the raccord's donor
// and receiver cells are all distinct; otherwise one thread takes the points last to first, the
// transpose of an injection whose later points overwrite the earlier ones.
__device__ inline void fold_point(const TransferPoint& q, const TransferArgs& g) {
    double* __restrict__ d = g.f[q.zd];
    double* __restrict__ r = g.f[q.zr];
    long ndd = g.nd[q.zd], ndr = g.nd[q.zr];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        double x = r[q.rcv + v * ndr];
        r[q.rcv + v * ndr] = 0.;
        d[q.dnr + v * ndd] += x;
    }
}

__global__ void k_transfer_vjp(const TransferPoint* __restrict__ pts, long npts, TransferArgs g) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts) return;
    fold_point(pts[t], g);
}

__global__ void k_transfer_vjp_serial(const TransferPoint* __restrict__ pts, long npts, TransferArgs g) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (long t = npts - 1; t >= 0; t--) fold_point(pts[t], g);
}

void launch_transfer_vjp(const TransferPoint* pts, long npts, const TransferArgs& g, int serial, hipStream_t st) {
    if (npts <= 0) return;
    if (serial) {
        hipLaunchKernelGGL(k_transfer_vjp_serial, dim3(1), dim3(64), 0, st, pts, npts, g);
        return;
    }
    long nb = (npts + 255) / 256;
    hipLaunchKernelGGL(k_transfer_vjp, dim3((unsigned)nb), dim3(256), 0, st, pts, npts, g);
}

__global__ void k_transfer_serial(const TransferPoint* __restrict__ pts, long npts, TransferArgs a) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (long t = 0; t < npts; t++) {
        TransferPoint q = pts[t];
        long ndd = a.nd[q.zd], ndr = a.nd[q.zr];
        for (int v = 0; v < 5; v++) a.f[q.zr][q.rcv + v * ndr] = a.f[q.zd][q.dnr + v * ndd];
    }
}

// One raccord's points; point by point, first to last, when it repeats a receiver cell.
void launch_transfer_part(const TransferPoint* pts, long npts, const TransferArgs& a, int serial, hipStream_t st) {
    if (npts <= 0) return;
    if (serial) {
        hipLaunchKernelGGL(k_transfer_serial, dim3(1), dim3(64), 0, st, pts, npts, a);
        return;
    }
    launch_transfer(pts, npts, a, st);
}

void launch_transfer(const TransferPoint* pts, long npts, const TransferArgs& a, hipStream_t st) {
    if (npts <= 0) return;
    long nb = (npts + 255) / 256;
    hipLaunchKernelGGL(k_transfer, dim3((unsigned)nb), dim3(256), 0, st, pts, npts, a);
}

}  // namespace fm

// ---------------------------------------------------------------------------------------
// Wall efforts, eff_flu<scheme>_<model>_<slope>_<mesh> of template_effort.for:100-440: the
// balance's own face flux on every face of a window, minus the free-stream pressure, scaled by
// 2 / (rho_inf V_inf^2) and signed outward; with the face gradients, pressure coefficient,
// viscosity, density and pressure.  One thread per face; the host sums the 12 contributions
// in the reference's loop order (k, j, i).
namespace fm {

template <int KFLU, int SLOPE, bool LAMIN, bool CART, int D>
__global__ void k_effort(KZ z, EffArgs a) {
    const int n0 = a.w[1] - a.w[0] + 1, n1 = a.w[3] - a.w[2] + 1, n2 = a.w[5] - a.w[4] + 1;
    const long nf = (long)n0 * n1 * n2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf) return;
    const int i = a.w[0] + (int)(t % n0), j = a.w[2] + (int)((t / n0) % n1), k = a.w[4] + (int)(t / ((long)n0 * n1));
    const double* __restrict__ rop = a.rop;
    FluxConst f = flux_const(z);
    const long l = IDX(z, i, j, k), nd = z.nd, ndm = z.ndm;
    const long inc = D == 0 ? 1 : D == 1 ? z.incj : z.inck;
    double qm[5], qp[5], tn[3], sn, fl[5], gv[13];
#pragma unroll
    for (int q = 0; q < 12; q++) gv[q] = 0.;
    gv[12] = f.volinv;
    long lt = 0;
    if (CART) {
        tn[0] = f.tc[0]; tn[1] = f.tc[1]; tn[2] = f.tc[2]; sn = f.s[D];
    } else {
        const double* tm = D == 0 ? z.ti : D == 1 ? z.tj : z.tk;
        lt = MTR(z, i, j, k);
        tn[0] = tm[lt]; tn[1] = tm[lt + ndm]; tn[2] = tm[lt + 2 * ndm];
        sn = sqrt(tn[0] * tn[0] + tn[1] * tn[1] + tn[2] * tn[2]);
    }
    slope_state<SLOPE>(f, rop, nd, l, inc, qm, qp);
    inviscid_flux<KFLU, CART, D>(f, rop, z.wig, nd, l, tn, sn, qm, qp, fl);
    if (LAMIN) {
        if (CART) visc_cart<D>(f, z, rop, z.xmut, l, fl, gv);
        else visc_full<D>(f, z, rop, z.xmut, l, lt, tn, fl, gv);
    }
    const double pinf = z.pinf;
    const double norm = 2. / (z.roinf * (z.vinf * z.vinf));
    const bool odd = (a.idir % 2) == 1;
    const double sens = odd ? -norm : norm, sens1 = odd ? -1. : 1.;
    const double flag = CART ? 0. : 1.;
    const double p1 = qp[0] * qp[4] * f.rgp, p2 = qm[0] * qm[4] * f.rgp;
    double v[21];
    v[0] = fl[0] * sens1;
    v[1] = (fl[1] - pinf * tn[0] * (D == 0 ? 1. : flag)) * sens;
    v[2] = (fl[2] - pinf * tn[1] * (D == 1 ? 1. : flag)) * sens;
    v[3] = (fl[3] - pinf * tn[2] * (D == 2 ? 1. : flag)) * sens;
    if (D == 0) v[1] = (fl[1] - pinf * tn[0]) * sens;
    if (D == 1) v[2] = (fl[2] - pinf * tn[1]) * sens;
    if (D == 2) v[3] = (fl[3] - pinf * tn[2]) * sens;
    v[4] = fl[4] * sens;
#pragma unroll
    for (int q = 0; q < 12; q++) v[5 + q] = gv[q] * 2. * gv[12];
    v[17] = (0.5 * (p1 + p2) - pinf) * norm;
    v[18] = LAMIN ? 0.5 * (z.xmut[l] + z.xmut[l - inc]) : 0.;
    // Roe: etat_roe_GD.for leaves the square roots of the densities in qm1, qp1
    v[19] = KFLU == 5 ? 0.5 * (sqrt(qm[0]) + sqrt(qp[0])) : 0.5 * (qm[0] + qp[0]);
    v[20] = 0.5 * (p1 + p2);
    long x1, x2;
    if (D == 0) { x1 = z.xi; x2 = (long)z.xi * z.xj; }
    else if (D == 1) { x1 = 1; x2 = (long)z.xi * z.xj; }
    else { x1 = 1; x2 = z.xi; }
    const long x3 = x1 + x2, lx = XYZ(z, i, j, k);
    const double f4 = 0.25 * (z.x[lx] + z.x[lx + x1] + z.x[lx + x2] + z.x[lx + x3]);
    const double f5 = 0.25 * (z.y[lx] + z.y[lx + x1] + z.y[lx + x2] + z.y[lx + x3]);
    const double f6 = 0.25 * (z.z[lx] + z.z[lx + x1] + z.z[lx + x2] + z.z[lx + x3]);
    double* o = a.out;
#pragma unroll
    for (int q = 0; q < 21; q++) o[t + q * nf] = v[q];
    o[t + 21 * nf] = v[1]; o[t + 22 * nf] = v[2]; o[t + 23 * nf] = v[3];
    o[t + 24 * nf] = v[3] * (f5 - a.pos[1]) - v[2] * (f6 - a.pos[2]);
    o[t + 25 * nf] = v[1] * (f6 - a.pos[2]) - v[3] * (f4 - a.pos[0]);
    o[t + 26 * nf] = v[2] * (f4 - a.pos[0]) - v[1] * (f5 - a.pos[1]);
    o[t + 27 * nf] = sn;
    o[t + 28 * nf] = v[0];
#pragma unroll
    for (int e = 1; e < 5; e++) o[t + (28 + e) * nf] = fl[e] * sens1;
}

template <int KFLU, int SLOPE, bool LAMIN, bool CART>
static void effort_d(const KZ& z, const EffArgs& a, dim3 g, hipStream_t st) {
    int d = (a.idir - 1) / 2;
    if (d == 0) hipLaunchKernelGGL((k_effort<KFLU, SLOPE, LAMIN, CART, 0>), g, dim3(256), 0, st, z, a);
    else if (d == 1) hipLaunchKernelGGL((k_effort<KFLU, SLOPE, LAMIN, CART, 1>), g, dim3(256), 0, st, z, a);
    else hipLaunchKernelGGL((k_effort<KFLU, SLOPE, LAMIN, CART, 2>), g, dim3(256), 0, st, z, a);
}

template <int KFLU, int SLOPE>
static void effort_m(const KZ& z, const EffArgs& a, dim3 g, hipStream_t st) {
    bool lamin = z.iflow == 2;
    if (lamin) { if (z.typ == 2) effort_d<KFLU, SLOPE, true, true>(z, a, g, st); else effort_d<KFLU, SLOPE, true, false>(z, a, g, st); }
    else { if (z.typ == 2) effort_d<KFLU, SLOPE, false, true>(z, a, g, st); else effort_d<KFLU, SLOPE, false, false>(z, a, g, st); }
}

// The 12 sums of a window's effort contributions (rows 21..32 of out), face by face in the
// window order, as fm_wall_effort adds them on the host: thread q adds row q sequentially, so
// the order is fixed.
__global__ void k_effort_sum(const double* __restrict__ out, long nf, double* __restrict__ sums) {
    const int q = threadIdx.x;
    if (q >= 12 || blockIdx.x != 0) return;
    const double* __restrict__ row = out + (21 + q) * nf;
    double s = 0.;
    for (long t = 0; t < nf; t++) s = s + row[t];
    sums[q] = s;
}

void launch_effort_sum(const double* out, long nf, double* sums, hipStream_t st) {
    hipLaunchKernelGGL(k_effort_sum, dim3(1), dim3(64), 0, st, out, nf, sums);
}

int launch_effort(const KZ& z, const EffArgs& a, hipStream_t st) {
    long nf = (long)(a.w[1] - a.w[0] + 1) * (a.w[3] - a.w[2] + 1) * (a.w[5] - a.w[4] + 1);
    if (nf <= 0) return 0;
    dim3 g((unsigned)((nf + 255) / 256));
    if (z.kflu == 5) {
        if (z.slope == 1) effort_m<5, 1>(z, a, g, st);
        else if (z.slope == 2) effort_m<5, 2>(z, a, g, st);
        else if (z.slope == 3) effort_m<5, 3>(z, a, g, st);
        else return -1;
    } else if (z.kflu == 1 && z.slope == 2) effort_m<1, 2>(z, a, g, st);
    else if (z.kflu == 2 && z.slope == 2) effort_m<2, 2>(z, a, g, st);
    else return -1;
    return 0;
}

}  // namespace fm
