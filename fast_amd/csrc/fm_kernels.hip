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

// ---------------------------------------------------------------------------------------
// Residual: fused init_rhs + 6 face fluxes + balance + update, one thread per interior cell.
// Accumulation order per cell as the reference sweep (template_FluxAndBalance.for:192-271):
//   init, +Fk(k), +Fj(j), +Fi(i), -Fi(i+1), -Fj(j+1), -Fk(k+1).
// Update: core3as2.for (implicit, writes the Newton residual into drodm) or core3ark3.for
// (explicit, keeps drodm and writes the new stage state).
template <int KFLU, int SLOPE, bool LAMIN, bool CART>
__global__ void __launch_bounds__(256) k_residual(KZ z, RhsArgs a) {
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
    face_flux<KFLU, SLOPE, LAMIN, CART, 1>(f, z, rop, i, j + 1, k, flu);
#pragma unroll
    for (int e = 0; e < 5; e++) acc[e] = acc[e] - flu[e];
    face_flux<KFLU, SLOPE, LAMIN, CART, 2>(f, z, rop, i, j, k + 1, flu);
#pragma unroll
    for (int e = 0; e < 5; e++) acc[e] = acc[e] - flu[e];

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

// ---------------------------------------------------------------------------------------
// LU-SGS on i+j+k hyperplanes (invlu_l.for / invlu_u.for).  Cells of one hyperplane are
// independent; per cell the neighbour terms are added in the reference order.
struct LUC { double gam1, gamm1, cp, gam2; };

template <bool CART, int D, bool LOWER>
__device__ __forceinline__ void lu_term(const KZ& z, const LUC& c, const double* __restrict__ rop,
                                        const double* __restrict__ dw, long l1, long lt,
                                        double signe, double b[5]) {
    long nd = z.nd;
    double u = rop[l1 + nd], v = rop[l1 + 2 * nd], w = rop[l1 + 3 * nd], t = rop[l1 + 4 * nd];
    double q2 = .5 * (u * u + v * v + w * w);
    double h = c.cp * t + q2;
    double ph2 = c.gamm1 * q2;
    double cs = z.coe[l1 + (1 + D) * nd] * signe;
    double gam1 = c.gam1, gamm1 = c.gamm1, gam2 = c.gam2;
    double d1 = dw[l1], d2 = dw[l1 + nd], d3 = dw[l1 + 2 * nd], d4 = dw[l1 + 3 * nd], d5 = dw[l1 + 4 * nd];
    if (CART) {
        double tc = D == 0 ? z.ti[0] : D == 1 ? z.tj[0] : z.tk[0];
        if (D == 0) {
            double qn = tc * u;
            double b21 = (tc * ph2 - u * qn), b22 = (qn - tc * gam2 * u) + cs, b23 = (-tc * gamm1 * v);
            double b24 = (-tc * gamm1 * w), b25 = tc * gamm1;
            double b51 = qn * (ph2 - h), b52 = (tc * h - gamm1 * u * qn), b53 = (-gamm1 * v * qn);
            double b54 = (-gamm1 * w * qn), b55 = (gam1 * qn) + cs;
            b[0] = cs * d1 + tc * d2;
            b[1] = b21 * d1 + b22 * d2 + b23 * d3 + b24 * d4 + b25 * d5;
            b[2] = (-v * qn) * d1 + (tc * v) * d2 + ((qn) + cs) * d3;
            b[3] = (-w * qn) * d1 + (tc * w) * d2 + ((qn) + cs) * d4;
            b[4] = b51 * d1 + b52 * d2 + b53 * d3 + b54 * d4 + b55 * d5;
        } else if (D == 1) {
            double qn = tc * v;
            double b31 = (tc * ph2 - v * qn), b32 = (-tc * gamm1 * u), b33 = (qn - tc * gam2 * v) + cs;
            double b34 = (-tc * gamm1 * w), b35 = tc * gamm1;
            double b51 = qn * (ph2 - h), b52 = (-gamm1 * u * qn), b53 = (tc * h - gamm1 * v * qn);
            double b54 = (-gamm1 * w * qn), b55 = (gam1 * qn) + cs;
            b[0] = cs * d1 + tc * d3;
            b[1] = (-u * qn) * d1 + (qn + cs) * d2 + (tc * u) * d3;
            b[2] = b31 * d1 + b32 * d2 + b33 * d3 + b34 * d4 + b35 * d5;
            b[3] = (-w * qn) * d1 + (tc * w) * d3 + ((qn) + cs) * d4;
            b[4] = b51 * d1 + b52 * d2 + b53 * d3 + b54 * d4 + b55 * d5;
        } else {
            double qn = tc * w;
            double b41 = tc * ph2 - w * qn, b42 = -tc * gamm1 * u, b43 = -tc * gamm1 * v;
            double b44 = (qn - tc * gam2 * w) + cs, b45 = tc * gamm1;
            double b51 = qn * (ph2 - h), b52 = -gamm1 * u * qn, b53 = -gamm1 * v * qn;
            double b54 = tc * h - gamm1 * w * qn, b55 = gam1 * qn + cs;
            b[0] = cs * d1 + tc * d4;
            b[1] = (-u * qn) * d1 + (qn + cs) * d2 + (tc * u) * d4;
            b[2] = (-v * qn) * d1 + (qn + cs) * d3 + (tc * v) * d4;
            b[3] = b41 * d1 + b42 * d2 + b43 * d3 + b44 * d4 + b45 * d5;
            b[4] = b51 * d1 + b52 * d2 + b53 * d3 + b54 * d4 + b55 * d5;
        }
        return;
    }
    const double* tm = D == 0 ? z.ti : D == 1 ? z.tj : z.tk;
    long inc = D == 0 ? z.mi : D == 1 ? z.mj : z.mk;
    long ndm = z.ndm, lt1, lt2;
    if (LOWER) { lt1 = lt - inc; lt2 = lt; }
    else { lt1 = lt + inc; lt2 = lt + 2 * inc; }
    double tcx = 0.5 * (tm[lt2] + tm[lt1]);
    double tcy = 0.5 * (tm[lt2 + ndm] + tm[lt1 + ndm]);
    double tcz = 0.5 * (tm[lt2 + 2 * ndm] + tm[lt1 + 2 * ndm]);
    double qn = tcx * u + tcy * v + tcz * w;
    double b21 = (tcx * ph2 - u * qn), b22 = (qn - tcx * gam2 * u) + cs;
    double b23 = (tcy * u - tcx * gamm1 * v), b24 = (tcz * u - tcx * gamm1 * w), b25 = tcx * gamm1;
    double b31 = (tcy * ph2 - v * qn), b32 = (tcx * v - tcy * gamm1 * u);
    double b33 = (qn - tcy * gam2 * v) + cs, b34 = (tcz * v - tcy * gamm1 * w), b35 = tcy * gamm1;
    double b41 = (tcz * ph2 - w * qn), b42 = (tcx * w - tcz * gamm1 * u);
    double b43 = (tcy * w - tcz * gamm1 * v), b44 = (qn - tcz * gam2 * w) + cs, b45 = tcz * gamm1;
    double b51 = qn * (ph2 - h), b52 = (tcx * h - gamm1 * u * qn), b53 = (tcy * h - gamm1 * v * qn);
    double b54 = (tcz * h - gamm1 * w * qn), b55 = (gam1 * qn) + cs;
    b[0] = cs * d1 + tcx * d2 + tcy * d3 + tcz * d4;
    b[1] = b21 * d1 + b22 * d2 + b23 * d3 + b24 * d4 + b25 * d5;
    b[2] = b31 * d1 + b32 * d2 + b33 * d3 + b34 * d4 + b35 * d5;
    b[3] = b41 * d1 + b42 * d2 + b43 * d3 + b44 * d4 + b45 * d5;
    b[4] = b51 * d1 + b52 * d2 + b53 * d3 + b54 * d4 + b55 * d5;
}

template <bool CART>
__global__ void k_lu_lower_plane(KZ z, LuArgs a, int s) {
    int jj = blockIdx.x * blockDim.x + threadIdx.x;
    int kk = blockIdx.y;
    int nj = a.j1 - a.j0 + 1, nk = a.k1 - a.k0 + 1, ni = a.i1 - a.i0 + 1;
    if (jj >= nj || kk >= nk) return;
    int ii = s - jj - kk;
    if (ii < 0 || ii >= ni) return;
    int i = a.i0 + ii, j = a.j0 + jj, k = a.k0 + kk;
    long nd = z.nd, l = IDX(z, i, j, k);
    long lt = CART ? 0 : MTR(z, i, j, k);
    LUC c; c.gam1 = z.gam; c.gamm1 = c.gam1 - 1.; c.cp = z.gam * z.cv; c.gam2 = c.gamm1 - 1.;
    double* dw = a.dw;
    const double* rop = a.rop;
    double xal = z.coe[l] * 0.5;
    double x[5], b[5];
#pragma unroll
    for (int e = 0; e < 5; e++) x[e] = dw[l + e * nd];
    if (CART) {
        if (jj > 0) { lu_term<CART, 1, true>(z, c, rop, dw, l - z.incj, lt, 0.5, b);
#pragma unroll
            for (int e = 0; e < 5; e++) x[e] = x[e] + b[e] * xal; }
        if (kk > 0) { lu_term<CART, 2, true>(z, c, rop, dw, l - z.inck, lt, 0.5, b);
#pragma unroll
            for (int e = 0; e < 5; e++) x[e] = x[e] + b[e] * xal; }
    } else {
        if (kk > 0) { lu_term<CART, 2, true>(z, c, rop, dw, l - z.inck, lt, 0.5, b);
#pragma unroll
            for (int e = 0; e < 5; e++) x[e] = x[e] + b[e] * xal; }
        if (jj > 0) { lu_term<CART, 1, true>(z, c, rop, dw, l - z.incj, lt, 0.5, b);
#pragma unroll
            for (int e = 0; e < 5; e++) x[e] = x[e] + b[e] * xal; }
    }
    if (ii > 0) { lu_term<CART, 0, true>(z, c, rop, dw, l - 1, lt, 0.5, b);
#pragma unroll
        for (int e = 0; e < 5; e++) x[e] = x[e] + b[e] * xal; }
    double diag = 1. / z.coe[l + 4 * nd];
#pragma unroll
    for (int e = 0; e < 5; e++) dw[l + e * nd] = x[e] * diag;
}

// Upper sweep + Newton update (mjr_newton.for).  J and I terms and the Newton base state use
// rop; the K term uses rop_k (the already updated state of plane k+1 when the sweep works in
// place, as the reference does for sub-iterations >= 2); the new state goes to out.
template <bool CART>
__global__ void k_lu_upper_plane(KZ z, LuArgs a, int s) {
    int jj = blockIdx.x * blockDim.x + threadIdx.x;
    int kk = blockIdx.y;
    int nj = a.j1 - a.j0 + 1, nk = a.k1 - a.k0 + 1, ni = a.i1 - a.i0 + 1;
    if (jj >= nj || kk >= nk) return;
    int ii = s - jj - kk;
    if (ii < 0 || ii >= ni) return;
    // reversed coordinates: plane s counts from the top corner
    int i = a.i1 - ii, j = a.j1 - jj, k = a.k1 - kk;
    long nd = z.nd, l = IDX(z, i, j, k);
    long lt = CART ? 0 : MTR(z, i, j, k);
    LUC c; c.gam1 = z.gam; c.gamm1 = c.gam1 - 1.; c.cp = z.gam * z.cv; c.gam2 = c.gamm1 - 1.;
    double* dw = a.dw;
    double xal = z.coe[l] * -0.5;
    double dg = z.coe[l + 4 * nd];
    double x[5], b[5];
#pragma unroll
    for (int e = 0; e < 5; e++) x[e] = dw[l + e * nd] * dg;
    if (kk > 0) { lu_term<CART, 2, false>(z, c, a.rop_k, dw, l + z.inck, lt, -0.5, b);
#pragma unroll
        for (int e = 0; e < 5; e++) x[e] = x[e] + b[e] * xal; }
    if (jj > 0) { lu_term<CART, 1, false>(z, c, a.rop, dw, l + z.incj, lt, -0.5, b);
#pragma unroll
        for (int e = 0; e < 5; e++) x[e] = x[e] + b[e] * xal; }
    if (ii > 0) { lu_term<CART, 0, false>(z, c, a.rop, dw, l + 1, lt, -0.5, b);
#pragma unroll
        for (int e = 0; e < 5; e++) x[e] = x[e] + b[e] * xal; }
    double diag = 1. / z.coe[l + 4 * nd];
#pragma unroll
    for (int e = 0; e < 5; e++) { x[e] = x[e] * diag; dw[l + e * nd] = x[e]; }
    if (a.newton) {
        const double* R = a.rop;
        double* O = a.out;
        double cvinv = 1. / z.cv, cvinv2 = 0.5 * cvinv;
        double ro_old = R[l], u_old = R[l + nd], v_old = R[l + 2 * nd], w_old = R[l + 3 * nd], t_old = R[l + 4 * nd];
        double r1 = R[l] + x[0];
        double r_1 = 1. / r1;
        double u1 = (ro_old * u_old + x[1]) * r_1, v1 = (ro_old * v_old + x[2]) * r_1, w1 = (ro_old * w_old + x[3]) * r_1;
        double roe_old = ro_old * (z.cv * t_old + 0.5 * (u_old * u_old + v_old * v_old + w_old * w_old));
        O[l] = r1; O[l + nd] = u1; O[l + 2 * nd] = v1; O[l + 3 * nd] = w1;
        O[l + 4 * nd] = (roe_old + x[4]) * r_1 * cvinv - cvinv2 * (u1 * u1 + v1 * v1 + w1 * w1);
    }
}

// ---------------------------------------------------------------------------------------
// Ghost cells.  Face BC over a window (BCExtrapolate.for, BCPeriodic.for,
// BCWallViscous_{i,j,k}.for), then correct_coins.for edges and corners.
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
    else if (a.kind == 1) c[d] = up ? 2 * a.w[2 * d] - 1 - c[d] : 2 * a.w[2 * d + 1] + 1 - c[d];
    else c[d] = up ? c[d] - n : c[d] + n;
    long l1 = IDX(z, c[0], c[1], c[2]);
    double c1 = a.lrhs ? 0. : 1.;
    if (a.kind == 0) {
#pragma unroll
        for (int e = 0; e < 5; e++) rop[l + e * nd] = dmax(-1.e30, rop[l1 + e * nd]) * c1;
    } else if (a.kind == 2) {
#pragma unroll
        for (int e = 0; e < 5; e++) rop[l + e * nd] = rop[l1 + e * nd] * c1;
    } else {
        double u = rop[l1 + nd], v = rop[l1 + 2 * nd], w = rop[l1 + 3 * nd];
        rop[l] = rop[l1];
        rop[l + nd] = 0. - u; rop[l + 2 * nd] = 0. - v; rop[l + 3 * nd] = 0. - w;
        rop[l + 4 * nd] = rop[l1 + 4 * nd];
    }
}

// Edges: 12 bands of nf x nf ghosts along the zone edges; mean of the two face ghosts.
__global__ void k_coins_edges(KZ z, double* __restrict__ rop, Box L) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int band = blockIdx.y;       // 0..11
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

__global__ void k_coins_corners(KZ z, double* __restrict__ rop, Box L) {
    int corner = blockIdx.x;   // 0..7: bit0 i, bit1 j, bit2 k
    int t = threadIdx.x;
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
__global__ void k_tke_partial(KZ z, const double* __restrict__ ro, double* __restrict__ part) {
    __shared__ double sk[256];
    __shared__ double sv[256];
    int tid = threadIdx.x;
    long ncell = (long)z.ni * z.nj * z.nk;
    double ak = 0., av = 0.;
    for (long c = blockIdx.x * (long)blockDim.x + tid; c < ncell; c += (long)gridDim.x * blockDim.x) {
        int i = 1 + (int)(c % z.ni); long r = c / z.ni; int j = 1 + (int)(r % z.nj); int k = 1 + (int)(r / z.nj);
        long l = IDX(z, i, j, k);
        double vol = z.typ == 2 ? z.vol[0] : z.vol[MTR(z, i, j, k)];
        double u = ro[l + z.nd], v = ro[l + 2 * z.nd], w = ro[l + 3 * z.nd];
        ak += 0.5 * ro[l] * (u * u + v * v + w * w) * vol;
        av += vol;
    }
    sk[tid] = ak; sv[tid] = av;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w) { sk[tid] += sk[tid + w]; sv[tid] += sv[tid + w]; }
        __syncthreads();
    }
    if (tid == 0) { part[blockIdx.x * 2] = sk[0]; part[blockIdx.x * 2 + 1] = sv[0]; }
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

int launch_residual(const KZ& z, const RhsArgs& a, hipStream_t st) {
    dim3 g((z.ni + 63) / 64, (z.nj + 3) / 4, z.nk);
    bool lamin = z.iflow == 2;
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

void launch_lu(const KZ& z, const LuArgs& a, hipStream_t st) {
    int nj = a.j1 - a.j0 + 1, nk = a.k1 - a.k0 + 1, ni = a.i1 - a.i0 + 1;
    int nplanes = ni + nj + nk - 2;
    dim3 g((nj + 63) / 64, nk), b(64);
    for (int s = 0; s < nplanes; s++) {
        if (z.typ == 2) hipLaunchKernelGGL(k_lu_lower_plane<true>, g, b, 0, st, z, a, s);
        else hipLaunchKernelGGL(k_lu_lower_plane<false>, g, b, 0, st, z, a, s);
    }
    for (int s = 0; s < nplanes; s++) {
        if (z.typ == 2) hipLaunchKernelGGL(k_lu_upper_plane<true>, g, b, 0, st, z, a, s);
        else hipLaunchKernelGGL(k_lu_upper_plane<false>, g, b, 0, st, z, a, s);
    }
}

void launch_bc_face(const KZ& z, double* rop, const BcArgs& a, hipStream_t st) {
    int n0 = a.w[1] - a.w[0] + 1, n1 = a.w[3] - a.w[2] + 1, n2 = a.w[5] - a.w[4] + 1;
    if (n0 <= 0 || n1 <= 0 || n2 <= 0) return;
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

void launch_norms(const KZ& z, const double* dr, double* part, int nblk, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_norm_partial, dim3(nblk), dim3(256), 0, st, z, dr, part);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(64), 0, st, part, nblk, out);
}

void launch_tke(const KZ& z, const double* ro, double* part, int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_tke_partial, dim3(nblk), dim3(256), 0, st, z, ro, part);
}

}  // namespace fm
