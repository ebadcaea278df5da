// Device-side numerics of the FastS structured NS iteration for CDNA4 (gfx950).
//
// Every function reproduces the per-cell floating-point operation order of the reference
// kernel it cites, so that with floating-point contraction disabled the results are bit-for-bit
// those of the reference FastS kernels.  Layout: structure of arrays, i fastest, ghost layers
// included (Fast/FastS/FastS/formule_param.h: inddm).
#pragma once
#include <hip/hip_runtime.h>

namespace fm {

// Everything a kernel needs about one zone, passed by value as a kernel argument.
struct KZ {
    int ni, nj, nk;        // interior cells (IJKV)
    int nit, njt, nkt;     // storage cells per direction (NIJK)
    int nf, nfk;           // ghost layers
    long nd;               // NDIMDX
    long incj, inck;
    int typ;               // ITYPZONE: 0 3dfull, 2 3dcart
    long mi, mj, mk;       // metric strides (NIJK_MTR)
    int mf, mfk;
    long ndm;              // NDIMDX_MTR
    int xi, xj, xf, xfk;   // vertex grid (NIJK_XYZ)
    int kflu, slope, iflow, itypcp, dtloc, rk, nssiter, sensortype;
    // cache blocks of the reference's flux and gradient passes in j and k (CACHEBLCKJ/K and
    // the block counts max(1, n / size) of crsdm_scater.for; one block in i): they fix the
    // order in which a cell receives the fluxes of its max faces
    int cbj, cbk, nbj, nbk;
    double dtc, gam, cv, roinf, pinf, vinf, tinf, prandt, xmul0, temp0, cs, cfl, psiroe, wig_amp;
    // fields
    const double* x; const double* y; const double* z;
    double* ti; double* tj; double* tk; double* vol;
    double* xmut; double* drodm; double* coe; double* wig;
};

__device__ __forceinline__ long IDX(const KZ& z, int i, int j, int k) {
    return (long)(i + z.nf - 1) + (long)(j + z.nf - 1) * z.incj + (long)(k + z.nfk - 1) * z.inck;
}
__device__ __forceinline__ long MTR(const KZ& z, int i, int j, int k) {
    return (long)(i + z.mf - 1) * z.mi + (long)(j + z.mf - 1) * z.mj + (long)(k + z.mfk - 1) * z.mk;
}
__device__ __forceinline__ long XYZ(const KZ& z, int i, int j, int k) {
    return (long)(i + z.xf - 1) + (long)(j + z.xf - 1) * z.xi + (long)(k + z.xfk - 1) * (long)z.xi * z.xj;
}
// Fortran MAX/MIN as compiled by the reference: a > b ? a : b
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dsign(double a, double b) { return copysign(fabs(a), b); }

// Constants of the flux template header (template_FluxAndBalance.for:100-190)
struct FluxConst {
    double gam, rgp, gam1, gam3, psiroe, tmin_1, c1, c2, c3, opt0, c4, c5, c6, cvisq, wig_cte;
    double tc[3], s[3], volinv;   // 3dcart metric
};

__device__ __forceinline__ FluxConst flux_const(const KZ& z) {
    FluxConst f;
    f.gam = z.gam;
    f.rgp = z.cv * (f.gam - 1.);
    f.gam1 = f.gam / (f.gam - 1.);
    f.gam3 = f.gam1 / z.prandt * f.rgp;
    f.psiroe = z.psiroe;
    f.tmin_1 = 100. / z.tinf;
    f.c1 = 0.02 * z.vinf;
    f.c2 = 0.02 / (z.vinf * z.roinf);
    f.c3 = -2.;
    f.opt0 = (double)z.sensortype;
    f.wig_cte = z.wig_amp;
    f.c6 = 1. / 6.; f.c4 = 5. * f.c6; f.c5 = 2. * f.c6; f.c6 = -1. * f.c6;
    f.cvisq = 1. / 3;
    if (z.typ == 2) {
        f.tc[0] = z.ti[0]; f.tc[1] = z.tj[0]; f.tc[2] = z.tk[0];
        f.s[0] = fabs(f.tc[0]); f.s[1] = fabs(f.tc[1]); f.s[2] = fabs(f.tc[2]);
        f.volinv = 0.5 / z.vol[0];
    } else {
        f.tc[0] = f.tc[1] = f.tc[2] = 0.; f.s[0] = f.s[1] = f.s[2] = 0.; f.volinv = 0.;
    }
    return f;
}

// MUSCL states (Slope/{o1,o3,minmod}_slope_var.for); qm right, qp left of face (l-inc, l)
template <int SLOPE>
__device__ __forceinline__ void slope_state(const FluxConst& f, const double* __restrict__ rop,
                                            long nd, long l, long inc, double qm[5], double qp[5]) {
    long nm = l - inc, nm2 = l - 2 * inc, np = l + inc;
    if (SLOPE == 1) {
#pragma unroll
        for (int e = 0; e < 5; e++) { qm[e] = rop[l + e * nd]; qp[e] = rop[nm + e * nd]; }
    } else if (SLOPE == 2) {
#pragma unroll
        for (int e = 0; e < 5; e++) {
            const double* q = rop + e * nd;
            double a = q[l], b = q[nm], c = q[np], d = q[nm2];
            qm[e] = f.c4 * a + f.c5 * b + f.c6 * c;
            qp[e] = f.c4 * b + f.c6 * d + f.c5 * a;
        }
    } else {
        double roff = .5 * dmax(0., dmin(rop[l + 4 * nd] * f.tmin_1 - .5, 1.));
#pragma unroll
        for (int e = 0; e < 5; e++) {
            const double* q = rop + e * nd;
            double a = q[l], b = q[nm], c = q[np], d = q[nm2];
            double delm = a - b, delp = c - a, delq = b - d;
            double sg = dsign(1., delm);
            double slp = roff * (sg * dmax(0., dmin(fabs(delm), sg * delp)));
            sg = dsign(1., delq);
            double slq = roff * (sg * dmax(0., dmin(fabs(delq), sg * delm)));
            qm[e] = a - slp;
            qp[e] = b + slq;
        }
    }
}

// Inviscid face flux: Roe (etat_roe_GD.for + Vit_ent/fludiffer_*), AUSM+(P)
// (AUSM/fluFaceEuler.for) and SENSEUR (SENSOR/fluFaceEuler.for), 3dcart or 3dfull.
template <int KFLU, bool CART, int D>
__device__ __forceinline__ void inviscid_flux(const FluxConst& f, const double* __restrict__ rop,
                                              const double* __restrict__ wig, long nd, long l,
                                              const double tn[3], double sn, const double qm[5],
                                              const double qp[5], double flu[5]) {
    if (KFLU == 5) {
        double gam = f.gam, rgp = f.rgp, gam1 = f.gam1, psiroe = f.psiroe;
        double r1 = qp[0];
        double p1 = r1 * qp[4] * rgp;
        double h1 = gam1 * qp[4] * rgp + .5 * (qp[1] * qp[1] + qp[2] * qp[2] + qp[3] * qp[3]);
        double r2 = qm[0];
        double p2 = r2 * qm[4] * rgp;
        double h2 = gam1 * qm[4] * rgp + .5 * (qm[1] * qm[1] + qm[2] * qm[2] + qm[3] * qm[3]);
        double sp1 = sqrt(r1), sm1 = sqrt(r2);
        double r = sp1 * sm1;
        double r_1 = 1. / (sp1 + sm1);
        double a[3];
        a[0] = (sp1 * qp[1] + sm1 * qm[1]) * r_1;
        a[1] = (sp1 * qp[2] + sm1 * qm[2]) * r_1;
        a[2] = (sp1 * qp[3] + sm1 * qm[3]) * r_1;
        double h = (sp1 * h1 + sm1 * h2) * r_1;
        double q = .5 * (a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        double c = sqrt((gam - 1.) * fabs(h - q));
        double qn1, qn2;
        if (CART) { qn1 = qp[1 + D] * tn[D]; qn2 = qm[1 + D] * tn[D]; }
        else {
            qn1 = qp[1] * tn[0] + qp[2] * tn[1] + qp[3] * tn[2];
            qn2 = qm[1] * tn[0] + qm[2] * tn[1] + qm[3] * tn[2];
        }
        double du[3] = {(qm[1] - qp[1]) * sn, (qm[2] - qp[2]) * sn, (qm[3] - qp[3]) * sn};
        double dp = p2 - p1;
        double dqn = qn2 - qn1;
        double s_1 = 1. / dmax(sn, 1.e-30);
        double nn[3], qn;
        if (CART) { nn[0] = nn[1] = nn[2] = 0.; nn[D] = tn[D] * s_1; qn = a[D] * nn[D]; }
        else {
            nn[0] = tn[0] * s_1; nn[1] = tn[1] * s_1; nn[2] = tn[2] * s_1;
            qn = a[0] * nn[0] + a[1] * nn[1] + a[2] * nn[2];
        }
        double f1 = qn;
        double f4 = psiroe * (fabs(f1) + c);
        double f2 = .5 + dsign(.5, fabs(f1) - f4);
        double f3 = (f2 - 1.) * (f1 * f1 / f4 + f4) * 0.5 - f2 * fabs(f1);
        f1 = (r2 - r1 - dp / (c * c)) * sn;
        flu[0] = f3 * (f1);
#pragma unroll
        for (int m = 0; m < 3; m++) {
            bool dir = CART ? (m == D) : true;
            flu[1 + m] = f3 * (f1 * a[m] + r * (dir ? du[m] - nn[m] * dqn : du[m]));
        }
        flu[4] = f3 * (f1 * q + r * (a[0] * du[0] + a[1] * du[1] + a[2] * du[2] - qn * dqn));
        f1 = qn + c;
        f2 = .5 + dsign(.5, fabs(f1) - f4);
        f3 = (f2 - 1.) * (f1 * f1 / f4 + f4) * 0.5 - f2 * fabs(f1);
        f1 = (dp + r * c * dqn * s_1) / (2. * c * c) * sn * f3;
        flu[0] = flu[0] + f1;
#pragma unroll
        for (int m = 0; m < 3; m++) {
            bool dir = CART ? (m == D) : true;
            flu[1 + m] = flu[1 + m] + f1 * (dir ? a[m] + nn[m] * c : a[m]);
        }
        flu[4] = flu[4] + f1 * (h + qn * c);
        f1 = qn - c;
        f2 = .5 + dsign(.5, fabs(f1) - f4);
        f3 = (f2 - 1.) * (f1 * f1 / f4 + f4) * 0.5 - f2 * fabs(f1);
        f1 = (dp - r * c * dqn * s_1) / (2. * c * c) * f3 * sn;
        flu[0] = flu[0] + f1;
#pragma unroll
        for (int m = 0; m < 3; m++) {
            bool dir = CART ? (m == D) : true;
            flu[1 + m] = flu[1 + m] + f1 * (dir ? a[m] - nn[m] * c : a[m]);
        }
        flu[4] = flu[4] + f1 * (h - qn * c);
        double rou1 = r1 * qn1, rou2 = r2 * qn2, p1p2 = p1 + p2;
        flu[0] = (flu[0] + rou1 + rou2) * 0.5;
#pragma unroll
        for (int m = 0; m < 3; m++) {
            bool dir = CART ? (m == D) : true;
            flu[1 + m] = dir ? (flu[1 + m] + rou1 * qp[1 + m] + rou2 * qm[1 + m] + tn[m] * p1p2) * 0.5
                             : (flu[1 + m] + rou1 * qp[1 + m] + rou2 * qm[1 + m]) * 0.5;
        }
        flu[4] = (flu[4] + rou1 * h1 + rou2 * h2) * 0.5;
    } else {
        double rgp = f.rgp, gam1 = f.gam1;
        double r1 = qp[0];
        double rou1 = r1 * qp[1], rov1 = r1 * qp[2], row1 = r1 * qp[3];
        double p1 = r1 * qp[4] * rgp;
        double h1 = gam1 * p1 + .5 * (rou1 * qp[1] + rov1 * qp[2] + row1 * qp[3]);
        double r2 = qm[0];
        double rou2 = r2 * qm[1], rov2 = r2 * qm[2], row2 = r2 * qm[3];
        double p2 = r2 * qm[4] * rgp;
        double h2 = gam1 * p2 + .5 * (rou2 * qm[1] + rov2 * qm[2] + row2 * qm[3]);
        double qn1, qn2;
        if (CART) { qn1 = qp[1 + D] * tn[D]; qn2 = qm[1 + D] * tn[D]; }
        else {
            qn1 = qp[1] * tn[0] + qp[2] * tn[1] + qp[3] * tn[2];
            qn2 = qm[1] * tn[0] + qm[2] * tn[1] + qm[3] * tn[2];
        }
        double u, tdu;
        if (KFLU == 1) {
            double c = rgp * f.gam * rop[l + 4 * nd];
            double son = sqrt(qn1 * qn1 / c);
            double tam = f.c3 * son + sn;
            double tam1 = dmax(0., tam) * f.c2;
            u = 0.25 * (qn1 + qn2) - tam1 * (p2 - p1);
            tdu = dmax(fabs(u), f.c1 * sn) * f.wig_cte;
        } else {
            double wg = wig[l + D * nd];
            u = 0.25 * (qn1 + qn2) - f.c2 * sn * (p2 - p1) * (f.opt0 * wg + 1. - f.opt0);
            tdu = dmax(fabs(u), f.c1 * sn) * wg;
        }
        double p1p2 = (p1 + p2) * 0.5;
        double mom1[3] = {rou1, rov1, row1}, mom2[3] = {rou2, rov2, row2};
        flu[0] = u * (r1 + r2) - tdu * (r2 - r1);
#pragma unroll
        for (int m = 0; m < 3; m++) {
            bool dir = CART ? (m == D) : true;
            flu[1 + m] = dir ? u * (mom1[m] + mom2[m]) - tdu * (mom2[m] - mom1[m]) + tn[m] * p1p2
                             : u * (mom1[m] + mom2[m]) - tdu * (mom2[m] - mom1[m]);
        }
        flu[4] = u * (h1 + h2) - tdu * (h2 - h1);
    }
}

// Centred viscous flux on a Cartesian zone (fluvisq_3dcart_{i,j,k}.for)
template <int D>
// gv (optional): the face gradients (u, v, w, T) x (x, y, z) the flux was built from, and the
// face's volinv in gv[12]; components this direction does not form stay as they were
__device__ __forceinline__ void visc_cart(const FluxConst& f, const KZ& z, const double* __restrict__ rop,
                                          const double* __restrict__ xmut, long l, double flu[5],
                                          double* gv = nullptr) {
    long nd = z.nd, incj = z.incj, inck = z.inck;
    const double* U = rop + nd; const double* V = rop + 2 * nd;
    const double* W = rop + 3 * nd; const double* T = rop + 4 * nd;
    double tcx = f.tc[0], tcy = f.tc[1], tcz = f.tc[2];
    double xmutvol, xktvol, fv, fv5, f1, f2, f3, f4, f5, f6, div, fvs;
    if (D == 0) {
        long il = l - 1, ir = l;
        long l220 = l - 1 - incj, l020 = l - incj, l210 = l - 1 + incj, l010 = l + incj;
        long l202 = l - 1 - inck, l002 = l - inck, l201 = l - 1 + inck, l001 = l + inck;
        double tix = tcx, tjy = .25 * tcy, tkz = .25 * tcz;
        double gUx, gUy, gUz, gVx, gVy, gWx, gWz, gTx;
        f1 = U[il]; f2 = U[ir]; fvs = f1 + f2;
        f3 = fvs + U[l220] + U[l020]; f4 = fvs + U[l210] + U[l010];
        f5 = fvs + U[l202] + U[l002]; f6 = fvs + U[l201] + U[l001];
        gUx = (f2 - f1) * tix; gUy = (f4 - f3) * tjy; gUz = (f6 - f5) * tkz;
        f1 = V[il]; f2 = V[ir]; fvs = f1 + f2;
        f3 = fvs + V[l220] + V[l020]; f4 = fvs + V[l210] + V[l010];
        gVx = (f2 - f1) * tix; gVy = (f4 - f3) * tjy;
        f1 = W[il]; f2 = W[ir]; fvs = f1 + f2;
        f5 = fvs + W[l202] + W[l002]; f6 = fvs + W[l201] + W[l001];
        gWx = (f2 - f1) * tix; gWz = (f6 - f5) * tkz;
        f1 = T[il]; f2 = T[ir];
        gTx = (f2 - f1) * tix;
        div = (gUx + gVy + gWz) * f.cvisq;
        f1 = gUx - div; f2 = gUy + gVx; f3 = gUz + gWx;
        xmutvol = (xmut[ir] + xmut[il]) * f.volinv;
        xktvol = xmutvol * f.gam3;
        fv = -(2. * f1 * tcx) * xmutvol; fv5 = fv * (U[ir] + U[il]); flu[1] = flu[1] + fv;
        fv = -(f2 * tcx) * xmutvol; fv5 = fv5 + fv * (V[ir] + V[il]); flu[2] = flu[2] + fv;
        fv = -(f3 * tcx) * xmutvol; fv5 = fv5 + fv * (W[ir] + W[il]); flu[3] = flu[3] + fv;
        fv = fv5 * 0.5 - (gTx * tcx) * xktvol; flu[4] = flu[4] + fv;
        if (gv) {
            gv[0] = gUx; gv[1] = gUy; gv[2] = gUz; gv[3] = gVx; gv[4] = gVy;
            gv[6] = gWx; gv[8] = gWz; gv[9] = gTx;
        }
    } else if (D == 1) {
        long il = l - incj, ir = l;
        long l022 = l - incj - inck, l002 = l - inck, l021 = l - incj + inck, l001 = l + inck;
        long l220 = l - 1 - incj, l200 = l - 1, l120 = l + 1 - incj, l100 = l + 1;
        double tiy = tcy, tjz = .25 * tcz, tkx = .25 * tcx;
        double gUx, gUy, gVx, gVy, gVz, gWy, gWz, gTy;
        f1 = U[il]; f2 = U[ir]; fvs = f1 + f2;
        f5 = fvs + U[l220] + U[l200]; f6 = fvs + U[l120] + U[l100];
        gUx = (f6 - f5) * tkx; gUy = (f2 - f1) * tiy;
        f1 = V[il]; f2 = V[ir]; fvs = f1 + f2;
        f3 = fvs + V[l022] + V[l002]; f4 = fvs + V[l021] + V[l001];
        f5 = fvs + V[l220] + V[l200]; f6 = fvs + V[l120] + V[l100];
        gVx = (f6 - f5) * tkx; gVy = (f2 - f1) * tiy; gVz = (f4 - f3) * tjz;
        f1 = W[il]; f2 = W[ir]; fvs = f1 + f2;
        f3 = fvs + W[l022] + W[l002]; f4 = fvs + W[l021] + W[l001];
        gWy = (f2 - f1) * tiy; gWz = (f4 - f3) * tjz;
        f1 = T[il]; f2 = T[ir];
        gTy = (f2 - f1) * tiy;
        div = (gUx + gVy + gWz) * f.cvisq;
        f2 = gUy + gVx; f4 = gVy - div; f5 = gVz + gWy;
        xmutvol = (xmut[ir] + xmut[il]) * f.volinv;
        xktvol = xmutvol * f.gam3;
        fv = -(f2 * tcy) * xmutvol; fv5 = fv * (U[ir] + U[il]); flu[1] = flu[1] + fv;
        fv = -(2. * f4 * tcy) * xmutvol; fv5 = fv5 + fv * (V[ir] + V[il]); flu[2] = flu[2] + fv;
        fv = -(f5 * tcy) * xmutvol; fv5 = fv5 + fv * (W[ir] + W[il]); flu[3] = flu[3] + fv;
        fv = fv5 * 0.5 - (gTy * tcy) * xktvol; flu[4] = flu[4] + fv;
        if (gv) {
            gv[0] = gUx; gv[1] = gUy; gv[3] = gVx; gv[4] = gVy; gv[5] = gVz;
            gv[7] = gWy; gv[8] = gWz; gv[10] = gTy;
        }
    } else {
        long il = l - inck, ir = l;
        long l200 = l - 1, l202 = l - 1 - inck, l102 = l + 1 - inck, l100 = l + 1;
        long l022 = l - incj - inck, l020 = l - incj, l012 = l + incj - inck, l010 = l + incj;
        double tiz = tcz, tky = .25 * tcy, tjx = .25 * tcx;
        double gUx, gUz, gVy, gVz, gWx, gWy, gWz, gTz;
        f1 = U[il]; f2 = U[ir]; fvs = f1 + f2;
        f3 = fvs + U[l202] + U[l200]; f4 = fvs + U[l102] + U[l100];
        gUx = (f4 - f3) * tjx; gUz = (f2 - f1) * tiz;
        f1 = V[il]; f2 = V[ir]; fvs = f1 + f2;
        f5 = fvs + V[l022] + V[l020]; f6 = fvs + V[l012] + V[l010];
        gVy = (f6 - f5) * tky; gVz = (f2 - f1) * tiz;
        f1 = W[il]; f2 = W[ir]; fvs = f1 + f2;
        f3 = fvs + W[l202] + W[l200]; f4 = fvs + W[l102] + W[l100];
        f5 = fvs + W[l022] + W[l020]; f6 = fvs + W[l012] + W[l010];
        gWx = (f4 - f3) * tjx; gWy = (f6 - f5) * tky; gWz = (f2 - f1) * tiz;
        f1 = T[il]; f2 = T[ir];
        gTz = (f2 - f1) * tiz;
        div = (gUx + gVy + gWz) * f.cvisq;
        f3 = gUz + gWx; f5 = gVz + gWy; f6 = gWz - div;
        xmutvol = (xmut[ir] + xmut[il]) * f.volinv;
        xktvol = xmutvol * f.gam3;
        fv = -f3 * tcz * xmutvol; fv5 = fv * (U[ir] + U[il]); flu[1] = flu[1] + fv;
        fv = -f5 * tcz * xmutvol; fv5 = fv5 + fv * (V[ir] + V[il]); flu[2] = flu[2] + fv;
        fv = -2. * f6 * tcz * xmutvol; fv5 = fv5 + fv * (W[ir] + W[il]); flu[3] = flu[3] + fv;
        fv = fv5 * 0.5 - gTz * tcz * xktvol; flu[4] = flu[4] + fv;
        if (gv) {
            gv[0] = gUx; gv[2] = gUz; gv[4] = gVy; gv[5] = gVz;
            gv[6] = gWx; gv[7] = gWy; gv[8] = gWz; gv[11] = gTz;
        }
    }
    if (gv) gv[12] = f.volinv;
}

// Centred viscous flux on a curvilinear zone (fluvisq_3dfull_{i,j,k}.for)
template <int D>
__device__ __forceinline__ void visc_full(const FluxConst& f, const KZ& z, const double* __restrict__ rop,
                                          const double* __restrict__ xmut, long l, long lt,
                                          const double tn[3], double flu[5], double* gv = nullptr) {
    long nd = z.nd, ndm = z.ndm, incj = z.incj, inck = z.inck, mi = z.mi, mj = z.mj, mk = z.mk;
    const double* A; const double* B; const double* C;
    long il, lvol, a0, a1, b0, b1, b2, b3, c0, c1, c2, c3, n3a, n3b, n4a, n4b, n5a, n5b, n6a, n6b;
    if (D == 0) {
        A = z.ti; B = z.tj; C = z.tk; il = l - 1; lvol = lt - mi;
        a0 = lt - mi; a1 = lt + mi; b0 = lt; b1 = lt - mi; b2 = lt + mj; b3 = lt - mi + mj;
        c0 = lt; c1 = lt - mi; c2 = lt + mk; c3 = lt - mi + mk;
        n3a = l - 1 - incj; n3b = l - incj; n4a = l - 1 + incj; n4b = l + incj;
        n5a = l - 1 - inck; n5b = l - inck; n6a = l - 1 + inck; n6b = l + inck;
    } else if (D == 1) {
        A = z.tj; B = z.tk; C = z.ti; il = l - incj; lvol = lt - mj;
        a0 = lt - mj; a1 = lt + mj; b0 = lt; b1 = lt - mj; b2 = lt + mk; b3 = lt - mj + mk;
        c0 = lt; c1 = lt - mj; c2 = lt + mi; c3 = lt + mi - mj;
        n3a = l - incj - inck; n3b = l - inck; n4a = l - incj + inck; n4b = l + inck;
        n5a = l - 1 - incj; n5b = l - 1; n6a = l + 1 - incj; n6b = l + 1;
    } else {
        A = z.tk; B = z.ti; C = z.tj; il = l - inck; lvol = lt - mk;
        a0 = lt - mk; a1 = lt + mk; b0 = lt; b1 = lt - mk; b2 = lt + mi; b3 = lt + mi - mk;
        c0 = lt; c1 = lt - mk; c2 = lt + mj; c3 = lt + mj - mk;
        n3a = l - 1 - inck; n3b = l - 1; n4a = l + 1 - inck; n4b = l + 1;
        n5a = l - incj - inck; n5b = l - incj; n6a = l + incj - inck; n6b = l + incj;
    }
    long ir = l;
    double t0[3], t1[3], u0[3], u1[3], w0[3], w1[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        t0[m] = .5 * (A[lt + m * ndm] + A[a0 + m * ndm]);
        t1[m] = .5 * (A[lt + m * ndm] + A[a1 + m * ndm]);
        u0[m] = .125 * (B[b0 + m * ndm] + B[b1 + m * ndm]);
        u1[m] = .125 * (B[b2 + m * ndm] + B[b3 + m * ndm]);
        w0[m] = .125 * (C[c0 + m * ndm] + C[c1 + m * ndm]);
        w1[m] = .125 * (C[c2 + m * ndm] + C[c3 + m * ndm]);
    }
    double g[4][3];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const double* q = rop + (e + 1) * nd;
        double f1 = q[il], f2 = q[ir];
        double f3 = q[ir] + q[il] + q[n3a] + q[n3b];
        double f4 = q[ir] + q[il] + q[n4a] + q[n4b];
        double f5 = q[ir] + q[il] + q[n5a] + q[n5b];
        double f6 = q[ir] + q[il] + q[n6a] + q[n6b];
#pragma unroll
        for (int m = 0; m < 3; m++)
            g[e][m] = (f2 * t1[m] - f1 * t0[m]) + (f4 * u1[m] - f3 * u0[m]) + (f6 * w1[m] - f5 * w0[m]);
    }
    double div = (g[0][0] + g[1][1] + g[2][2]) * f.cvisq;
    double f1 = g[0][0] - div, f2 = g[0][1] + g[1][0], f3 = g[0][2] + g[2][0];
    double f4 = g[1][1] - div, f5 = g[1][2] + g[2][1], f6 = g[2][2] - div;
    double volinv = 1. / (z.vol[lt] + z.vol[lvol]);
    double xmutvol = (xmut[ir] + xmut[il]) * volinv;
    double xktvol = xmutvol * f.gam3;
    const double* U = rop + nd; const double* V = rop + 2 * nd; const double* W = rop + 3 * nd;
    double tcx = tn[0], tcy = tn[1], tcz = tn[2], fv, fv5;
    fv = -(2. * f1 * tcx + f2 * tcy + f3 * tcz) * xmutvol;
    fv5 = fv * (U[ir] + U[il]); flu[1] = flu[1] + fv;
    fv = -(f2 * tcx + 2. * f4 * tcy + f5 * tcz) * xmutvol;
    fv5 = fv5 + fv * (V[ir] + V[il]); flu[2] = flu[2] + fv;
    fv = -(f3 * tcx + f5 * tcy + 2. * f6 * tcz) * xmutvol;
    fv5 = fv5 + fv * (W[ir] + W[il]); flu[3] = flu[3] + fv;
    fv = fv5 * 0.5 - (g[3][0] * tcx + g[3][1] * tcy + g[3][2] * tcz) * xktvol;
    flu[4] = flu[4] + fv;
    if (gv) {
#pragma unroll
        for (int e = 0; e < 4; e++)
#pragma unroll
            for (int m = 0; m < 3; m++) gv[3 * e + m] = g[e][m];
        gv[12] = volinv;
    }
}

// Full face flux through the face (l - inc_D, l) of cell (i,j,k)
template <int KFLU, int SLOPE, bool LAMIN, bool CART, int D>
__device__ __forceinline__ void face_flux(const FluxConst& f, const KZ& z, const double* __restrict__ rop,
                                          int i, int j, int k, double flu[5]) {
    long l = IDX(z, i, j, k);
    long inc = D == 0 ? 1 : D == 1 ? z.incj : z.inck;
    double qm[5], qp[5], tn[3], sn;
    long lt = 0;
    if (CART) {
        tn[0] = f.tc[0]; tn[1] = f.tc[1]; tn[2] = f.tc[2]; sn = f.s[D];
    } else {
        const double* t = D == 0 ? z.ti : D == 1 ? z.tj : z.tk;
        lt = MTR(z, i, j, k);
        tn[0] = t[lt]; tn[1] = t[lt + z.ndm]; tn[2] = t[lt + 2 * z.ndm];
        sn = sqrt(tn[0] * tn[0] + tn[1] * tn[1] + tn[2] * tn[2]);
    }
    slope_state<SLOPE>(f, rop, z.nd, l, inc, qm, qp);
    inviscid_flux<KFLU, CART, D>(f, rop, z.wig, z.nd, l, tn, sn, qm, qp, flu);
    if (LAMIN) {
        if (CART) visc_cart<D>(f, z, rop, z.xmut, l, flu);
        else visc_full<D>(f, z, rop, z.xmut, l, lt, tn, flu);
    }
}

}  // namespace fm
