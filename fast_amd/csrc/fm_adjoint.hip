// Adjoint pieces, for 3-D zones:
// * gradients of the wall pressure-force coefficients with respect to the state
//   (compute_dpJ_dpW.cpp:33-261 -> dpj_dpw_calc_meth.for -> dpJ_dpW_fluausm_select.for ->
//   dpJ_dpW_fluausm_euler_o3_2d.for);
// * the transposed Jacobian product of the AUSM+(P) o3 Euler face balance (the reverse mode of
//   fluausm_euler_o3_2d_b.for).
//
// Every floating-point operation follows the order of the oracle's fo_dpj_dpw and
// fo_residual_vjp, compiled without contraction, so that the two agree bit for bit.
#include "fm_device.hpp"
#include "fm_kernels.hpp"

namespace fm {

// primitive gradient g of one cell (rho, u, v, w, T) -> gradient with respect to the
// conservative state (rho, rho u, rho v, rho w, rho E) (dpP_dpW_fluausm_euler_2d_idir_min.for)
__device__ __forceinline__ void dpj_cons(const double q[5], double cvm, const double g[5], double out[5]) {
    double w1 = q[0], w2 = q[0] * q[1], w3 = q[0] * q[2], w4 = q[0] * q[3];
    double ke = 0.5 * (w2 * w2 + w3 * w3 + w4 * w4);
    double w5 = cvm * w1 * q[4] + ke / w1;
    double w11 = w1 * w1;
    double t1 = 1. / cvm * (-w5 / w11 + 2. * ke / (w11 * w1));
    out[0] = g[0] + g[1] * (-w2 / w11) + g[2] * (-w3 / w11) + g[3] * (-w4 / w11) + g[4] * t1;
    out[1] = g[1] * (1. / w1) + g[4] * (-1. / cvm * w2 / w11);
    out[2] = g[2] * (1. / w1) + g[4] * (-1. / cvm * w3 / w11);
    out[3] = g[3] * (1. / w1) + g[4] * (-1. / cvm * w4 / w11);
    out[4] = g[4] * (1. / (cvm * w1));
}

// Reverse pass of the AUSM+(P) o3 momentum flux of one face (AUSM fluFaceEuler.for, the
// flux inviscid_flux<1> forms): q[s] = primitive states of the stencil (nm2, nm, l, np),
// fb = weights of (F_x, F_y, F_z); wb[s] += d(fb . F)/dq[s], the sound speed c^2 = rgp gam T_l
// included (dpJ_dpW_fluausm_euler_o3_2d.for:468-614).
__device__ __forceinline__ void dpj_face(const FluxConst& f, const double q[4][5], const double tn[3],
                                         double sn, const double fb[3], double wb[4][5]) {
    double qp[5], qm[5];
#pragma unroll
    for (int e = 0; e < 5; e++) {
        qm[e] = f.c4 * q[2][e] + f.c5 * q[1][e] + f.c6 * q[3][e];
        qp[e] = f.c4 * q[1][e] + f.c6 * q[0][e] + f.c5 * q[2][e];
    }
    double rgp = f.rgp, r1 = qp[0], r2 = qm[0];
    double p1 = r1 * qp[4] * rgp, p2 = r2 * qm[4] * rgp;
    double qn1 = qp[1] * tn[0] + qp[2] * tn[1] + qp[3] * tn[2];
    double qn2 = qm[1] * tn[0] + qm[2] * tn[1] + qm[3] * tn[2];
    double c = rgp * f.gam * q[2][4];
    double son = sqrt(qn1 * qn1 / c);
    double tam = f.c3 * son + sn;
    double tam1 = dmax(0., tam) * f.c2;
    double u = 0.25 * (qn1 + qn2) - tam1 * (p2 - p1);
    double tdu = dmax(fabs(u), f.c1 * sn) * f.wig_cte;
    double ub = 0., tdub = 0., p1p2b = 0., m1b[3], m2b[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        double mom1 = r1 * qp[1 + m], mom2 = r2 * qm[1 + m];
        ub = ub + fb[m] * (mom1 + mom2);
        tdub = tdub + fb[m] * (mom1 - mom2);
        m1b[m] = fb[m] * (u + tdu);
        m2b[m] = fb[m] * (u - tdu);
        p1p2b = p1p2b + fb[m] * tn[m];
    }
    double p1b = 0.5 * p1p2b, p2b = 0.5 * p1p2b;
    if (fabs(u) >= f.c1 * sn) ub = ub + tdub * f.wig_cte * (u >= 0. ? 1. : -1.);
    double qn1b = 0.25 * ub, qn2b = 0.25 * ub;
    double tam1b = -ub * (p2 - p1);
    p1b = p1b + ub * tam1;
    p2b = p2b - ub * tam1;
    double tamb = tam >= 0. ? tam1b * f.c2 : 0.;
    double sonb = f.c3 * tamb;
    double sc = sqrt(c);
    qn1b = qn1b + sonb * (qn1 >= 0. ? 1. / sc : -1. / sc);
    double cb = sonb * (-0.5 * sqrt(qn1 * qn1) / sqrt(c * c * c));
    double qpb[5], qmb[5];
    qpb[0] = p1b * qp[4] * rgp; qmb[0] = p2b * qm[4] * rgp;
#pragma unroll
    for (int m = 0; m < 3; m++) {
        qpb[0] = qpb[0] + m1b[m] * qp[1 + m];
        qmb[0] = qmb[0] + m2b[m] * qm[1 + m];
        qpb[1 + m] = qn1b * tn[m] + m1b[m] * r1;
        qmb[1 + m] = qn2b * tn[m] + m2b[m] * r2;
    }
    qpb[4] = p1b * r1 * rgp; qmb[4] = p2b * r2 * rgp;
#pragma unroll
    for (int e = 0; e < 5; e++) {
        wb[0][e] = wb[0][e] + f.c6 * qpb[e];
        wb[1][e] = wb[1][e] + (f.c4 * qpb[e] + f.c5 * qmb[e]);
        wb[2][e] = wb[2][e] + (f.c5 * qpb[e] + f.c4 * qmb[e]);
        wb[3][e] = wb[3][e] + f.c6 * qmb[e];
    }
    wb[2][4] = wb[2][4] + cb * (rgp * f.gam);
}

// ghost gradient g carried onto the interior cell it mirrors (bvbs_wall_inviscid.for: density
// and temperature copied, velocity reflected about the plane of unit normal n): x += M^T g
__device__ __forceinline__ void dpj_mirror(const double n[3], const double g[5], double x[5]) {
    x[0] = x[0] + g[0];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double s = 0.;
#pragma unroll
        for (int b = 0; b < 3; b++) s = s + g[1 + b] * ((a == b ? 1. : 0.) - 2. * n[b] * n[a]);
        x[1 + a] = x[1 + a] + s;
    }
    x[4] = x[4] + g[4];
}

// One thread per face of the window: J_D = s (cos a F_x + sin a F_y), J_L = s (cos a F_y -
// sin a F_x), s = sens * surfinv (dpJ_dpW_fluausm_euler_o3_2d.for:359-366, 855-890); the two
// interior cells of the stencil are assigned their conservative gradients (:958-978).
template <bool CART, int D>
__global__ void k_dpj_dpw(KZ z, DpjArgs a) {
    const int n0 = a.w[1] - a.w[0] + 1, n1 = a.w[3] - a.w[2] + 1, n2 = a.w[5] - a.w[4] + 1;
    const long nf = (long)n0 * n1 * n2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf) return;
    const int i = a.w[0] + (int)(t % n0), j = a.w[2] + (int)((t / n0) % n1), k = a.w[4] + (int)(t / ((long)n0 * n1));
    const FluxConst f = flux_const(z);
    const long l = IDX(z, i, j, k), nd = z.nd, ndm = z.ndm;
    const long inc = D == 0 ? 1 : D == 1 ? z.incj : z.inck;
    const bool up = (a.idir % 2) == 0;
    const double norm = 2. / (z.roinf * (z.vinf * z.vinf));
    const double sens = up ? norm : -norm;
    const double cc = sens * a.surfinv * a.cosa, cs = sens * a.surfinv * a.sina;
    const double cvm = f.rgp / (f.gam - 1.);
    const long cell[4] = {l - 2 * inc, l - inc, l, l + inc};
    double tn[3] = {0., 0., 0.}, sn;
    if (CART) {
        tn[D] = f.tc[D]; sn = f.s[D];
    } else {
        const double* tm = D == 0 ? z.ti : D == 1 ? z.tj : z.tk;
        const long lt = MTR(z, i, j, k);
        tn[0] = tm[lt]; tn[1] = tm[lt + ndm]; tn[2] = tm[lt + 2 * ndm];
        sn = sqrt(tn[0] * tn[0] + tn[1] * tn[1] + tn[2] * tn[2]);
    }
    const double nrm = dmax(sqrt(tn[0] * tn[0] + tn[1] * tn[1] + tn[2] * tn[2]), 1e-30);
    const double n[3] = {tn[0] / nrm, tn[1] / nrm, tn[2] / nrm};
    double q[4][5];
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
        for (int e = 0; e < 5; e++) q[s][e] = a.rop[cell[s] + e * nd];
    const int A = up ? 1 : 2, B = up ? 0 : 3, gA = up ? 2 : 1, gB = up ? 3 : 0;
#pragma unroll
    for (int o = 0; o < 2; o++) {
        const double fb[3] = {o == 0 ? cc : -cs, o == 0 ? cs : cc, 0.};
        double wb[4][5];
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int e = 0; e < 5; e++) wb[s][e] = 0.;
        dpj_face(f, q, tn, sn, fb, wb);
        dpj_mirror(n, wb[gA], wb[A]);
        dpj_mirror(n, wb[gB], wb[B]);
        double ga[5], gb[5];
        dpj_cons(q[A], cvm, wb[A], ga);
        dpj_cons(q[B], cvm, wb[B], gb);
        double* out = o == 0 ? a.dcd : a.dcl;
#pragma unroll
        for (int e = 0; e < 5; e++) {
            out[cell[A] + e * nd] = ga[e];
            out[cell[B] + e * nd] = gb[e];
        }
    }
}

template <bool CART>
static void dpj_d(const KZ& z, const DpjArgs& a, dim3 g, hipStream_t st) {
    int d = (a.idir - 1) / 2;
    if (d == 0) hipLaunchKernelGGL((k_dpj_dpw<CART, 0>), g, dim3(256), 0, st, z, a);
    else if (d == 1) hipLaunchKernelGGL((k_dpj_dpw<CART, 1>), g, dim3(256), 0, st, z, a);
    else hipLaunchKernelGGL((k_dpj_dpw<CART, 2>), g, dim3(256), 0, st, z, a);
}

int launch_dpj_dpw(const KZ& z, const DpjArgs& a, hipStream_t st) {
    if (z.kflu != 1 || z.slope != 2 || z.iflow != 1) return -1;     // option 213 only
    long nf = (long)(a.w[1] - a.w[0] + 1) * (a.w[3] - a.w[2] + 1) * (a.w[5] - a.w[4] + 1);
    if (nf <= 0) return 0;
    dim3 g((unsigned)((nf + 255) / 256));
    if (z.typ == 2) dpj_d<true>(z, a, g, st);
    else dpj_d<false>(z, a, g, st);
    return 0;
}

// ------------------------------------------------------------------------------------------
// Transposed residual Jacobian product (the reverse mode Tapenade gives fluausm_euler_o3_2d in
// ADJOINT/AUSM/2d/fluausm_euler_o3_2d_b.for: the gradient of the face balance drodm with
// respect to rop), for 3-D zones: out = (dR/dW)^T lambda, R the AUSM+(P) o3 Euler face balance
// of the interior cells (R_l = sum_d F_d(l) - F_d(l + inc_d)), W the primitive state of every
// cell, ghost cells included.
//
// Every face of direction D (face index 1..n_D + 1, interior across) is weighted by
// fb = lambda_l - lambda_{l - inc} (each term only where its cell is interior); its reverse pass
// gives its contributions to the four stencil cells, and each cell (index -1..n_D + 2 along D)
// adds those of the faces c - 1, c, c + 1, c + 2 in that order, then adds the sum to out.
// Directions run i, j, k.  Kernels:
// * i: one workgroup per (j, k) row, the row's face contributions staged in LDS;
// * j, k: one thread per line, marching up the line with the four-cell stencil and the four
//   partial sums in registers (a cell is complete, and written, once its last face is done);
// * rows too long for LDS: face contributions through a global buffer, then a gather pass.

// reverse pass of the whole AUSM+(P) o3 face flux (AUSM fluFaceEuler.for): fb = weights of the
// five flux components; wb[s] = d(fb . F)/dq[s] for the stencil (nm2, nm, l, np)
__device__ __forceinline__ void ausm_face_rev(const FluxConst& f, const double q[4][5], const double tn[3],
                                              double sn, const double fb[5], double wb[4][5]) {
    double qp[5], qm[5];
#pragma unroll
    for (int e = 0; e < 5; e++) {
        qm[e] = f.c4 * q[2][e] + f.c5 * q[1][e] + f.c6 * q[3][e];
        qp[e] = f.c4 * q[1][e] + f.c6 * q[0][e] + f.c5 * q[2][e];
    }
    double rgp = f.rgp, gam1 = f.gam1, r1 = qp[0], r2 = qm[0];
    double mom1[3], mom2[3];
#pragma unroll
    for (int m = 0; m < 3; m++) { mom1[m] = r1 * qp[1 + m]; mom2[m] = r2 * qm[1 + m]; }
    double p1 = r1 * qp[4] * rgp, p2 = r2 * qm[4] * rgp;
    double h1 = gam1 * p1 + .5 * (mom1[0] * qp[1] + mom1[1] * qp[2] + mom1[2] * qp[3]);
    double h2 = gam1 * p2 + .5 * (mom2[0] * qm[1] + mom2[1] * qm[2] + mom2[2] * qm[3]);
    double qn1 = qp[1] * tn[0] + qp[2] * tn[1] + qp[3] * tn[2];
    double qn2 = qm[1] * tn[0] + qm[2] * tn[1] + qm[3] * tn[2];
    double c = rgp * f.gam * q[2][4];
    double son = sqrt(qn1 * qn1 / c);
    double tam = f.c3 * son + sn;
    double tam1 = dmax(0., tam) * f.c2;
    double u = 0.25 * (qn1 + qn2) - tam1 * (p2 - p1);
    double tdu = dmax(fabs(u), f.c1 * sn) * f.wig_cte;
    // F0 = u (r1 + r2) - tdu (r2 - r1); F_{1+m} = u (mom1 + mom2) - tdu (mom2 - mom1) + tn_m p1p2;
    // F4 = u (h1 + h2) - tdu (h2 - h1)
    double ub = fb[0] * (r1 + r2), tdub = -(fb[0] * (r2 - r1)), p1p2b = 0.;
    double m1b[3], m2b[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        ub = ub + fb[1 + m] * (mom1[m] + mom2[m]);
        tdub = tdub - fb[1 + m] * (mom2[m] - mom1[m]);
        m1b[m] = fb[1 + m] * (u + tdu);
        m2b[m] = fb[1 + m] * (u - tdu);
        p1p2b = p1p2b + fb[1 + m] * tn[m];
    }
    ub = ub + fb[4] * (h1 + h2);
    tdub = tdub - fb[4] * (h2 - h1);
    double r1b = fb[0] * (u + tdu), r2b = fb[0] * (u - tdu);
    double h1b = fb[4] * (u + tdu), h2b = fb[4] * (u - tdu);
    double qpb[5], qmb[5];
    // h = gam1 p + 1/2 mom . q
    double p1b = gam1 * h1b + 0.5 * p1p2b, p2b = gam1 * h2b + 0.5 * p1p2b;
#pragma unroll
    for (int m = 0; m < 3; m++) {
        m1b[m] = m1b[m] + .5 * h1b * qp[1 + m];
        m2b[m] = m2b[m] + .5 * h2b * qm[1 + m];
        qpb[1 + m] = .5 * h1b * mom1[m];
        qmb[1 + m] = .5 * h2b * mom2[m];
    }
    if (fabs(u) >= f.c1 * sn) ub = ub + tdub * f.wig_cte * (u >= 0. ? 1. : -1.);
    double qn1b = 0.25 * ub, qn2b = 0.25 * ub;
    double tam1b = -ub * (p2 - p1);
    p1b = p1b + ub * tam1;
    p2b = p2b - ub * tam1;
    double tamb = tam >= 0. ? tam1b * f.c2 : 0.;
    double sonb = f.c3 * tamb;
    double sc = sqrt(c);
    qn1b = qn1b + sonb * (qn1 >= 0. ? 1. / sc : -1. / sc);
    double cb = sonb * (-0.5 * sqrt(qn1 * qn1) / sqrt(c * c * c));
    r1b = r1b + p1b * qp[4] * rgp;
    r2b = r2b + p2b * qm[4] * rgp;
#pragma unroll
    for (int m = 0; m < 3; m++) {
        r1b = r1b + m1b[m] * qp[1 + m];
        r2b = r2b + m2b[m] * qm[1 + m];
        qpb[1 + m] = qpb[1 + m] + qn1b * tn[m] + m1b[m] * r1;
        qmb[1 + m] = qmb[1 + m] + qn2b * tn[m] + m2b[m] * r2;
    }
    qpb[0] = r1b; qmb[0] = r2b;
    qpb[4] = p1b * r1 * rgp; qmb[4] = p2b * r2 * rgp;
#pragma unroll
    for (int e = 0; e < 5; e++) {
        wb[0][e] = f.c6 * qpb[e];
        wb[1][e] = f.c4 * qpb[e] + f.c5 * qmb[e];
        wb[2][e] = f.c5 * qpb[e] + f.c4 * qmb[e];
        wb[3][e] = f.c6 * qmb[e];
    }
    wb[2][4] = wb[2][4] + cb * (rgp * f.gam);
}

template <int D>
__device__ __forceinline__ void face_extent(const KZ& z, int& n0, int& n1, int& n2) {
    n0 = z.ni + (D == 0); n1 = z.nj + (D == 1); n2 = z.nk + (D == 2);
}

template <bool CART, int D>
__global__ void k_vjp_faces(KZ z, VjpArgs a) {
    int n0, n1, n2;
    face_extent<D>(z, n0, n1, n2);
    const long nf = (long)n0 * n1 * n2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf) return;
    const int i = 1 + (int)(t % n0), j = 1 + (int)((t / n0) % n1), k = 1 + (int)(t / ((long)n0 * n1));
    const FluxConst f = flux_const(z);
    const long l = IDX(z, i, j, k), nd = z.nd, ndm = z.ndm;
    const long inc = D == 0 ? 1 : D == 1 ? z.incj : z.inck;
    const int fi = D == 0 ? i : D == 1 ? j : k;
    const int nD = D == 0 ? z.ni : D == 1 ? z.nj : z.nk;
    double fb[5];
#pragma unroll
    for (int e = 0; e < 5; e++) {
        double p = fi <= nD ? a.lam[l + e * nd] : 0.;
        double m = fi >= 2 ? a.lam[l - inc + e * nd] : 0.;
        fb[e] = p - m;
    }
    double tn[3] = {0., 0., 0.}, sn;
    if (CART) {
        tn[D] = f.tc[D]; sn = f.s[D];
    } else {
        const double* tm = D == 0 ? z.ti : D == 1 ? z.tj : z.tk;
        const long lt = MTR(z, i, j, k);
        tn[0] = tm[lt]; tn[1] = tm[lt + ndm]; tn[2] = tm[lt + 2 * ndm];
        sn = sqrt(tn[0] * tn[0] + tn[1] * tn[1] + tn[2] * tn[2]);
    }
    double q[4][5];
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
        for (int e = 0; e < 5; e++) q[s][e] = a.rop[l + (s - 2) * inc + e * nd];
    double wb[4][5];
    ausm_face_rev(f, q, tn, sn, fb, wb);
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
        for (int e = 0; e < 5; e++) a.buf[(long)(s * 5 + e) * nf + t] = wb[s][e];
}

template <int D>
__global__ void k_vjp_gather(KZ z, VjpArgs a) {
    int n0, n1, n2;
    face_extent<D>(z, n0, n1, n2);
    const long nf = (long)n0 * n1 * n2;
    // cells: index -1 .. n_D + 2 along D, interior across
    const int c0 = D == 0 ? z.ni + 4 : z.ni, c1 = D == 1 ? z.nj + 4 : z.nj;
    const int c2 = D == 2 ? z.nk + 4 : z.nk;
    const long nc = (long)c0 * c1 * c2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc) return;
    int i = (int)(t % c0), j = (int)((t / c0) % c1), k = (int)(t / ((long)c0 * c1));
    i += D == 0 ? -1 : 1; j += D == 1 ? -1 : 1; k += D == 2 ? -1 : 1;
    const int c = D == 0 ? i : D == 1 ? j : k;
    const int nD = D == 0 ? z.ni : D == 1 ? z.nj : z.nk;
    const long l = IDX(z, i, j, k), nd = z.nd;
    double acc[5] = {0., 0., 0., 0., 0.};
#pragma unroll
    for (int s = 3; s >= 0; s--) {                // faces c - 1, c, c + 1, c + 2
        const int fidx = c + 2 - s;               // the face whose stencil slot s is this cell
        if (fidx < 1 || fidx > nD + 1) continue;
        const int fi = D == 0 ? fidx : i, fj = D == 1 ? fidx : j, fk = D == 2 ? fidx : k;
        const long tf = (long)(fi - 1) + (long)(fj - 1) * n0 + (long)(fk - 1) * n0 * n1;
#pragma unroll
        for (int e = 0; e < 5; e++) acc[e] = acc[e] + a.buf[(long)(s * 5 + e) * nf + tf];
    }
#pragma unroll
    for (int e = 0; e < 5; e++) a.out[l + e * nd] = a.out[l + e * nd] + acc[e];
}

// weight of face fi (index along the line) from the adjoint of its two cells
__device__ __forceinline__ void face_weight(const double lp[5], const double lm[5], int fi, int n, double fb[5]) {
#pragma unroll
    for (int e = 0; e < 5; e++) {
        double p = fi <= n ? lp[e] : 0.;
        double m = fi >= 2 ? lm[e] : 0.;
        fb[e] = p - m;
    }
}

template <bool CART, int D>
__device__ __forceinline__ void face_normal(const KZ& z, const FluxConst& f, int i, int j, int k,
                                            double tn[3], double& sn) {
    tn[0] = tn[1] = tn[2] = 0.;
    if (CART) {
        tn[D] = f.tc[D]; sn = f.s[D];
    } else {
        const double* tm = D == 0 ? z.ti : D == 1 ? z.tj : z.tk;
        const long lt = MTR(z, i, j, k), ndm = z.ndm;
        tn[0] = tm[lt]; tn[1] = tm[lt + ndm]; tn[2] = tm[lt + 2 * ndm];
        sn = sqrt(tn[0] * tn[0] + tn[1] * tn[1] + tn[2] * tn[2]);
    }
}

// i direction: one workgroup per (j, k) row; LDS holds the row's 20 x (ni + 1) contributions
template <bool CART>
__global__ void __launch_bounds__(256) k_vjp_row(KZ z, VjpArgs a) {
    extern __shared__ double sh[];
    const int j = 1 + (int)(blockIdx.x % z.nj), k = 1 + (int)(blockIdx.x / z.nj);
    const int n = z.ni, nf = n + 1;
    const long nd = z.nd;
    const FluxConst f = flux_const(z);
    for (int fi = 1 + (int)threadIdx.x; fi <= nf; fi += blockDim.x) {
        const long l = IDX(z, fi, j, k);
        double lp[5], lm[5], fb[5];
#pragma unroll
        for (int e = 0; e < 5; e++) {
            lp[e] = fi <= n ? a.lam[l + e * nd] : 0.;
            lm[e] = fi >= 2 ? a.lam[l - 1 + e * nd] : 0.;
        }
        face_weight(lp, lm, fi, n, fb);
        double tn[3], sn;
        face_normal<CART, 0>(z, f, fi, j, k, tn, sn);
        double q[4][5], wb[4][5];
#pragma unroll
        for (int s2 = 0; s2 < 4; s2++)
#pragma unroll
            for (int e = 0; e < 5; e++) q[s2][e] = a.rop[l + (s2 - 2) + e * nd];
        ausm_face_rev(f, q, tn, sn, fb, wb);
#pragma unroll
        for (int s2 = 0; s2 < 4; s2++)
#pragma unroll
            for (int e = 0; e < 5; e++) sh[(s2 * 5 + e) * nf + (fi - 1)] = wb[s2][e];
    }
    __syncthreads();
    for (int c = -1 + (int)threadIdx.x; c <= n + 2; c += blockDim.x) {
        const long l = IDX(z, c, j, k);
        double acc[5] = {0., 0., 0., 0., 0.};
#pragma unroll
        for (int s2 = 3; s2 >= 0; s2--) {
            const int fidx = c + 2 - s2;
            if (fidx < 1 || fidx > nf) continue;
#pragma unroll
            for (int e = 0; e < 5; e++) acc[e] = acc[e] + sh[(s2 * 5 + e) * nf + (fidx - 1)];
        }
#pragma unroll
        for (int e = 0; e < 5; e++) a.out[l + e * nd] = a.out[l + e * nd] + acc[e];
    }
}

// j or k direction: one thread per line (i fastest across threads), marching the faces of the line
template <bool CART, int D>
__global__ void __launch_bounds__(256) k_vjp_march(KZ z, VjpArgs a) {
    const int no = D == 1 ? z.nk : z.nj;
    const long nl = (long)z.ni * no;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nl) return;
    const int i = 1 + (int)(t % z.ni), o = 1 + (int)(t / z.ni);
    const int n = D == 1 ? z.nj : z.nk;
    const long nd = z.nd, inc = D == 1 ? z.incj : z.inck;
    const FluxConst f = flux_const(z);
    auto cell = [&](int c) -> long { return D == 1 ? IDX(z, i, c, o) : IDX(z, i, o, c); };
    double q[4][5], acc[4][5], lm[5];
    const long l0 = cell(1);
#pragma unroll
    for (int s2 = 0; s2 < 4; s2++)
#pragma unroll
        for (int e = 0; e < 5; e++) { q[s2][e] = a.rop[l0 + (s2 - 2) * inc + e * nd]; acc[s2][e] = 0.; }
#pragma unroll
    for (int e = 0; e < 5; e++) lm[e] = 0.;
    for (int fi = 1; fi <= n + 1; fi++) {
        const long l = l0 + (long)(fi - 1) * inc;
        double lp[5], fb[5];
#pragma unroll
        for (int e = 0; e < 5; e++) lp[e] = fi <= n ? a.lam[l + e * nd] : 0.;
        face_weight(lp, lm, fi, n, fb);
        double tn[3], sn;
        if (D == 1) face_normal<CART, 1>(z, f, i, fi, o, tn, sn);
        else face_normal<CART, 2>(z, f, i, o, fi, tn, sn);
        double wb[4][5];
        ausm_face_rev(f, q, tn, sn, fb, wb);
#pragma unroll
        for (int s2 = 0; s2 < 4; s2++)
#pragma unroll
            for (int e = 0; e < 5; e++) acc[s2][e] = acc[s2][e] + wb[s2][e];
        // cell fi - 2 has had its last face
        const long lc = l - 2 * inc;
#pragma unroll
        for (int e = 0; e < 5; e++) a.out[lc + e * nd] = a.out[lc + e * nd] + acc[0][e];
#pragma unroll
        for (int e = 0; e < 5; e++) {
            acc[0][e] = acc[1][e]; acc[1][e] = acc[2][e]; acc[2][e] = acc[3][e]; acc[3][e] = 0.;
            q[0][e] = q[1][e]; q[1][e] = q[2][e]; q[2][e] = q[3][e];
            lm[e] = lp[e];
        }
        if (fi <= n) {
#pragma unroll
            for (int e = 0; e < 5; e++) q[3][e] = a.rop[l + 2 * inc + e * nd];
        }
    }
    // cells n, n + 1, n + 2
    const long ln = cell(n);
#pragma unroll
    for (int s2 = 0; s2 < 3; s2++)
#pragma unroll
        for (int e = 0; e < 5; e++) a.out[ln + s2 * inc + e * nd] = a.out[ln + s2 * inc + e * nd] + acc[s2][e];
}

constexpr int kVjpRowMax = 400;    // longest i row staged in LDS (20 x (ni + 1) doubles <= 64 KiB)

template <bool CART, int D>
static void vjp_two_pass(const KZ& z, const VjpArgs& a, hipStream_t st);

template <bool CART, int D>
static void vjp_dir(const KZ& z, const VjpArgs& a, hipStream_t st) {
    if (D == 0 && z.ni <= kVjpRowMax) {
        size_t lds = sizeof(double) * 20 * (size_t)(z.ni + 1);
        hipLaunchKernelGGL((k_vjp_row<CART>), dim3((unsigned)((long)z.nj * z.nk)), dim3(256), lds, st, z, a);
        return;
    }
    if (D == 1 || D == 2) {
        long nl = (long)z.ni * (D == 1 ? z.nk : z.nj);
        hipLaunchKernelGGL((k_vjp_march<CART, D>), dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, st, z, a);
        return;
    }
    vjp_two_pass<CART, D>(z, a, st);
}

template <bool CART, int D>
static void vjp_two_pass(const KZ& z, const VjpArgs& a, hipStream_t st) {
    int n0, n1, n2;
    n0 = z.ni + (D == 0); n1 = z.nj + (D == 1); n2 = z.nk + (D == 2);
    long nf = (long)n0 * n1 * n2;
    hipLaunchKernelGGL((k_vjp_faces<CART, D>), dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, st, z, a);
    long nc = (long)(z.ni + (D == 0 ? 4 : 0)) * (z.nj + (D == 1 ? 4 : 0)) * (z.nk + (D == 2 ? 4 : 0));
    hipLaunchKernelGGL((k_vjp_gather<D>), dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, z, a);
}

long vjp_buffer_doubles(const KZ& z) {
    // only the i direction of rows too long for LDS goes through the buffer
    return z.ni <= kVjpRowMax ? 0 : 20 * (long)(z.ni + 1) * z.nj * z.nk;
}

int launch_residual_vjp(const KZ& z, const VjpArgs& a, hipStream_t st) {
    if (z.kflu != 1 || z.slope != 2 || z.iflow != 1) return -1;
    if (z.typ == 2) { vjp_dir<true, 0>(z, a, st); vjp_dir<true, 1>(z, a, st); vjp_dir<true, 2>(z, a, st); }
    else { vjp_dir<false, 0>(z, a, st); vjp_dir<false, 1>(z, a, st); vjp_dir<false, 2>(z, a, st); }
    return 0;
}

}  // namespace fm
