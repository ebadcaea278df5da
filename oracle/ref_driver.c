/*
 * Test-oracle harness around the reference FastS Fortran kernels (compiled from
 * /root/reference sources by oracle/Makefile into oracle/_ref/).  This file is the
 * harness only: it owns no numerics.  It reproduces, for ONE zone, ONE OpenMP thread and
 * ONE cache block, the call sequence of the reference C++ drivers that cannot be built
 * here (they need Cassiopee's KCore):
 *
 *   allocate_metric / init_metric  Fast/FastS/FastS/Metric/allocate_metric.cpp:120-176,
 *                                  Fast/FastC/FastC/Metric/init_metric.cpp:200-330
 *   gsdr3 (one sub-iteration)      Fast/FastS/FastS/Compute/gsdr3.cpp:152-173,205-444
 *     rhs.cpp -> navier_stokes_struct_   Compute/rhs.cpp:33-51
 *     lhs.cpp -> BCzone(lrhs=1), cprdu3s1_, invlu_   Compute/lhs.cpp:53-185
 *     bcs.cpp -> BCzone(lrhs=0), correct_coins_      Compute/bcs.cpp:1-41
 *   BCzone dispatch                Fast/FastS/FastS/BC/BCzone.cpp:60-865
 *
 * Test infrastructure only: nothing in the shipped package links or loads this.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* param_solver.h slots used by the harness */
enum { NIJK = 0, NIJK_MTR = 5, NIJK_XYZ = 10, IJKV = 20, MXSSDOM_LU = 24, ITYPZONE = 25,
       IFLOW = 27, ITYPCP = 29, SIZE_SSDOM = 30, LALE = 34, NEQ = 36, NEQ_IJ = 37, NEQ_K = 38,
       NEQ_COE = 39, NDIMDX = 41, NDIMDX_XYZ = 42, NDIMDX_MTR = 43, DTLOC = 46, NSSITER = 58,
       PT_BC = 70, LU_MATCH = 75, BC_TYPE = 0, BC_IDIR = 1, BC_FEN = 2, BC_NBDATA = 8 };
enum { R_GAMMA = 1, R_ROINF = 3, R_PINF = 4, R_RONUTILDEINF = 8, R_EPSI_NEWTON = 14,
       R_VXINF = 19, R_VYINF = 20, R_VZINF = 21 };

typedef int I;
typedef double F;

/* ---- reference Fortran entry points (all arguments by reference) ---- */
extern void nature_geom_dom_(I* nijk_xyz, I* ndimt_xyz, F* x, F* y, F* z, I* degener, I* lale,
                             F* param_real, I* nijk_mtr, I* ndimdx_mtr, I* neq_ij, I* neq_k,
                             I* type_zone);
extern void cp_tijk_(I* param_int, F* x, F* y, F* z, F* ti, F* tj, F* tk, F* ti0, F* tj0, F* tk0,
                     I* ind_loop);
extern void cp_vol_(I* param_int, F* x, F* y, F* z, F* ti, F* tj, F* tk, F* ti0, F* tj0, F* tk0,
                    F* vol, I* ind_loop);
extern void tijk_extrap_(I* ndimdx_mtr, I* ndimt_xyz, I* nijk_xyz, I* nijk_mtr, I* nijk, I* neq_ij,
                         I* neq_k, I* ind_dm, I* degener, F* ti, F* tj, F* tk, F* ti0, F* tj0,
                         F* tk0, F* vol);
extern void navier_stokes_struct_(
    I* ndo, I* nthread, I* ithread, I* ithread_io, I* omp_mode, I* layer_mode, I* nsocket,
    I* socket, I* mx_synchro, I* lssiter_verif, I* lexit_lu, I* nptpsi, I* nitcfg, I* nssiter,
    I* nitrun, I* first_it, I* nb_pulse, I* flagcelln, I* param_int, F* param_real, F* temps,
    I* ijkv_sdm, I* ind_dm_zone, I* ind_dm_socket, I* ind_dm_omp, I* socket_topology, I* lok,
    I* topo_s, F* timer_omp, F* krylov, F* norm_kry, F* cfl, F* x, F* y, F* z, F* celln,
    F* celln_ibc, F* rop, F* rop_m1, F* rop_tmp, F* rop_ssiter, F* xmut, F* ti, F* tj, F* tk,
    F* vol, F* ti_df, F* tj_df, F* tk_df, F* vol_df, F* venti, F* ventj, F* ventk, F* wig,
    F* stat_wig, F* rot, F* drodm, F* coe, F* delta, F* ro_res, F* ro_src);
extern void invlu_(I* ndo, I* nitcfg, I* nitrun, I* param_int, F* param_real, I* ind_loop_lu,
                   I* ind_loop_sdm, I* mjrnewton, F* rotmp, F* rop_ssiter, F* drodm_in,
                   F* drodm_out, F* ti, F* tj, F* tk, F* venti, F* ventj, F* ventk, F* coe,
                   F* ssor, I* ssor_size);
extern void cprdu3s1_(I* ndom, I* nitcfg, I* nssiter, I* neq, I* ndimdx, I* ndim_rdm, I* nitrun,
                      I* iflw, I* les, I* ithread, I* nthread, I* omp_mode, I* mx_ssdom_lu,
                      I* it_bloc, F* epsi, I* size_ssdom, I* nisdom_residu, I* nijk, I* ijkv,
                      I* ijkv_lu, I* ind_loop, I* it_lu_ssdom, I* it_target_ssdom,
                      I* it_target_old, I* it_temp_ssdom, I* no_lu, F* drodm, F* rdm_sdm);
extern void indice_cl_sdm_(I* dir, I* npass, I* lskip, I* typbc, I* ific, I* kfic, I* ijkv,
                           I* window, I* ind_dm, I* ind_sdm, I* ind_cl, I* ind_cl119);
extern void correct_coins_(I* ndo, I* param_int, I* ind_loop, F* rop);

#define BCARGS I* idir, I* lrhs, I* nstep, I* neq_mtr, F* mobile_coef, I* param_int, \
    I* ind_loop, F* param_real, F* c4, F* c5, F* c6, F* x, F* y, F* z, F* ventijk, F* tijk, \
    F* rop, F* xmut, F* state
extern void bvbs_extrapolate_(BCARGS);
extern void bvbs_farfield_(BCARGS);
extern void bvbs_outflow_(BCARGS);
extern void bvbs_inflow_(BCARGS);
extern void bvbs_wall_inviscid_(BCARGS);
extern void bvbs_wall_viscous_adia_(BCARGS);
extern void bvbs_periodique_(BCARGS);
extern void bvbs_outpres_(BCARGS, F* data1, I* size_data, I* inc_bc);

/* ------------------------------------------------------------------------------------ */
/* allocate_metric.cpp:158: zone classification (3dfull / 3dhomo / 3dcart / 2d)        */
int ref_nature_geom_dom(I* p, F* pr, F* x, F* y, F* z, I* degen)
{
    I lale = p[LALE], ndimdx_mtr = 0, neq_ij = 0, neq_k = 0, typ = 0;
    nature_geom_dom_(p + NIJK_XYZ, &p[NDIMDX_XYZ], x, y, z, degen, &lale, pr, p + NIJK_MTR,
                     &ndimdx_mtr, &neq_ij, &neq_k, &typ);
    p[NDIMDX_MTR] = ndimdx_mtr; p[ITYPZONE] = typ; p[NEQ_IJ] = neq_ij; p[NEQ_K] = neq_k;
    return 0;
}

/* init_metric.cpp:255-330 for one thread owning the whole zone */
int ref_init_metric(I* p, F* x, F* y, F* z, I* degen, F* ti, F* tj, F* tk, F* vol)
{
    I ind_mtr[6] = {1, p[IJKV], 1, p[IJKV + 1], 1, p[IJKV + 2]};
    ind_mtr[0] -= p[NIJK + 3]; ind_mtr[2] -= p[NIJK + 3]; ind_mtr[4] -= p[NIJK + 4];
    ind_mtr[1] += p[NIJK + 3]; ind_mtr[3] += p[NIJK + 3]; ind_mtr[5] += p[NIJK + 4];
    if (p[ITYPZONE] == 1) { ind_mtr[4] = 1; ind_mtr[5] = 1; }
    else if (p[ITYPZONE] == 2) { for (int i = 0; i < 6; i++) ind_mtr[i] = 1; }
    /* no mesh motion: ti0 aliases ti (allocate_metric.cpp: lale == 0) */
    cp_tijk_(p, x, y, z, ti, tj, tk, ti, tj, tk, ind_mtr);
    cp_vol_(p, x, y, z, ti, tj, tk, ti, tj, tk, vol, ind_mtr);
    if (p[ITYPZONE] != 2) {
        I ind_dm[6] = {1, p[IJKV], 1, p[IJKV + 1], 1, p[IJKV + 2]};
        if (p[ITYPZONE] == 1 || p[ITYPZONE] == 3) ind_dm[5] = 1;
        tijk_extrap_(&p[NDIMDX_MTR], &p[NDIMDX_XYZ], p + NIJK_XYZ, p + NIJK_MTR, p + NIJK,
                     &p[NEQ_IJ], &p[NEQ_K], ind_dm, degen, ti, tj, tk, ti, tj, tk, vol);
    }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* BCzone.cpp:60-865 for one thread whose box is the whole zone, npass = 0 (gsdr3.cpp:115) */
static void default_state(I* p, F* pr, I nbdata, F* ipt_data, F* loc, F** out)
{
    (void)p;
    if (nbdata == 0) {
        loc[0] = pr[R_ROINF];
        loc[1] = pr[R_VXINF] * pr[R_ROINF];
        loc[2] = pr[R_VYINF] * pr[R_ROINF];
        loc[3] = pr[R_VZINF] * pr[R_ROINF];
        loc[4] = pr[R_PINF] / (pr[R_GAMMA] - 1.) +
                 0.5 * (pr[R_VXINF] * loc[1] + pr[R_VYINF] * loc[2] + pr[R_VZINF] * loc[3]);
        loc[5] = pr[R_RONUTILDEINF];
        *out = loc;
    } else {
        *out = ipt_data;
    }
}

int ref_bczone(I* p, F* pr, I lrhs, I nstep, I* shift_lu, F* rop, F* ti, F* tj, F* tk,
               F* x, F* y, F* z, F* xmut, I* lcorner_out)
{
    I ind_dm[6] = {1, p[IJKV], 1, p[IJKV + 1], 1, p[IJKV + 2]};
    I itest[6] = {1, p[IJKV], 1, p[IJKV + 1], 1, p[IJKV + 2]};
    I lskip[6] = {1, 1, 1, 1, 1, 1};
    I ificmax[6];
    I npass = 0, lcorner = 1;
    I pt_bcs = p[PT_BC];
    I nb_bc = (pt_bcs > 0) ? p[pt_bcs] : 0;
    F vent[3] = {0., 0., 0.};
    F c4 = 5. / 6., c5 = 2. / 6., c6 = -1. / 6.;
    I ind_cl[6], ind_cl119[6];
    if (p[LU_MATCH] != 0) p[LU_MATCH] = 1;
    for (int i = 0; i < 4; i++) ificmax[i] = p[NIJK + 3] * p[LU_MATCH];
    for (int i = 4; i < 6; i++) ificmax[i] = p[NIJK + 4] * p[LU_MATCH];
    for (I ndf = 0; ndf < nb_bc; ndf++) {
        I pt_bc = p[pt_bcs + 1 + ndf];
        I idir = p[pt_bc + BC_IDIR];
        I ip = idir - 1;
        I t = p[pt_bc + BC_TYPE];
        if (lrhs == 1) {
            if (ind_dm[ip] == itest[ip] && ((t >= 3 && t <= 7) || t == 12 || t == 18 || t == 30)) {
                lskip[ip] = 0;
                lcorner = 0;
                if ((t >= 3 && t <= 7) || t == 12) ificmax[ip] = 1;
            }
        } else {
            lskip[ip] = 1;
        }
    }
    if (lrhs == 1) {
        shift_lu[0] = ind_dm[0] - (1 - lskip[0]) * ificmax[0];
        shift_lu[1] = ind_dm[1] + (1 - lskip[1]) * ificmax[1];
        shift_lu[2] = ind_dm[2] - (1 - lskip[2]) * ificmax[2];
        shift_lu[3] = ind_dm[3] + (1 - lskip[3]) * ificmax[3];
        shift_lu[4] = ind_dm[4] - (1 - lskip[4]) * ificmax[4];
        shift_lu[5] = ind_dm[5] + (1 - lskip[5]) * ificmax[5];
        if (p[LU_MATCH] == 1) { fprintf(stderr, "ref_bczone: lu_match unsupported\n"); return -1; }
        int idirmax = (p[ITYPZONE] == 3) ? 4 : 6;
        for (int idir = 0; idir < idirmax; idir++) {
            I ind_rhs[6];
            for (int l = 0; l < 6; l++) ind_rhs[l] = ind_dm[l];
            if (idir == 0) { ind_rhs[0] = 1; ind_rhs[1] = 1; }
            if (idir == 2) { ind_rhs[2] = 1; ind_rhs[3] = 1; }
            if (idir == 4) { ind_rhs[4] = 1; ind_rhs[5] = 1; }
            if (idir == 1) ind_rhs[0] = ind_rhs[1];
            if (idir == 3) ind_rhs[2] = ind_rhs[3];
            if (idir == 5) ind_rhs[4] = ind_rhs[5];
            if (lskip[idir] == 0) {
                I idl = idir + 1, typ = 0, lsk;
                indice_cl_sdm_(&idl, &npass, &lsk, &typ, &p[NIJK + 3], &p[NIJK + 4], p + IJKV,
                               ind_rhs, ind_dm, ind_dm, ind_cl, ind_cl119);
                if (lsk == 0) {
                    F loc[6], *st; F mob = 1.;
                    I neq_mtr = (idl <= 4) ? p[NEQ_IJ] : p[NEQ_K];
                    F* tijk = (idl <= 2) ? ti : (idl <= 4) ? tj : tk;
                    default_state(p, pr, 0, NULL, loc, &st);
                    bvbs_extrapolate_(&idl, &lrhs, &nstep, &neq_mtr, &mob, p, ind_cl119, pr, &c4,
                                      &c5, &c6, x, y, z, vent, tijk, rop, xmut, st);
                }
            }
        }
    }
    for (I ndf = 0; ndf < nb_bc; ndf++) {
        I pt_bc = p[pt_bcs + 1 + ndf];
        I idir = p[pt_bc + BC_IDIR];
        I nbdata = p[pt_bc + BC_NBDATA];
        I bc_type = p[pt_bc + BC_TYPE];
        I* size_data = p + pt_bc + BC_NBDATA + 1;
        I ip = idir - 1;
        if (!(lskip[ip] == 0 || lrhs == 0)) continue;
        F* tijk = (idir <= 2) ? ti : (idir <= 4) ? tj : tk;
        I neq_mtr = (idir <= 4) ? p[NEQ_IJ] : p[NEQ_K];
        I lsk;
        indice_cl_sdm_(&idir, &npass, &lsk, &p[pt_bc + BC_TYPE], &p[NIJK + 3], &p[NIJK + 4],
                       p + IJKV, p + pt_bc + BC_FEN, ind_dm, ind_dm, ind_cl, ind_cl119);
        F* ipt_data = NULL;
        if (nbdata != 0) ipt_data = pr + p[pt_bcs + 1 + ndf + nb_bc];
        if (lsk != 0) continue;
        if (lrhs == 1 && (bc_type == 0 || bc_type == 1 || bc_type == 2 || bc_type == 10 ||
                          bc_type == 11 || bc_type == 14 || bc_type == 16 || bc_type == 17 ||
                          bc_type == 20 || bc_type == 21))
            bc_type = 0;
        F loc[6], *st; F mob = 1.;
        default_state(p, pr, nbdata, ipt_data, loc, &st);
        if (bc_type == 0) {
            bvbs_extrapolate_(&idir, &lrhs, &nstep, &neq_mtr, &mob, p, ind_cl119, pr, &c4, &c5,
                              &c6, x, y, z, vent, tijk, rop, xmut, st);
        } else if (bc_type == 1) {
            bvbs_farfield_(&idir, &lrhs, &nstep, &neq_mtr, &mob, p, ind_cl, pr, &c4, &c5, &c6,
                           x, y, z, vent, tijk, rop, xmut, st);
        } else if (bc_type == 10) {
            bvbs_outflow_(&idir, &lrhs, &nstep, &neq_mtr, &mob, p, ind_cl, pr, &c4, &c5, &c6,
                          x, y, z, vent, tijk, rop, xmut, st);
        } else if (bc_type == 13) {
            if (!(nbdata == 0 || size_data[0] <= 6)) { fprintf(stderr, "inflow file data\n"); return -1; }
            bvbs_inflow_(&idir, &lrhs, &nstep, &neq_mtr, &mob, p, ind_cl, pr, &c4, &c5, &c6,
                         x, y, z, vent, tijk, rop, xmut, st);
        } else if (bc_type == 3 || (bc_type == 4 && p[IFLOW] == 1) || bc_type == 5) {
            if (nbdata != 0) mob = ipt_data[0];
            bvbs_wall_inviscid_(&idir, &lrhs, &nstep, &neq_mtr, &mob, p, ind_cl, pr, &c4, &c5,
                                &c6, x, y, z, vent, tijk, rop, xmut, st);
        } else if ((bc_type == 6 || bc_type == 4) && p[IFLOW] > 1) {
            if (nbdata != 0) mob = ipt_data[0];
            bvbs_wall_viscous_adia_(&idir, &lrhs, &nstep, &neq_mtr, &mob, p, ind_cl, pr, &c4,
                                    &c5, &c6, x, y, z, vent, tijk, rop, xmut, st);
        } else if (bc_type == 11) {
            bvbs_periodique_(&idir, &lrhs, &nstep, &neq_mtr, &mob, p, ind_cl119, pr, &c4, &c5,
                             &c6, x, y, z, vent, tijk, rop, xmut, st);
        } else if (bc_type == 16) {
            I* fen = p + pt_bc + BC_FEN;
            I inc_bc[3];
            if (idir <= 2) { inc_bc[0] = fen[3] - fen[2] + 1; inc_bc[1] = fen[2]; inc_bc[2] = fen[4]; }
            else if (idir <= 4) { inc_bc[0] = fen[1] - fen[0] + 1; inc_bc[1] = fen[0]; inc_bc[2] = fen[4]; }
            else { inc_bc[0] = fen[1] - fen[0] + 1; inc_bc[1] = fen[0]; inc_bc[2] = fen[2]; }
            I sz = 1; F* d1 = NULL;
            if (nbdata != 0) {
                if (size_data[0] == 1) { mob = ipt_data[0]; d1 = ipt_data + 1; sz = size_data[1]; }
                else { d1 = ipt_data; if (nbdata == 2) { mob = ipt_data[size_data[0]]; sz = size_data[0]; } }
            }
            I zero = 0;
            bvbs_outpres_(&idir, &lrhs, &nstep, &neq_mtr, &mob, p, ind_cl, pr, &c4, &c5, &c6,
                          x, y, z, vent, tijk, rop, xmut, st, d1, nbdata ? size_data : &zero, inc_bc);
        } else {
            fprintf(stderr, "ref_bczone: BC type %d unsupported by the harness\n", bc_type);
            return -1;
        }
    }
    if (lcorner_out) *lcorner_out = lcorner;
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* One sub-iteration of gsdr3 (gsdr3.cpp:152-444) for one zone, one thread.
 * stop: 0 = full sub-iteration, 1 = return after the RHS (navier_stokes_struct),
 *       2 = return after the implicit RHS BC + norms (before invlu), 3 = after invlu (before
 *       the ghost-cell fill).
 * rdm: residual map, size nssiter*neq*2*ndim_rdm (cprdu3s1.for: rdm_sdm).               */
int ref_substep(I* p, F* pr, I nitcfg, I nitrun, I first_it, I lssiter_verif, I lexit_lu,
                F* x, F* y, F* z, F* ro, F* ro_m1, F* ro_p1, F* xmut, F* ti, F* tj, F* tk,
                F* vol, F* drodm, F* coe, F* wig, F* rdm, F* cfl, I stop)
{
    I nd = 0, nthread = 1, ithread = 1, ithread_io = 1, omp_mode = 0, layer_mode = 1;
    I nsocket = 1, socket = 1, mx_synchro = 1000, nptpsi = 1, nb_pulse = 0, flagcelln = 0;
    I nssiter = p[NSSITER];
    F temps = 0.;
    I ijkv_sdm[3] = {1, 1, 1};
    I ind_dm[6] = {1, p[IJKV], 1, p[IJKV + 1], 1, p[IJKV + 2]};
    I ind_socket[6], ind_omp[6];
    memcpy(ind_socket, ind_dm, sizeof ind_dm);
    memcpy(ind_omp, ind_dm, sizeof ind_dm);
    I topo[3] = {1, 1, 1}, topo_s[3] = {1, 1, 1};
    I* lok = (I*)calloc(mx_synchro * 4, sizeof(I));
    F timer[16] = {0}, krylov[4] = {0}, norm_kry = 0., dummy[8] = {0};
    F vent[3] = {0., 0., 0.};
    I ndimdx = p[NDIMDX];
    F* stat_wig = (F*)calloc((size_t)3 * ndimdx, sizeof(F));
    F* rot = (F*)calloc((size_t)3 * ndimdx, sizeof(F));
    F* delta = (F*)calloc((size_t)ndimdx, sizeof(F));
    F* ro_res = (F*)calloc((size_t)ndimdx * 6, sizeof(F));
    F* ro_src = (F*)calloc((size_t)ndimdx * 7, sizeof(F));
    F *ro_ssiter, *ro_cl;
    int rc = 0;
    /* gsdr3.cpp:152-173: which time level feeds the RHS and receives the BC */
    if (nitcfg == 1) { ro_ssiter = ro; ro_cl = ro_p1; }
    else if (p[ITYPCP] < 2) { ro_ssiter = ro_p1; ro_cl = ro_p1; }
    else if (nitcfg % 2 == 0) { ro_ssiter = ro_p1; ro_cl = ro; }
    else { ro_ssiter = ro; ro_cl = ro_p1; }

    navier_stokes_struct_(&nd, &nthread, &ithread, &ithread_io, &omp_mode, &layer_mode, &nsocket,
                          &socket, &mx_synchro, &lssiter_verif, &lexit_lu, &nptpsi, &nitcfg,
                          &nssiter, &nitrun, &first_it, &nb_pulse, &flagcelln, p, pr, &temps,
                          ijkv_sdm, ind_dm, ind_socket, ind_omp, topo_s, lok, topo, timer, krylov,
                          &norm_kry, cfl, x, y, z, ro, dummy, ro, ro_m1, ro_p1, ro_ssiter, xmut,
                          ti, tj, tk, vol, ti, tj, tk, vol, vent, vent, vent, wig, stat_wig, rot,
                          drodm, coe, delta, ro_res, ro_src);
    if (stop == 1) goto done;

    if (p[ITYPCP] <= 1) {
        /* lhs.cpp:53-69 then residus_navier.h then invlu (lhs.cpp:140-161) */
        I shift_lu[6]; I lcorner = 1;
        rc = ref_bczone(p, pr, 1, nitcfg, shift_lu, drodm, ti, tj, tk, x, y, z, xmut, &lcorner);
        if (rc) goto done;
        if (lcorner == 0) correct_coins_(&nd, p, shift_lu, drodm);
        if (lssiter_verif == 1 && rdm != NULL && (p[ITYPCP] != 2 || p[DTLOC] == 1)) {
            I ijkv_lu[3];
            for (int i = 0; i < 3; i++) {
                ijkv_lu[i] = p[IJKV + i] / p[SIZE_SSDOM + i];
                if (ijkv_lu[i] < 1) ijkv_lu[i] = 1;
            }
            I ndim_rdm = ijkv_lu[0] * ijkv_lu[1] * ijkv_lu[2];
            I mx = p[MXSSDOM_LU] > ndim_rdm ? p[MXSSDOM_LU] : ndim_rdm;
            I* itab = (I*)calloc((size_t)mx * 5 + 4 + 2 * nssiter, sizeof(I));
            I* nisdom = itab + mx * 5;
            I it_bloc = 1, iles = 0, neq = p[NEQ];
            I* no_lu = itab + 4 * mx;
            /* every residual sub-block is computed (souszones_list.cpp: lssiter_verif) */
            for (int i = 0; i < nssiter; i++) nisdom[i] = ndim_rdm;
            for (int i = 0; i < ndim_rdm; i++) no_lu[i] = i + 1;
            cprdu3s1_(&nd, &nitcfg, &nssiter, &neq, &ndimdx, &ndim_rdm, &nitrun, &p[IFLOW], &iles,
                      &ithread, &nthread, &omp_mode, &mx, &it_bloc, &pr[R_EPSI_NEWTON],
                      p + SIZE_SSDOM, nisdom, p + NIJK, p + IJKV, ijkv_lu, ind_dm, itab,
                      itab + mx, itab + 2 * mx, itab + 3 * mx, itab + 4 * mx, drodm, rdm);
            free(itab);
        }
        if (stop == 2) goto done;
        if (lexit_lu == 0) {
            I mjrnewton = 1, ssor_size = 1;
            invlu_(&nd, &nitcfg, &nitrun, p, pr, shift_lu, ind_dm, &mjrnewton, ro_p1, ro_ssiter,
                   drodm, drodm, ti, tj, tk, vent, vent, vent, coe, dummy, &ssor_size);
        }
    }
    if (stop == 3) goto done;
    /* gsdr3.cpp:422-444 -> bcs.cpp */
    if (lexit_lu == 0) {
        I shift_lu[6];
        rc = ref_bczone(p, pr, 0, nitcfg, shift_lu, ro_cl, ti, tj, tk, x, y, z, xmut, NULL);
        if (rc) goto done;
        correct_coins_(&nd, p, ind_dm, ro_cl);
    }
done:
    free(lok); free(stat_wig); free(rot); free(delta); free(ro_res); free(ro_src);
    return rc;
}

/* POST/post.for for one thread owning the zone (computePT_enstrophy.cpp:170-196):
   out[0] = tke sum, out[1] = enstrophy sum, out[2] = volume sum (the counter). */
extern void post_(I* ndo, I* nthread, I* ithread, I* nsocket, I* socket, I* mx_synchro,
                  I* neq_grad, I* param_int, F* param_real, F* tke, F* enst, F* compteur,
                  I* ijkv_sdm, I* ind_dm_zone, I* ind_dm_socket, I* ind_dm_omp, I* topo_s,
                  I* socket_topology, I* lok, F* rop, F* ti, F* tj, F* tk, F* vol, F* grad);

int ref_enstrophy(I* p, F* pr, F* rop, F* ti, F* tj, F* tk, F* vol, F* out, F* grad_out)
{
    I nd = 0, nthread = 1, ithread = 1, nsocket = 1, socket = 1, mx_synchro = 1000, neq_grad = 3;
    I ijkv_sdm[3] = {1, 1, 1};
    I ind_dm[6] = {1, p[IJKV], 1, p[IJKV + 1], 1, p[IJKV + 2]};
    I ind_socket[6], ind_omp[6];
    memcpy(ind_socket, ind_dm, sizeof ind_dm);
    memcpy(ind_omp, ind_dm, sizeof ind_dm);
    I topo[3] = {1, 1, 1}, topo_s[3] = {1, 1, 1};
    I* lok = (I*)calloc(mx_synchro * 4, sizeof(I));
    F* grad = (F*)calloc((size_t)p[NDIMDX] * neq_grad * 3, sizeof(F));
    F tke = 0., enst = 0., cpt = 0.;
    post_(&nd, &nthread, &ithread, &nsocket, &socket, &mx_synchro, &neq_grad, p, pr, &tke, &enst,
          &cpt, ijkv_sdm, ind_dm, ind_socket, ind_omp, topo_s, topo, lok, rop, ti, tj, tk, vol, grad);
    out[0] = tke; out[1] = enst; out[2] = cpt;
    if (grad_out) memcpy(grad_out, grad, sizeof(F) * (size_t)p[NDIMDX] * neq_grad * 3);
    free(lok); free(grad);
    return 0;
}

/* Ghost-cell fill alone (FastS.PyTree._applyBC -> applyBC.cpp -> BCzone(lrhs=0) + corners). */
int ref_apply_bc(I* p, F* pr, I nstep, F* rop, F* ti, F* tj, F* tk, F* x, F* y, F* z, F* xmut)
{
    I nd = 0, shift_lu[6];
    I ind_dm[6] = {1, p[IJKV], 1, p[IJKV + 1], 1, p[IJKV + 2]};
    int rc = ref_bczone(p, pr, 0, nstep, shift_lu, rop, ti, tj, tk, x, y, z, xmut, NULL);
    if (rc) return rc;
    correct_coins_(&nd, p, ind_dm, rop);
    return 0;
}

/* Wall efforts of one window (FastS.PyTree._computeStress -> compute_effort.cpp:200-258 ->
   bceffort.for -> eff_flu<scheme>_<model>_<slope>_<mesh>, one thread owning the window).
   ind: face window (the face index in the normal direction, cells across it), idir 1..6.
   The effort descriptor holds what createStressNodes (PyTree.py:2405-2497) stores: NEQ 21,
   NDIMDX = number of faces, the window (EFF_LOOP 23-28), the flux-array origin and strides
   (EFF_I0 29-33), the zone number (EFF_NONZ 34) and the direction (EFF_IDIR 35).
   out: the 12 sums of the routine; flu: 21 fields per face (may be NULL). */
extern void bceffort_(I* ndom, I* ithread, I* param_int, F* param_real, I* param_int_eff,
                      I* ind_loop, F* effort, F* xyz_ref, F* rop, F* flu, F* wig, F* xmut,
                      F* x, F* y, F* z, F* ti, F* tj, F* tk, F* vol, F* venti, F* ventj, F* ventk);

int ref_effort(I* p, F* pr, I idir, I* ind, F* pos, F* rop, F* xmut, F* wig, F* x, F* y, F* z,
               F* ti, F* tj, F* tk, F* vol, F* out, F* flu_out)
{
    I nd = 0, ithread = 1;
    I pe[64];
    memset(pe, 0, sizeof pe);
    I n0 = ind[1] - ind[0] + 1, n1 = ind[3] - ind[2] + 1, n2 = ind[5] - ind[4] + 1;
    I nf = n0 * n1 * n2;
    pe[NEQ] = 21; pe[NDIMDX] = nf;
    for (int q = 0; q < 6; q++) pe[23 + q] = ind[q];
    pe[29] = ind[0]; pe[30] = ind[2]; pe[31] = ind[4]; pe[32] = n0; pe[33] = n0 * n1;
    pe[34] = 0; pe[35] = idir;
    F eff[12];
    memset(eff, 0, sizeof eff);
    F* flu = (F*)calloc((size_t)nf * 21, sizeof(F));
    F vent[3] = {0., 0., 0.};
    F* mut = p[IFLOW] > 1 ? xmut : rop;       /* compute_effort.cpp:124-129 */
    bceffort_(&nd, &ithread, p, pr, pe, ind, eff, pos, rop, flu, wig, mut, x, y, z, ti, tj, tk,
              vol, vent, vent, vent);
    memcpy(out, eff, sizeof eff);
    if (flu_out) memcpy(flu_out, flu, sizeof(F) * (size_t)nf * 21);
    free(flu);
    return 0;
}

/* compute_dpJ_dpW.cpp:220-230 -> dpj_dpw_calc_meth -> dpJ_dpW_fluausm_select (option 213),
   called here at its 2-D routine directly: the reference codes no 3-D option.  The routine
   is run on 3-D arrays whose k planes are independent 2-D problems (zero w, face vectors in
   the (x, y) plane); NEQ_IJ is given as 2, the 2-D zone's metric count, so that shape_tab_mtr
   drops the z component of the wall normal (ck_mtr = 0), as for a 2-D zone. */
void dpj_dpw_fluausm_euler_o3_2d_(I* ndom, I* ithread, I* p, F* pr, I* pe, I* ind_loop, F* pos,
                                  F* rop, F* wig, F* x, F* y, F* z, F* dcd, F* dcl, F* venti,
                                  F* ventj, F* ventk, F* ti, F* tj, F* tk, F* vol, F* xmut,
                                  F* cosa, F* sina, F* surfinv);

int ref_dpj_dpw(I* p, F* pr, I idir, I* ind, F* rop, F* wig, F* x, F* y, F* z, F* ti, F* tj,
                F* tk, F* vol, F cosa, F sina, F surfinv, F* dcd, F* dcl)
{
    I nd = 0, ithread = 1;
    I pe[64], p2[137];               /* the header slots of Parameter_int */
    memset(pe, 0, sizeof pe);
    memcpy(p2, p, sizeof p2);
    p2[NEQ_IJ] = 2;
    I n0 = ind[1] - ind[0] + 1, n1 = ind[3] - ind[2] + 1, n2 = ind[5] - ind[4] + 1;
    pe[NEQ] = 21; pe[NDIMDX] = n0 * n1 * n2;
    for (int q = 0; q < 6; q++) pe[23 + q] = ind[q];
    pe[29] = ind[0]; pe[30] = ind[2]; pe[31] = ind[4]; pe[32] = n0; pe[33] = n0 * n1;
    pe[34] = 0; pe[35] = idir;
    I loop[6];
    memcpy(loop, ind, sizeof loop);
    F vent[3] = {0., 0., 0.}, pos[3] = {0., 0., 0.};
    dpj_dpw_fluausm_euler_o3_2d_(&nd, &ithread, p2, pr, pe, loop, pos, rop, wig, x, y, z, dcd, dcl,
                                 vent, vent, vent, ti, tj, tk, vol, rop, &cosa, &sina, &surfinv);
    return 0;
}
