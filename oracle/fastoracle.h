/*
 * fastoracle -- CPU restatement of the FastS structured Navier-Stokes iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP kernels of the
 * fast_amd package; it is never linked, loaded or called by the shipped code.  Every routine
 * restates the reference Fortran kernel cited next to it, in the same per-cell floating-point
 * operation order, so that it reproduces the reference bit for bit when compiled without
 * floating-point contraction.  It is itself pinned against the compiled reference kernels
 * (oracle/_ref, built by oracle/Makefile) and against the fixtures under tests/golden/.
 *
 * Zone description: the reference's own Parameter_int / Parameter_real vectors
 * (Fast/FastS/FastS/param_solver.h).  Arrays are the reference layout: structure of arrays,
 * Fortran order (i fastest), ghost layers included (formule_param.h: inddm).
 */
#ifndef FASTORACLE_H
#define FASTORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Metric of one zone: face vectors (normal x area) and volumes, including ghost cells.
   Fast/FastC/FastC/Metric/{cp_tijk,cp_vol,tijk_extrap}.for via init_metric.cpp:255-330. */
int fo_init_metric(int* param_int, const double* x, const double* y, const double* z,
                   double* ti, double* tj, double* tk, double* vol);

/* One sub-iteration of gsdr3 for one zone (gsdr3.cpp:152-444), same contract as
   ref_substep in ref_driver.c: stop = 0 full, 1 after RHS, 2 after RHS BC + norms,
   3 after LU-SGS (before the ghost-cell fill). */
int fo_substep(int* param_int, double* param_real, int nitcfg, int nitrun, int first_it,
               int lssiter_verif, int lexit_lu, const double* x, const double* y,
               const double* z, double* ro, double* ro_m1, double* ro_p1, double* xmut,
               double* ti, double* tj, double* tk, double* vol, double* drodm, double* coe,
               double* wig, double* rdm, double* cfl, int stop);

/* Ghost-cell fill (BCzone(lrhs=0) + correct_coins). */
int fo_apply_bc(int* param_int, double* param_real, int nstep, double* rop, double* ti,
                double* tj, double* tk, const double* x, const double* y, const double* z,
                double* xmut);

/* Pieces exposed for unit tests */
void fo_invist(const int* p, const double* pr, const int* ind, const double* rop, double* xmut);
int fo_lu_sgs(int* p, double* pr, const int* ind_lu, double* rotmp, const double* rop,
              double* drodm, const double* ti, const double* tj, const double* tk,
              const double* coe, int mjrnewton);
void fo_enstrophy(const int* p, const double* pr, const double* rop, const double* ti,
                  const double* tj, const double* tk, const double* vol, double* out,
                  double* grad);
void fo_residual_blocks(const int* p, int nitcfg, int nssiter, const double* drodm, double* rdm);

/* wall efforts of one window (bceffort.for); eff[12], flu 21 x nfaces or NULL */
int fo_effort(const int* p, const double* pr, int idir, const int* ind, const double* pos,
              const double* rop, const double* xmut, const double* wig, const double* x,
              const double* y, const double* z, const double* ti, const double* tj,
              const double* tk, const double* vol, double* eff, double* flu);

/* gradients of the pressure drag / lift coefficients with respect to the conservative state
   of the wall cells of one slip-wall window (dpJ_dpW_fluausm_euler_o3_2d.for in 3-D) */
int fo_dpj_dpw(const int* p, const double* pr, int idir, const int* ind, const double* rop,
               const double* ti, const double* tj, const double* tk, double cosa, double sina,
               double surfinv, double* dcd, double* dcl);

/* out = (dR/dW)^T lam for the AUSM+(P) o3 Euler face balance R of the interior cells, W the
   primitive state of every cell (5 x NDIMDX each) */
int fo_residual_vjp(const int* p, const double* pr, const double* rop, const double* ti,
                    const double* tj, const double* tk, const double* lam, double* out);
/* the face balance alone (drodm zeroed first) */
int fo_flux_balance(const int* p, const double* pr, const double* rop, const double* wig,
                    const double* ti, const double* tj, const double* tk, const double* vol,
                    const double* xmut, double* drodm);

#ifdef __cplusplus
}
#endif
#endif
