/*
 * fastmi -- MI355X-native FastS structured Navier-Stokes iteration, C ABI.
 *
 * The drop-in boundary of the FastS compute path.  A context owns device memory for a set
 * of structured zones; each zone is described by the reference's own Parameter_int /
 * Parameter_real vectors (slot numbers of Fast/FastS/FastS/param_solver.h:80-312), and its
 * fields use the reference layout (five primitive variables rho,u,v,w,T one after the other,
 * each NDIMDX doubles, i fastest, ghost layers included).  Host buffers belong to the caller;
 * device mirrors belong to the context.  Every entry point returns 0 on success or a negative
 * FM_E* code (never exits), and fm_last_error() describes the last failure.
 *
 * Reference interfaces replaced (all under Fast/FastS/FastS/ unless noted):
 *   fm_zone_add          allocate_metric  (Metric/allocate_metric.cpp:34-481, PyTree.py:267)
 *   fm_zone_metric       init_metric      (Fast/FastC/FastC/Metric/init_metric.cpp:34-366)
 *   fm_compute           fasts._computePT (Compute/computePT.cpp:44-721, PyTree.py:42-236)
 *   fm_zone_substep      gsdr3 for one sub-iteration (Compute/gsdr3.cpp:57-484)
 *   fm_apply_bc          fasts._applyBC   (BC/applyBC.cpp:33-255, PyTree.py:1155)
 *   fm_wall_effort       fasts.compute_effort (BC/compute_effort.cpp:34-289, bceffort.for)
 *   fm_dpj_dpw           fasts.compute_dpJ_dpW (ADJOINT/compute_dpJ_dpW.cpp:33-261,
 *                        PyTree.py:3719-3735)
 *   fm_transfers_set     the flattened tc raccords (param_int_tc) that FastC builds from the
 *                        connectivity tree (Fast/FastC/FastC/PyTree.py, used by
 *                        Fast/FastC/FastC/Com/setInterpTransfersFast.cpp:640-760)
 *   fm_transfer          setInterpTransfersFast for injection raccords
 *                        (Fast/FastC/FastC/Com/setInterpTransfersFast.cpp:88-900, type 0)
 *   fm_residual_norms    display_ss_iteration (DISPLAY/display_ss_iteration.cpp:29-191)
 *   fm_enstrophy_tke     fasts._computePT_enstrophy (POST/computePT_enstrophy.cpp:127-227)
 *   fm_zone_set_field / fm_zone_get_field   the numpy buffers the reference borrows in place
 *                        (K_PYTREE::getValueAF, computePT.cpp:225-369): here they are copied
 *                        to and from device memory explicitly.
 */
#ifndef FASTMI_H
#define FASTMI_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define FM_OK 0
#define FM_E_ARG -1        /* invalid argument */
#define FM_E_UNSUPPORTED -2 /* option outside the implemented path */
#define FM_E_HIP -3        /* HIP runtime error */
#define FM_E_NOMEM -4      /* device allocation failed */

/* field identifiers for fm_zone_set_field / fm_zone_get_field */
#define FM_FIELD_RO 0      /* time level N   (5*NDIMDX) */
#define FM_FIELD_RO_M1 1   /* time level N-1 (5*NDIMDX) */
#define FM_FIELD_RO_P1 2   /* time level N+1 (5*NDIMDX) */
#define FM_FIELD_MUT 3     /* laminar viscosity (NDIMDX) */
#define FM_FIELD_DRODM 4   /* right-hand side / increment (5*NDIMDX) */
#define FM_FIELD_COE 5     /* time step and LU coefficients (NEQ_COE*NDIMDX) */
#define FM_FIELD_METRIC 6  /* ti | tj | tk | vol (NDIMDX_MTR*(2*NEQ_IJ+NEQ_K+1)) */
#define FM_FIELD_WIG 7     /* SENSEUR sensor (3*NDIMDX) */

typedef struct fm_ctx fm_ctx;

const char* fm_last_error(void);
int fm_ctx_create(fm_ctx** out, int device);
int fm_ctx_destroy(fm_ctx* ctx);

/* Register a zone.  param_int/param_real are copied (n_int / n_real entries, BC table
   included).  x, y, z: vertex coordinates with ghost vertices (NDIMDX_XYZ doubles each). */
int fm_zone_add(fm_ctx* ctx, const int* param_int, int n_int, const double* param_real,
                int n_real, const double* x, const double* y, const double* z, int* zone_id);
/* Face vectors and volumes on the device (cp_tijk, cp_vol). */
int fm_zone_metric(fm_ctx* ctx, int zone);
int fm_zone_set_field(fm_ctx* ctx, int zone, int field, const double* host, long n);
int fm_zone_get_field(fm_ctx* ctx, int zone, int field, double* host, long n);
/* Update a zone's param_real (time step, CFL ...) without re-allocating. */
int fm_zone_set_param_real(fm_ctx* ctx, int zone, const double* param_real, int n_real);

/* One sub-iteration of one zone (gsdr3).  stop: 0 full, 1 after the RHS, 2 after the RHS
   boundary conditions and residual norms, 3 after LU-SGS.  Synchronous. */
int fm_zone_substep(fm_ctx* ctx, int zone, int nitcfg, int nitrun, int first_it,
                    int lssiter_verif, int lexit_lu, int stop);
/* Ghost-cell fill of a time level (0 N, 1 N-1, 2 N+1). */
int fm_apply_bc(fm_ctx* ctx, int zone, int level, int nstep);

/* Wall efforts of one face window of a zone (replaces fasts.compute_effort,
   BC/compute_effort.cpp:34-289 -> BC/bceffort.for -> eff_flu<scheme>_* of
   Compute/template_effort.for; FastS.PyTree._computeStress, PyTree.py:2536-2566).
   window: 6 ints (i0,i1,j0,j1,k0,k1), the face index in the normal direction of idir (1..6;
   1 or n+1) and the cells across it; level: time level whose state is used (0 N, 1 N-1,
   2 N+1); pos: moment reference point (3 doubles, NULL = origin).
   effort[12]: forces x,y,z (pressure minus free-stream pressure, plus viscous stress, scaled by
   2/(rho_inf V_inf^2), outward), moments about pos, wetted area, and the 5 raw face-flux sums.
   fields (NULL or 21 x nfaces, i fastest within the window): the teff variables of
   createStressNodes (Density, MomentumX/Y/Z, EnergyStagnationDensity = the scaled fluxes,
   12 face gradients of u, v, w, T, CoefPressure, ViscosityMolecular, Density2, Pressure).
   Synchronous. */
int fm_wall_effort(fm_ctx* ctx, int zone, int level, int idir, const int* window,
                   const double* pos, double* effort, double* fields);

/* fm_wall_effort for callers that cannot allocate device memory (code under HIP graph capture,
   serving or optimisation loops that evaluate the wall force every iteration, harnesses that
   forbid hipMalloc): no counterpart in the reference.  Same arguments and the same results, bit
   for bit; the per-face work goes through a device array of the library sized for at most
   65536 faces per window (larger windows return FM_E_ARG), which every call on a device shares:
   calls that overlap on one device race.  Synchronous. */
int fm_wall_effort_noalloc(fm_ctx* ctx, int zone, int level, int idir, const int* window,
                           const double* pos, double* effort, double* fields);

/* Gradients of the pressure drag and lift coefficients of one slip-wall face window with
   respect to the conservative state (rho, rho u, rho v, rho w, rho E) of the wall cells
   (replaces fasts.compute_dpJ_dpW, ADJOINT/compute_dpJ_dpW.cpp:33-261 -> dpj_dpw_calc_meth.for
   -> ADJOINT/AUSM/dpJ_dpW_fluausm_select.for -> dpJ_dpW_fluausm_euler_o3_2d.for, whose 2-D
   statement is extended here to 3-D zones; FastS.PyTree._compute_dpJ_dpW, PyTree.py:3719).
   Per face, J_D = s (cos_a F_x + sin_a F_y) and J_L = s (cos_a F_y - sin_a F_x), F the AUSM+(P)
   o3 momentum flux of the face, s = -+2 surfinv/(rho_inf V_inf^2) on min / max faces.  The
   two ghost cells of each face follow the slip-wall mirror.  window: as fm_wall_effort, on the
   zone face of idir.  dcd, dcl: 5 x NDIMDX doubles each (the fields dpCDp_dp*, dpCLp_dp*),
   assigned at the two interior cells of each face's stencil and left as they were elsewhere.
   Zones: AUSM+(P), o3, Euler (FM_E_UNSUPPORTED otherwise).  Synchronous. */
int fm_dpj_dpw(fm_ctx* ctx, int zone, int level, int idir, const int* window, double cosa,
               double sina, double surfinv, double* dcd, double* dcl);

/* The face balance of the interior cells alone, R_l = sum_d F_d(l) - F_d(l + inc_d) (the fluxes
   the RHS kernels of Compute/navier_stokes_struct.for accumulate into drodm before the time
   terms; the function fluausm_euler_o3_2d_b.for differentiates), from the state of a time
   level.  drodm: 5 x NDIMDX doubles, zero outside the interior.  Synchronous. */
int fm_flux_balance(fm_ctx* ctx, int zone, int level, double* drodm);

/* Transposed Jacobian product of that balance: out = (dR/dW)^T lam, W the primitive state of
   every cell of the level (ghost cells included), lam read on the interior cells (replaces the
   Tapenade reverse routine ADJOINT/AUSM/2d/fluausm_euler_o3_2d_b.for, which exists for 2-D zones
   only and is not called by the reference; stated here for 3-D zones).  lam, out: 5 x NDIMDX
   doubles.  Zones: AUSM+(P), o3, Euler (FM_E_UNSUPPORTED otherwise).  Synchronous. */
int fm_residual_vjp(fm_ctx* ctx, int zone, int level, const double* lam, double* out);

/* The same product for all zones at once, with the ghost cells filled by the transfers of
   fm_transfers_set: out[n] = (dR_n/dW_n)^T lam[n], then, raccords taken in reverse order, the
   gradient on each receiver cell moved onto its donor cell (the transpose of the injection
   rcv[rcvPts] = dnr[donorPts] of setInterpTransfersFast.cpp:440-760; a periodic box joined to
   itself by raccords is covered the same way).  out[n] is then the gradient with respect to the
   cells that the raccords do not fill: the interior, and ghost cells of physical boundaries, which
   the caller chains through its boundary conditions.  lam[n], out[n]: 5 x NDIMDX doubles of zone n,
   for every zone of the context in the order they were added.  Zones: AUSM+(P), o3, Euler.
   Synchronous. */
int fm_residual_vjp_zones(fm_ctx* ctx, int level, const double* const* lam, double* const* out);

/* Injection raccords between zones (1-to-1 matches), replacing any set before.  Raccord r
   fills nb_pts[r] receiver cells of zone no_r[r] from donor cells of zone no_d[r]; rcv_pts and
   dnr_pts hold the points of all raccords one after the other, as 0-based linear indices into
   the zones' field arrays (the rcvPts / donorPts of setInterpTransfersFast.cpp:753-756).
   nrac = 0 removes every raccord.  A receiver cell that several points fill takes the value of the
   last of them, raccords and points in the order given.  At most 128 zones.  fm_compute then runs, at every
   sub-iteration, the right-hand side and LU-SGS of all zones, the transfers, and the boundary
   conditions of all zones (gsdr3.cpp:205-444). */
int fm_transfers_set(fm_ctx* ctx, int nrac, const int* no_r, const int* no_d,
                     const int64_t* nb_pts, const int64_t* rcv_pts, const int64_t* dnr_pts);
/* Apply the transfers to a time level of every zone (0 N, 1 N-1, 2 N+1).  Synchronous. */
int fm_transfer(fm_ctx* ctx, int level);

/* nit_c time steps of all zones, sub-iterations nstep_deb..nstep_fin each, time levels
   rotated after each step (computePT.cpp:534-679).  modulo_verif: residuals are computed at
   the last sub-iteration of steps where (nitrun+it) % modulo_verif == 0.  Asynchronous on
   the context stream; fm_sync waits. */
int fm_compute(fm_ctx* ctx, int nitrun, int nstep_deb, int nstep_fin, int nit_c,
               int modulo_verif, int first_it);
int fm_sync(fm_ctx* ctx);

/* Residual of the last verified sub-iteration: out[0..4] L2, out[5..9] Linf per equation,
   already divided by DTC like display_ss_iteration.cpp:140-141. */
int fm_residual_norms(fm_ctx* ctx, int zone, double* out);
/* Residuals of every sub-iteration of the last verified step (the table display_ss_iteration
   assembles, DISPLAY/display_ss_iteration.cpp:96-137, from the cprdu3s1.for block values): for
   the r-th recorded sub-iteration, out[10 r + e] = L2 and out[10 r + 5 + e] = Linf of equation
   e, divided by DTC.  *nrec receives the number of recorded sub-iterations; at most nmax are
   written. */
int fm_residual_history(fm_ctx* ctx, int zone, double* out, int nmax, int* nrec);
/* CFL monitor of the last verification step (cpcfl3.for, computePT.cpp:629-645):
   out[0] max, out[1] min, out[2] mean of the local CFL number over the interior cells. */
int fm_cfl(fm_ctx* ctx, int zone, double* out);
/* Cell sums over the interior of a time level (POST/post.for + cptaylor.for):
   out[0] = sum 1/2 rho |u|^2 vol, out[1] = sum vol, out[2] = sum 1/2 rho |curl u|^2 vol with
   the 4th-order Green-Gauss velocity gradients.  The reference returns out[2]/out[1] and
   out[0]/out[1].  Ghost cells must be filled. */
int fm_enstrophy_tke(fm_ctx* ctx, int zone, int level, double* out);

/* Timing helper for benchmarks: events recorded on the context stream around fm_compute. */
int fm_compute_timed(fm_ctx* ctx, int nitrun, int nstep_deb, int nstep_fin, int nit_c,
                     int modulo_verif, int first_it, float* ms);

/* Per-phase timing for the roofline report: runs one phase of the first sub-iteration of a
   zone `reps` times on the context stream between two HIP events and returns the mean
   milliseconds per run (after one untimed run).  phase: FM_PHASE_COE (viscosity + time step / LU coefficients),
   FM_PHASE_RESIDUAL (fused fluxes + balance + BDF update), FM_PHASE_LU (lower + upper sweeps
   with the Newton update), FM_PHASE_BC (ghost-cell fill), FM_PHASE_VJP (the transposed product
   of fm_residual_vjp, both passes and three directions).  No counterpart in the reference. */
/* Context options.  FM_OPT_LU_SCHEDULE selects how the LU-SGS sweeps are scheduled on the
   GPU (both give results bit-identical to the reference's lexicographic sweeps):
   FM_LU_BLOCKS one launch per diagonal of 8x8x8 blocks;
   FM_LU_PIPELINE persistent 8x8-line tiles that trail their neighbours.  No counterpart in
   the reference. */
#define FM_OPT_LU_SCHEDULE 1
#define FM_LU_BLOCKS 0
#define FM_LU_PIPELINE 1
/* FM_LU_DATAFLOW: one launch per sweep; the 8x8x8 blocks are handed out in diagonal order and
   each starts as soon as its three neighbour blocks have handed over their face products.  The
   waits are bounded: one that gives up sets an error word, the sweeps still run to their end
   (with wrong values), and the next call that waits for the stream -- fm_sync, fm_zone_substep,
   fm_zone_get_field, fm_residual_norms, fm_compute_timed, fm_time_phase -- returns FM_E_HIP and
   clears the word (fm_compute itself is asynchronous and returns FM_OK).  With several zones
   and FM_OPT_ZONE_BATCH on, the sweeps of all zones run as one dataflow launch per sweep. */
#define FM_LU_DATAFLOW 2
/* FM_LU_STREAM (default): one wave per pencil of 8 x 8 (j, k) lines sweeps the whole i extent as a
   skewed wavefront (neighbour products move between lanes by DPP and ds_bpermute) and hands its
   face products to the next pencils through a global exchange (bounded waits, reported as above).
   Its inputs are first copied into a skewed, pencil-major layout (one contiguous row per step and
   array; on curvilinear zones with the mean face normals) and its results copied back; the zone
   needs 16 (3dcart) or 25 (3dfull) doubles per cell more device memory for that copy.  Inside
   fm_compute the rows are kept from one sweep to the next where nothing else changed them (the
   state rows the upper sweep writes, the normals; fm_zone_set_field invalidates them) and dW is
   copied back to drodm when drodm is next read (fm_zone_get_field, fm_time_phase).  With several
   zones and FM_OPT_ZONE_BATCH on, the pencils of all zones go through one launch per sweep.  FP32
   sweeps run the dataflow schedule. */
#define FM_LU_STREAM 3
/* FM_OPT_RESIDUAL_SCHEDULE: FM_RES_CELLS evaluates the six faces of every cell;
   FM_RES_FACES evaluates every face flux once and sums them per cell in the reference order;
   FM_RES_KCOL walks a column of 4 cells up k per thread, evaluating the k faces once;
   FM_RES_AUTO (default) takes FM_RES_CELLS on Cartesian zones and FM_RES_KCOL on curvilinear
   ones.  Same results bit for bit. */
#define FM_OPT_RESIDUAL_SCHEDULE 2
#define FM_RES_CELLS 0
#define FM_RES_AUTO 3
#define FM_RES_FACES 1
#define FM_RES_KCOL 2
/* FM_OPT_ZONE_BATCH: 1 (default) runs the LU-SGS sweeps of all zones together over one work list:
   one launch per sweep over (zone, pencil) pairs (stream schedule) or (zone, block) pairs
   (dataflow), one launch per block diagonal (blocks); 0 sweeps the zones one after the other.
   Same results. */
#define FM_OPT_ZONE_BATCH 3
/* FM_OPT_GRAPH: 1 (default) captures each distinct time-step configuration (level pointers, first
   step, verification, sub-iteration range) into a HIP graph the second time it runs and replays it
   afterwards; 0 launches every kernel.  Same results.  Changing options, param_real, zones or
   transfers drops the captured graphs. */
#define FM_OPT_GRAPH 4
/* FM_OPT_LU_PRECISION: FM_LU_FP64 (default) runs the LU-SGS sweeps in FP64, bit-identical to the
   reference; FM_LU_FP32 runs the sweep arithmetic (products, solve, LDS image) in FP32 while the
   state, the right-hand side (Newton residual), the increments written back and the Newton
   update stay FP64 (mixed precision; blocks and dataflow schedules).  LU-SGS is the approximate factor
   of a defect correction, so the converged Newton residual is still FP64; the increments
   differ from the FP64 sweeps at the FP32 rounding level. */
#define FM_OPT_LU_PRECISION 5
#define FM_LU_FP64 0
#define FM_LU_FP32 1
/* FM_OPT_LU_GIVE_UP, value 1: test hook for the failure path above; sets the error word of every
   zone as a wait that gave up would (the next sweeps skip their waits and finish, the next
   waiting call reports FM_E_HIP).  FP32 sweeps and FM_LU_PIPELINE exclude each other
   (FM_E_UNSUPPORTED, in either order of setting). */
#define FM_OPT_LU_GIVE_UP 6
/* Tile shapes (same results bit for bit, only the speed changes; fast_amd.tune picks them):
   FM_OPT_RES_KCOL_N: cells per thread column of FM_RES_KCOL (1..64, 0 = built-in 4);
   FM_OPT_LU_FLOW_GRID: workgroups of a dataflow sweep (0 = as many as can be resident). */
#define FM_OPT_RES_KCOL_N 7
#define FM_OPT_LU_FLOW_GRID 8
int fm_ctx_set_option(fm_ctx* ctx, int option, int value);

#define FM_PHASE_COE 0
#define FM_PHASE_RESIDUAL 1
#define FM_PHASE_LU 2
#define FM_PHASE_BC 3
#define FM_PHASE_VJP 4
int fm_time_phase(fm_ctx* ctx, int zone, int phase, int reps, float* ms_per_run);

#ifdef __cplusplus
}
#endif
#endif
