// Kernel launch interface shared by the kernels and the C ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include "fm_device.hpp"

namespace fm {

struct Box { int i0, i1, j0, j1, k0, k1; };

struct RhsArgs {
    const double* rop;      // state feeding the RHS (rop_ssiter)
    const double* rop_n;    // time level N   (implicit)
    const double* rop_n1;   // time level N-1 (implicit)
    const double* rk_in;    // explicit stage: base state
    double* rk_out;         // explicit stage: new state
    int zero_init;          // drodm starts at 0 (else coefH * drodm)
    double coefH, coefW;    // RK coefficients
    int implicit, nitcfg;
    double c1, c2, c3;      // BDF coefficients
    int schedule;           // RES_FACES (each face once) or RES_CELLS (six faces per cell)
};
enum { RES_CELLS = 0, RES_FACES = 1 };

struct LuArgs {
    int i0, i1, j0, j1, k0, k1;   // LU range (shift_lu)
    double* dw;                   // drodm in, dW out
    const double* rop;            // state for the Jacobians (J, I terms) and Newton base
    const double* rop_k;          // state for the K terms of the upper sweep
    double* out;                  // Newton-updated state
    int newton;
    // pipelined schedule (LU_PIPELINE): dW* buffer between the sweeps and the work counter /
    // error word of the persistent tiles
    double* dws;
    int* sync;
    int schedule;                 // LU_BLOCKS or LU_PIPELINE
};
enum { LU_BLOCKS = 0, LU_PIPELINE = 1 };

struct BcArgs {
    int w[6];     // ghost window (1-based cell indices)
    int idir;     // 1..6
    int kind;     // per ghost cell: 0 extrapolate, 1 viscous wall mirror, 2 periodic,
                  // 3 inviscid wall; per ghost column: 4 farfield / inflow, 5 outflow, 6 outpres
    int lrhs;     // 1: applied to the RHS (zeroes extrapolations)
    // face BCs that read the boundary metric and a reference state
    const double* tijk;   // ti, tj or tk for the face direction
    int neq_mtr;          // NEQ_IJ or NEQ_K
    double state[5];      // conservative reference state (BCzone.cpp DEFAULT_STATE)
    const double* pext;   // outpres: one pressure per face point
    int inc_bc[3];        // outpres: indbci strides of the BC window
};

void launch_metric(const KZ& z, hipStream_t st);
void launch_coe(const KZ& z, const double* rop, int first_it, int mode, Box b, int lamin, hipStream_t st);
void launch_wiggle(const KZ& z, const double* rop, hipStream_t st);
int launch_residual(const KZ& z, const RhsArgs& a, hipStream_t st);
void launch_lu(const KZ& z, const LuArgs& a, hipStream_t st);
int lu_pipeline_failed(const LuArgs& a, hipStream_t st);
void launch_copy_box(const KZ& z, const double* src, double* dst, Box b, hipStream_t st);
void launch_bc_face(const KZ& z, double* rop, const BcArgs& a, hipStream_t st);
void launch_coins(const KZ& z, double* rop, Box L, hipStream_t st);
void launch_norms(const KZ& z, const double* dr, double* part, int nblk, double* out, hipStream_t st);
void launch_cfl(const KZ& z, const double* rop, double* part, int nblk, double* out, hipStream_t st);
void launch_tke(const KZ& z, const double* ro, double* part, int nblk, hipStream_t st);

}  // namespace fm
