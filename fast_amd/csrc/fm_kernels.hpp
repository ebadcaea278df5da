// Kernel launch interface shared by the kernels and the C ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include "fm_device.hpp"

namespace fm {

struct Box { int i0, i1, j0, j1, k0, k1; };

// Roe wall-flux correction window (correct_flux.for, lapply = 4): the face cells of a wall BC
// (indice_cl_sdm with the normal range set to the face index) and which wall flux to use.
struct WallCorr { int idir, slip, neq_mtr, c[6]; };
constexpr int kMaxWallCorr = 12;

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
    int schedule;           // RES_CELLS (six faces per cell), RES_FACES (each face once) or
                            // RES_KCOL (a thread walks kcol_n cells up k, k faces once)
    int kcol_n;             // RES_KCOL column length (0: RES_KCOL_N)
    int ncorr;              // Roe: wall faces whose Euler flux is replaced, in BC order
    WallCorr corr[kMaxWallCorr];
    double ck_vent, c_ale;  // shape_tab_mtr.for mesh-motion factors of the wall flux
};
enum { RES_CELLS = 0, RES_FACES = 1, RES_KCOL = 2, RES_AUTO = 3 };
#ifndef FM_RES_KCOL_N
#define FM_RES_KCOL_N 4
#endif
constexpr int RES_KCOL_N = FM_RES_KCOL_N;

struct LuWork;
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
    double* xch;                  // pipelined schedule: tile-boundary exchange buffer
    int schedule;                 // LU_BLOCKS, LU_PIPELINE or LU_DATAFLOW
    int fp32;                     // blocks / dataflow: sweep arithmetic in FP32 (state, right-hand
                                  // side, results and Newton update stay FP64)
    // dataflow schedule: the blocks in diagonal order, and the products each block hands to
    // the blocks after it (face exchange, pre-filled with the pending pattern)
    const LuWork* flow_work;
    int flow_n;
    double* fx;
    int flow_step;                // dataflow: blocks take their halo products step by step (FP64)
    int flow_grid;                // dataflow: workgroups of a sweep (0: as many as are resident)
    // streaming schedule: the pencil-to-pencil exchange (pending pattern when idle) and the
    // skewed, pencil-major copy of the sweep inputs
    double* stx;
    double* skw;
    int keep_dw;                  // streaming: copy dW back to dw (else it stays in the skewed rows)
    int w_ready;                  // streaming: the state rows hold rop already (the last upper sweep
                                  // wrote them and nothing changed the state since)
    int n_ready;                  // streaming, 3dfull: the normal rows of this box are in place
};
enum { LU_BLOCKS = 0, LU_PIPELINE = 1, LU_DATAFLOW = 2, LU_STREAM = 3 };
// streaming: the arrays copied into the rows (16 on 3dcart, 25 on 3dfull; the state rows are
// skipped with w_ready, the normal rows with n_ready), and the arrays [first, end) copied back:
// dW (0-4) unless it stays in the rows (keep_dw off), the updated state (5-9) with the Newton update
__device__ __host__ __forceinline__ int lu_pack_count(bool cart, const LuArgs& a) {
    return (cart ? 16 : 25) - (a.w_ready ? 5 : 0) - (!cart && a.n_ready ? 9 : 0);
}
__device__ __host__ __forceinline__ int lu_unpack_first(const LuArgs& a) { return a.newton && !a.keep_dw ? 5 : 0; }
__device__ __host__ __forceinline__ int lu_unpack_end(const LuArgs& a) { return a.newton ? 10 : 5; }
// doubles of the streaming exchange, and of the skewed input copy, of an LU box of ni x nj x nk
// cells
__host__ __device__ long lu_stream_exchange_size(int ni, int nj, int nk);
long lu_skew_size(bool cart, int ni, int nj, int nk);
// fill n doubles of an exchange with the pending pattern (once when allocated, and after a wait
// gave up: the sweeps give back every slot they read)
void fill_pending(double* x, long n, hipStream_t st);
// doubles of the dataflow face exchange of an LU box of ni x nj x nk cells
long lu_flow_exchange_size(bool cart, int ni, int nj, int nk);

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

// Injection transfers between zones (FastC/Com/setInterpTransfersFast.cpp:640-900, type 0):
// rcv[zr][rcv_pt + v nd_r] = dnr[zd][dnr_pt + v nd_d] for the 5 variables of every point.
constexpr int kMaxTransferZones = 128;   // keeps the kernel arguments at about 2 KB
struct TransferPoint { int zr, zd; long rcv, dnr; };
struct TransferArgs {
    double* f[kMaxTransferZones];   // field level of every zone
    long nd[kMaxTransferZones];     // its NDIMDX
};
void launch_transfer(const TransferPoint* pts, long npts, const TransferArgs& a, hipStream_t st);
void launch_transfer_part(const TransferPoint* pts, long npts, const TransferArgs& a, int serial, hipStream_t st);
void launch_transfer_vjp(const TransferPoint* pts, long npts, const TransferArgs& g, int serial, hipStream_t st);

// Wall efforts (bceffort.for): per face of a window, 21 fields and the 12 contributions to the
// sums; out = 33 x nfaces (fields 0-20, then the 12 contributions)
struct EffArgs { int idir, w[6]; double pos[3]; const double* rop; double* out; };
int launch_effort(const KZ& z, const EffArgs& a, hipStream_t st);
// sums[0..11] = the 12 contributions of out (33 x nf) summed face by face
void launch_effort_sum(const double* out, long nf, double* sums, hipStream_t st);
// gradients of the pressure drag / lift coefficients at a slip-wall window (fm_adjoint.hip)
struct DpjArgs { int idir, w[6]; double cosa, sina, surfinv; const double* rop; double* dcd; double* dcl; };
int launch_dpj_dpw(const KZ& z, const DpjArgs& a, hipStream_t st);
// out += (dR/dW)^T lam for the AUSM+(P) o3 Euler face balance (fm_adjoint.hip); buf holds
// vjp_buffer_doubles(z) doubles
struct VjpArgs { const double* rop; const double* lam; double* out; double* buf; };
long vjp_buffer_doubles(const KZ& z);
int launch_residual_vjp(const KZ& z, const VjpArgs& a, hipStream_t st);

void launch_metric(const KZ& z, hipStream_t st);
void launch_coe(const KZ& z, const double* rop, int first_it, int mode, Box b, int lamin, hipStream_t st);
void launch_wiggle(const KZ& z, const double* rop, hipStream_t st);
int launch_residual(const KZ& z, const RhsArgs& a, hipStream_t st);
void launch_lu(const KZ& z, const LuArgs& a, hipStream_t st);
// the streaming schedule runs for these arguments; the dW its last run left in the skewed rows
bool lu_streams(const KZ& z, const LuArgs& a);
void launch_lu_unpack_dw(const KZ& z, const LuArgs& a, hipStream_t st);
// diagonal-batched LU-SGS (blocks schedule) over several zones of the same mesh kind
struct LuZone { KZ z; LuArgs a; };
struct LuWork { int zone, bi, bj, bk; };
void lu_work_list(bool cart, const std::vector<Box>& boxes, std::vector<LuWork>& work, std::vector<int>& offs);
void launch_lu_list(bool cart, bool lower, bool fp32, const LuZone* zones, const LuWork* work,
                    const std::vector<int>& offs, hipStream_t st);
// dataflow sweeps (lower then upper, FP64) of several zones of one mesh kind in one launch each:
// the work list in diagonal order, every zone's LuArgs with its face exchange (fx) and error word
// streaming sweeps of several zones of one mesh kind: pencils (zone, 0, jt, kt) in diagonal order,
// every zone's LuArgs with its rows (skw), exchange (stx) and error word; maxrb = the largest
// number of 64-row blocks of a zone; npack and nunpack = the largest lu_pack_count and
// lu_unpack_end - lu_unpack_first over the zones
void lu_pencil_list(const std::vector<Box>& boxes, std::vector<LuWork>& work, int& maxrb);
void launch_lu_stream_list(bool cart, const LuZone* zones, int nz, const LuWork* work, int nwork, int maxrb,
                           int npack, int nunpack, int* counter, hipStream_t st);
void launch_lu_flow_list(bool cart, const LuZone* zones, const LuWork* work, int nwork, int* counter,
                         hipStream_t st);
int lu_pipeline_failed(const LuArgs& a, hipStream_t st);
// doubles of the tile-boundary exchange buffer the pipelined sweeps of an LU box need
long lu_exchange_size(int ni, int nj, int nk);
void launch_copy_box(const KZ& z, const double* src, double* dst, Box b, hipStream_t st);
void launch_bc_face(const KZ& z, double* rop, const BcArgs& a, hipStream_t st);
void launch_coins(const KZ& z, double* rop, Box L, hipStream_t st);
// edges and corners of several zones in one launch (each zone's level and interior box)
struct CoinsZone { KZ z; double* rop; Box L; };
int coins_blocks(const KZ& z, const Box& L);     // workgroups per band a zone needs
void launch_coins_list(const CoinsZone* zones, int nz, int nblk, hipStream_t st);
void launch_norms(const KZ& z, const double* dr, double* part, int nblk, double* out, hipStream_t st);
void launch_cfl(const KZ& z, const double* rop, double* part, int nblk, double* out, hipStream_t st);
void launch_tke(const KZ& z, const double* ro, double* part, int nblk, hipStream_t st);

}  // namespace fm
