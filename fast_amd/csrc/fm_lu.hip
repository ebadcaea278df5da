// LU-SGS sweeps (invlu_l.for / invlu_u.for) and the Newton update (mjr_newton.for).
//
// The lower sweep solves (D + L) dW* = R cell by cell in increasing i, j, k; every cell needs
// the final dW* of its three lower neighbours.  The upper sweep is the mirror image.  Results
// must be bit-identical to the reference's lexicographic loops, so every cell evaluates the
// same expressions in the same order; only the schedule differs:
//
//  * the LU box is cut into blocks of CI x 8 x 8 cells (i x j x k).  Blocks with the same
//    bi + bj + bk are independent: either one launch per block diagonal, in sweep order, or (the
//    dataflow schedule) one launch per sweep whose workgroups take the blocks in diagonal order
//    and receive their halo products from the neighbour blocks through a face exchange;
//  * a block is first staged into LDS with coalesced loads (the latency of global memory is
//    paid once per block, not once per cell), together with the products of its halo cells,
//    which earlier launches finished;
//  * one wave then sweeps the block: lane (jj, kk) owns the i-line segment of (j, k) and at
//    step s handles cell ii = s - jj - kk (a skewed wavefront).  The lane that needs cell
//    (i, j-1, k) is the one just "below" in j, which produced it one step earlier: the
//    neighbour products B(W) dW of a finished cell travel through LDS (j, k) or stay in
//    registers (i).
//
// The off-diagonal products b = B(W_nb) dW_nb are formed by the cell that owns W_nb, exactly
// as the reference forms them for the cell that consumes them (same operands, same order).
#include "fm_kernels.hpp"

// Cells per line segment of an LU block (its i extent), per mesh kind.  With 8, a Cartesian
// block takes 80 KB of LDS (two blocks per CU) and a curvilinear one 118 KB; 16 Cartesian
// cells make 128-byte row segments at 148 KB (one block per CU).
#ifndef FM_LU_CI_CART
#define FM_LU_CI_CART 8
#endif
#ifndef FM_LU_CI_FULL
#define FM_LU_CI_FULL 8
#endif
template <bool CART> constexpr int lu_ci() { return CART ? FM_LU_CI_CART : FM_LU_CI_FULL; }
// Workgroups of a dataflow sweep (each takes blocks until none is left).
#ifndef FM_LU_FLOW_GRID
#define FM_LU_FLOW_GRID 2048
#endif
#include <algorithm>
#include <array>
#include <vector>

namespace fm {

// T: the arithmetic type of the sweeps (double; float for the mixed-precision option, whose
// state, right-hand side, Newton update and results stay double)
template <typename T = double>
struct LUC { T gam1, gamm1, cp, gam2; };

template <typename T = double>
__device__ __forceinline__ LUC<T> lu_const(const KZ& z) {
    double gam1 = z.gam, gamm1 = gam1 - 1., cp = z.gam * z.cv, gam2 = gamm1 - 1.;
    LUC<T> c;
    c.gam1 = (T)gam1; c.gamm1 = (T)gamm1; c.cp = (T)cp; c.gam2 = (T)gam2;
    return c;
}

// The matrix B_D(u, v, w, t; cs, face normal) of a neighbour term (invlu_l.for / invlu_u.for),
// row by row: m[r][c] multiplies d[c] in row r.  Only the entries the Cartesian rows use are set
// there; the expressions are the reference's (lu_i/j/k_3dcart.for, lu_*_3dfull.for).
template <typename T = double>
struct LuB { T m[5][5]; };

template <bool CART, int D, typename T = double>
__device__ __forceinline__ LuB<T> lu_coef(const LUC<T>& c, T u, T v, T w, T t, T cs, T tcx, T tcy, T tcz) {
    LuB<T> B;
    T q2 = (T).5 * (u * u + v * v + w * w);
    T h = c.cp * t + q2;
    T ph2 = c.gamm1 * q2;
    T gam1 = c.gam1, gamm1 = c.gamm1, gam2 = c.gam2;
    if (CART) {
        T tc = tcx;
        if (D == 0) {
            T qn = tc * u;
            B.m[0][0] = cs; B.m[0][1] = tc;
            B.m[1][0] = (tc * ph2 - u * qn); B.m[1][1] = (qn - tc * gam2 * u) + cs; B.m[1][2] = (-tc * gamm1 * v);
            B.m[1][3] = (-tc * gamm1 * w); B.m[1][4] = tc * gamm1;
            B.m[2][0] = (-v * qn); B.m[2][1] = (tc * v); B.m[2][2] = ((qn) + cs);
            B.m[3][0] = (-w * qn); B.m[3][1] = (tc * w); B.m[3][3] = ((qn) + cs);
            B.m[4][0] = qn * (ph2 - h); B.m[4][1] = (tc * h - gamm1 * u * qn); B.m[4][2] = (-gamm1 * v * qn);
            B.m[4][3] = (-gamm1 * w * qn); B.m[4][4] = (gam1 * qn) + cs;
        } else if (D == 1) {
            T qn = tc * v;
            B.m[0][0] = cs; B.m[0][2] = tc;
            B.m[1][0] = (-u * qn); B.m[1][1] = (qn + cs); B.m[1][2] = (tc * u);
            B.m[2][0] = (tc * ph2 - v * qn); B.m[2][1] = (-tc * gamm1 * u); B.m[2][2] = (qn - tc * gam2 * v) + cs;
            B.m[2][3] = (-tc * gamm1 * w); B.m[2][4] = tc * gamm1;
            B.m[3][0] = (-w * qn); B.m[3][2] = (tc * w); B.m[3][3] = ((qn) + cs);
            B.m[4][0] = qn * (ph2 - h); B.m[4][1] = (-gamm1 * u * qn); B.m[4][2] = (tc * h - gamm1 * v * qn);
            B.m[4][3] = (-gamm1 * w * qn); B.m[4][4] = (gam1 * qn) + cs;
        } else {
            T qn = tc * w;
            B.m[0][0] = cs; B.m[0][3] = tc;
            B.m[1][0] = (-u * qn); B.m[1][1] = (qn + cs); B.m[1][3] = (tc * u);
            B.m[2][0] = (-v * qn); B.m[2][2] = (qn + cs); B.m[2][3] = (tc * v);
            B.m[3][0] = tc * ph2 - w * qn; B.m[3][1] = -tc * gamm1 * u; B.m[3][2] = -tc * gamm1 * v;
            B.m[3][3] = (qn - tc * gam2 * w) + cs; B.m[3][4] = tc * gamm1;
            B.m[4][0] = qn * (ph2 - h); B.m[4][1] = -gamm1 * u * qn; B.m[4][2] = -gamm1 * v * qn;
            B.m[4][3] = tc * h - gamm1 * w * qn; B.m[4][4] = gam1 * qn + cs;
        }
        return B;
    }
    T qn = tcx * u + tcy * v + tcz * w;
    B.m[0][0] = cs; B.m[0][1] = tcx; B.m[0][2] = tcy; B.m[0][3] = tcz;
    B.m[1][0] = (tcx * ph2 - u * qn); B.m[1][1] = (qn - tcx * gam2 * u) + cs;
    B.m[1][2] = (tcy * u - tcx * gamm1 * v); B.m[1][3] = (tcz * u - tcx * gamm1 * w); B.m[1][4] = tcx * gamm1;
    B.m[2][0] = (tcy * ph2 - v * qn); B.m[2][1] = (tcx * v - tcy * gamm1 * u);
    B.m[2][2] = (qn - tcy * gam2 * v) + cs; B.m[2][3] = (tcz * v - tcy * gamm1 * w); B.m[2][4] = tcy * gamm1;
    B.m[3][0] = (tcz * ph2 - w * qn); B.m[3][1] = (tcx * w - tcz * gamm1 * u);
    B.m[3][2] = (tcy * w - tcz * gamm1 * v); B.m[3][3] = (qn - tcz * gam2 * w) + cs; B.m[3][4] = tcz * gamm1;
    B.m[4][0] = qn * (ph2 - h); B.m[4][1] = (tcx * h - gamm1 * u * qn); B.m[4][2] = (tcy * h - gamm1 * v * qn);
    B.m[4][3] = (tcz * h - gamm1 * w * qn); B.m[4][4] = (gam1 * qn) + cs;
    return B;
}

// b = B d, each row summed left to right over the entries the row has (the reference's order)
template <bool CART, int D, typename T = double>
__device__ __forceinline__ void lu_apply(const LuB<T>& B, const T d[5], T b[5]) {
    const T d1 = d[0], d2 = d[1], d3 = d[2], d4 = d[3], d5 = d[4];
    if (CART) {
        if (D == 0) {
            b[0] = B.m[0][0] * d1 + B.m[0][1] * d2;
            b[1] = B.m[1][0] * d1 + B.m[1][1] * d2 + B.m[1][2] * d3 + B.m[1][3] * d4 + B.m[1][4] * d5;
            b[2] = B.m[2][0] * d1 + B.m[2][1] * d2 + B.m[2][2] * d3;
            b[3] = B.m[3][0] * d1 + B.m[3][1] * d2 + B.m[3][3] * d4;
            b[4] = B.m[4][0] * d1 + B.m[4][1] * d2 + B.m[4][2] * d3 + B.m[4][3] * d4 + B.m[4][4] * d5;
        } else if (D == 1) {
            b[0] = B.m[0][0] * d1 + B.m[0][2] * d3;
            b[1] = B.m[1][0] * d1 + B.m[1][1] * d2 + B.m[1][2] * d3;
            b[2] = B.m[2][0] * d1 + B.m[2][1] * d2 + B.m[2][2] * d3 + B.m[2][3] * d4 + B.m[2][4] * d5;
            b[3] = B.m[3][0] * d1 + B.m[3][2] * d3 + B.m[3][3] * d4;
            b[4] = B.m[4][0] * d1 + B.m[4][1] * d2 + B.m[4][2] * d3 + B.m[4][3] * d4 + B.m[4][4] * d5;
        } else {
            b[0] = B.m[0][0] * d1 + B.m[0][3] * d4;
            b[1] = B.m[1][0] * d1 + B.m[1][1] * d2 + B.m[1][3] * d4;
            b[2] = B.m[2][0] * d1 + B.m[2][2] * d3 + B.m[2][3] * d4;
            b[3] = B.m[3][0] * d1 + B.m[3][1] * d2 + B.m[3][2] * d3 + B.m[3][3] * d4 + B.m[3][4] * d5;
            b[4] = B.m[4][0] * d1 + B.m[4][1] * d2 + B.m[4][2] * d3 + B.m[4][3] * d4 + B.m[4][4] * d5;
        }
        return;
    }
    b[0] = B.m[0][0] * d1 + B.m[0][1] * d2 + B.m[0][2] * d3 + B.m[0][3] * d4;
#pragma unroll
    for (int r = 1; r < 5; r++)
        b[r] = B.m[r][0] * d1 + B.m[r][1] * d2 + B.m[r][2] * d3 + B.m[r][3] * d4 + B.m[r][4] * d5;
}

// b = B_D(u, v, w, t; cs, face normal) d   (invlu_l.for / invlu_u.for neighbour term)
template <bool CART, int D, typename T = double>
__device__ __forceinline__ void lu_prod(const LUC<T>& c, T u, T v, T w, T t, T cs, T tcx, T tcy, T tcz,
                                        const T d[5], T b[5]) {
    lu_apply<CART, D, T>(lu_coef<CART, D, T>(c, u, v, w, t, cs, tcx, tcy, tcz), d, b);
}

// Face normal of direction D used for the product of the cell at metric index lt: the mean of
// its two faces (3dfull); the single Cartesian value otherwise.
template <bool CART, int D>
__device__ __forceinline__ void lu_normal(const KZ& z, long lt, double& tcx, double& tcy, double& tcz) {
    if (CART) {
        tcx = D == 0 ? z.ti[0] : D == 1 ? z.tj[0] : z.tk[0];
        tcy = tcz = 0.;
        return;
    }
    const double* tm = D == 0 ? z.ti : D == 1 ? z.tj : z.tk;
    long inc = D == 0 ? z.mi : D == 1 ? z.mj : z.mk;
    long ndm = z.ndm, l2 = lt + inc;
    tcx = 0.5 * (tm[l2] + tm[lt]);
    tcy = 0.5 * (tm[l2 + ndm] + tm[lt + ndm]);
    tcz = 0.5 * (tm[l2 + 2 * ndm] + tm[lt + 2 * ndm]);
}

// Product of a neighbour finished by an earlier launch, read back from global memory.
template <bool CART, int D, typename T = double>
__device__ __forceinline__ void lu_prod_global(const KZ& z, const LUC<T>& c, const double* __restrict__ rop,
                                               const double* __restrict__ dw, int i, int j, int k,
                                               double signe, T b[5]) {
    long nd = z.nd, l1 = IDX(z, i, j, k);
    double tcx, tcy, tcz;
    lu_normal<CART, D>(z, CART ? 0 : MTR(z, i, j, k), tcx, tcy, tcz);
    T d[5];
#pragma unroll
    for (int e = 0; e < 5; e++) d[e] = (T)dw[l1 + e * nd];
    double cs = z.coe[l1 + (1 + D) * nd] * signe;
    lu_prod<CART, D, T>(c, (T)rop[l1 + nd], (T)rop[l1 + 2 * nd], (T)rop[l1 + 3 * nd], (T)rop[l1 + 4 * nd],
                        (T)cs, (T)tcx, (T)tcy, (T)tcz, d, b);
}

// LDS image of one block: the block's own fields (state, coefficients, right-hand side and,
// for 3dfull, the three mean face normals), the products of the halo cells finished by earlier
// launches, and the exchange between the solving wave and the product waves.  Rows are indexed by lane so that
// the sweep reads (one cell per lane, lanes on different i) hit 16 distinct 8-byte banks.
template <bool CART, int CI, typename T = double>
struct LuLds {
    // W 0-4, coe1..4 times the sweep sign 5-8, 1/coe5 9, right-hand side 10-14 (upper sweep:
    // times coe5), [3dfull: normals 15-23]
    static constexpr int NV = CART ? 15 : 24;
    T st[NV][CI][64];
    T hI[5][64];
    T hJ[5][CI][8];
    T hK[5][CI][8];
    T xs[5][64];             // result of the cell each lane solved at this step
    T wk[4][64];             // its Newton-updated state (upper sweep in place)
    T px[3][5][64];          // its i, j, k products
};

// One block of CI x 8 x 8 cells (i x j x k, in sweep order) per launch diagonal.
//  1. the 4 waves stage the block into LDS with coalesced loads (16 consecutive i per line)
//     and form the products of the halo cells;
//  2. the skewed wavefront from LDS: lane (jj, kk) owns line (j, k) and handles cell
//     ii = s - jj - kk at step s; wave 0 solves, waves 1-3 form the products; results go
//     straight to global memory.
// Values handed between concurrently running blocks or tiles validate themselves: the buffer
// is pre-filled with a signalling-NaN pattern that arithmetic never produces, producers store
// with agent-scope stores (seen by every XCD) and consumers re-read with agent-scope loads until
// the pattern is gone.  Every wait is bounded: past the bound, or once any wave has given up,
// the error word is set and the kernel runs to its end, so the grid always drains.
constexpr long long kLuPending = 0x7FF4C0DEC0DE0001ll;
constexpr int kLuSpinMax = 1 << 22;

__device__ __forceinline__ double lu_ldc(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lu_stc(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool lu_any_pending(const double* v, int n) {
    bool p = false;
#pragma unroll
    for (int e = 0; e < 5; e++)
        if (e < n) p = p || __double_as_longlong(v[e]) == kLuPending;
    return p;
}
// re-read n values (stride nd) until none is the pending pattern; bounded
__device__ __forceinline__ void lu_settle(double* v, int n, const double* base, long nd, int* err) {
    int spins = 0;
    while (lu_any_pending(v, n)) {
        if (++spins > kLuSpinMax || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int e = 0; e < 5; e++)
            if (e < n) v[e] = lu_ldc(base + e * nd);
    }
}

// Dataflow face exchange of a block (sweep-order block coordinates bi, bj, bk): value e of the
// products it hands to block bi + 1 (lane (jj, kk)), bj + 1 ((ii, kk)) and bk + 1 ((ii, jj)).
template <int CI> struct LuFace {
    static constexpr int NI = 64, NJ = 8 * CI, NK = 8 * CI, N = NI + NJ + NK;
    __device__ static long at(long blk, int e, int face, int f) {
        return (blk * 5 + e) * N + (face == 0 ? 0 : face == 1 ? NI : NI + NJ) + f;
    }
};

template <bool CART, bool LOWER, int CI, typename T = double, bool FLOW = false, bool STEP = false>
__device__ __forceinline__ void lu_block(const KZ& z, const LuArgs& a, int bi, int bj, int bk,
                                         LuLds<CART, CI, T>& S, int nbi = 0, int nbj = 0, int nbk = 0) {
    constexpr int NV = LuLds<CART, CI, T>::NV;
    constexpr bool F64 = sizeof(T) == sizeof(double);
    const int tid = threadIdx.x;
    const int ni = a.i1 - a.i0 + 1, nj = a.j1 - a.j0 + 1, nk = a.k1 - a.k0 + 1;
    const int oi0 = bi * CI;
    const int ci = ni - oi0 < CI ? ni - oi0 : CI;
    const int oj0 = bj * 8, ok0 = bk * 8;
    const long nd = z.nd;
    const LUC<T> c = lu_const<T>(z);
    const double signe = LOWER ? 0.5 : -0.5;
    const double* __restrict__ rop = a.rop;
    const double* __restrict__ ropk = LOWER ? a.rop : a.rop_k;
    double* __restrict__ dw = a.dw;
    // sweep-order offsets -> zone indices
    auto I = [&](int o) { return LOWER ? a.i0 + o : a.i1 - o; };
    auto J = [&](int o) { return LOWER ? a.j0 + o : a.j1 - o; };
    auto K = [&](int o) { return LOWER ? a.k0 + o : a.k1 - o; };

    // 1a. stage the block
    for (int e = tid; e < 64 * CI; e += 256) {
        const int L = e / CI, ii = e % CI;
        const int oj = oj0 + (L & 7), ok = ok0 + (L >> 3);
        if (oj >= nj || ok >= nk || ii >= ci) continue;
        const int i = I(oi0 + ii), j = J(oj), k = K(ok);
        const long l = IDX(z, i, j, k);
        // the sweep's divisor, sign products and scaled right-hand side, formed here off the
        // wavefront's critical path (the same operations as in the solve)
        const double cf4 = z.coe[l + 4 * nd];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            S.st[v][ii][L] = (T)rop[l + v * nd];
            if (v < 4) S.st[5 + v][ii][L] = (T)(z.coe[l + v * nd] * signe);
            S.st[10 + v][ii][L] = (T)(LOWER ? dw[l + v * nd] : dw[l + v * nd] * cf4);
        }
        S.st[9][ii][L] = (T)(1. / cf4);
        if (!CART) {
            const long lt = MTR(z, i, j, k);
            double tx, ty, tz;
            lu_normal<CART, 0>(z, lt, tx, ty, tz);
            S.st[15][ii][L] = (T)tx; S.st[16][ii][L] = (T)ty; S.st[17][ii][L] = (T)tz;
            lu_normal<CART, 1>(z, lt, tx, ty, tz);
            S.st[18][ii][L] = (T)tx; S.st[19][ii][L] = (T)ty; S.st[20][ii][L] = (T)tz;
            lu_normal<CART, 2>(z, lt, tx, ty, tz);
            S.st[21][ii][L] = (T)tx; S.st[22][ii][L] = (T)ty; S.st[23][ii][L] = (T)tz;
        }
    }
    // step-by-step halo hand-over (FP64 sweeps only)
    constexpr bool FSTEP = FLOW && STEP && F64;
    // 1b. products of the halo cells (finished by earlier launches; dataflow: handed over by the
    //     neighbour blocks through the face exchange).  With FSTEP the block
    //     fetches what is there and the solver lane waits for a value still pending at the step
    //     that uses it, so a block starts while the blocks before it are still running; otherwise
    //     the block waits here for its whole halo.  Every slot read is stored back as pending (by
    //     the solver lane at its step, or here), so the exchange is ready for the next sweep
    //     without a fill.
    for (int e = tid; e < 64 + 16 * CI; e += 256) {
        T b[5];
        if (FLOW) {
            double v[5];
            long base;
            if (e < 64) {
                const int L = e, oj = oj0 + (L & 7), ok = ok0 + (L >> 3);
                if (oi0 == 0 || oj >= nj || ok >= nk) continue;
                base = LuFace<CI>::at(bi - 1 + (long)nbi * (bj + (long)nbj * bk), 0, 0, L);
            } else if (e < 64 + 8 * CI) {
                const int q = e - 64, ii = q / 8, kk = q % 8, ok = ok0 + kk;
                if (oj0 == 0 || ii >= ci || ok >= nk) continue;
                base = LuFace<CI>::at(bi + (long)nbi * (bj - 1 + (long)nbj * bk), 0, 1, ii * 8 + kk);
            } else {
                const int q = e - 64 - 8 * CI, ii = q / 8, jj = q % 8, oj = oj0 + jj;
                if (ok0 == 0 || ii >= ci || oj >= nj) continue;
                base = LuFace<CI>::at(bi + (long)nbi * (bj + (long)nbj * (bk - 1)), 0, 2, ii * 8 + jj);
            }
#pragma unroll
            for (int v2 = 0; v2 < 5; v2++) v[v2] = lu_ldc(a.fx + base + (long)v2 * LuFace<CI>::N);
            if (!FSTEP) {
                lu_settle(v, 5, a.fx + base, LuFace<CI>::N, a.sync + 1);
#pragma unroll
                for (int v2 = 0; v2 < 5; v2++) lu_stc(a.fx + base + (long)v2 * LuFace<CI>::N, __longlong_as_double(kLuPending));
            }
            if (e < 64) {
#pragma unroll
                for (int v2 = 0; v2 < 5; v2++) S.hI[v2][e] = (T)v[v2];
            } else if (e < 64 + 8 * CI) {
                const int q = e - 64;
#pragma unroll
                for (int v2 = 0; v2 < 5; v2++) S.hJ[v2][q / 8][q % 8] = (T)v[v2];
            } else {
                const int q = e - 64 - 8 * CI;
#pragma unroll
                for (int v2 = 0; v2 < 5; v2++) S.hK[v2][q / 8][q % 8] = (T)v[v2];
            }
            continue;
        }
        if (e < 64) {
            const int L = e, oj = oj0 + (L & 7), ok = ok0 + (L >> 3);
            if (oi0 == 0 || oj >= nj || ok >= nk) continue;
            lu_prod_global<CART, 0, T>(z, c, rop, dw, I(oi0 - 1), J(oj), K(ok), signe, b);
#pragma unroll
            for (int v = 0; v < 5; v++) S.hI[v][L] = b[v];
        } else if (e < 64 + 8 * CI) {
            const int q = e - 64, ii = q / 8, kk = q % 8, ok = ok0 + kk;
            if (oj0 == 0 || ii >= ci || ok >= nk) continue;
            lu_prod_global<CART, 1, T>(z, c, rop, dw, I(oi0 + ii), J(oj0 - 1), K(ok), signe, b);
#pragma unroll
            for (int v = 0; v < 5; v++) S.hJ[v][ii][kk] = b[v];
        } else {
            const int q = e - 64 - 8 * CI, ii = q / 8, jj = q % 8, oj = oj0 + jj;
            if (ok0 == 0 || ii >= ci || oj >= nj) continue;
            lu_prod_global<CART, 2, T>(z, c, ropk, dw, I(oi0 + ii), J(oj), K(ok0 - 1), signe, b);
#pragma unroll
            for (int v = 0; v < 5; v++) S.hK[v][ii][jj] = b[v];
        }
    }
    __syncthreads();

    // 2. skewed wavefront.  At step s, wave 0 solves cell ii = s - jj - kk of every line;
    //    then waves 1, 2, 3 form that cell's products for its i, j and k consumers, which
    //    wave 0 reads at step s + 1 (its own lane for i, the lane above in j or in k).  The
    //    two barriers per step replace one wave doing all of it: the critical path of a step
    //    is the solve plus one product instead of the solve plus three products.
    const int wv = tid >> 6;
    const int L = tid & 63, jj = L & 7, kk = L >> 3;
    const int oj = oj0 + jj, ok = ok0 + kk;
    const bool lane_ok = oj < nj && ok < nk;
    const int j = J(oj), k = K(ok);
    const bool kalias = !LOWER && a.rop_k != a.rop;
    // the Cartesian face normals are read once: inside the loop the compiler could not keep
    // them in registers across the stores to dw
    const T tci = (T)(CART ? z.ti[0] : 0.), tcj = (T)(CART ? z.tj[0] : 0.), tck = (T)(CART ? z.tk[0] : 0.);
    const int nsteps = ci + 14;
    // Newton update of the cell solved at a step: before the first barrier when the k products
    // read the updated state (in-place sub-iterations), else after it, beside the product waves
    const bool newton = !LOWER && a.newton;
    auto newton_update = [&](long l, const double x[5], double r0, double u0, double v0, double w0,
                             double t0) {
        double* __restrict__ O = a.out;
        const double cvinv = 1. / z.cv, cvinv2 = 0.5 * cvinv;
        const double r1 = r0 + x[0];
        const double r_1 = 1. / r1;
        const double u1 = (r0 * u0 + x[1]) * r_1, v1 = (r0 * v0 + x[2]) * r_1, w1 = (r0 * w0 + x[3]) * r_1;
        const double roe_old = r0 * (z.cv * t0 + 0.5 * (u0 * u0 + v0 * v0 + w0 * w0));
        const double t1 = (roe_old + x[4]) * r_1 * cvinv - cvinv2 * (u1 * u1 + v1 * v1 + w1 * w1);
        O[l] = r1; O[l + nd] = u1; O[l + 2 * nd] = v1; O[l + 3 * nd] = w1; O[l + 4 * nd] = t1;
        if (kalias) { S.wk[0][L] = (T)u1; S.wk[1][L] = (T)v1; S.wk[2][L] = (T)w1; S.wk[3][L] = (T)t1; }
    };
    // the product a wave forms: B_D of the cell it handles, from the staged state and coefficients
    auto coef = [&](int ii, T uu, T vv, T ww, T tt) {
        T tx, ty, tz;
        if (wv == 1) {
            if (CART) { tx = tci; ty = tz = (T)0.; }
            else { tx = S.st[15][ii][L]; ty = S.st[16][ii][L]; tz = S.st[17][ii][L]; }
            return lu_coef<CART, 0, T>(c, uu, vv, ww, tt, S.st[6][ii][L], tx, ty, tz);
        } else if (wv == 2) {
            if (CART) { tx = tcj; ty = tz = (T)0.; }
            else { tx = S.st[18][ii][L]; ty = S.st[19][ii][L]; tz = S.st[20][ii][L]; }
            return lu_coef<CART, 1, T>(c, uu, vv, ww, tt, S.st[7][ii][L], tx, ty, tz);
        }
        if (CART) { tx = tck; ty = tz = (T)0.; }
        else { tx = S.st[21][ii][L]; ty = S.st[22][ii][L]; tz = S.st[23][ii][L]; }
        return lu_coef<CART, 2, T>(c, uu, vv, ww, tt, S.st[8][ii][L], tx, ty, tz);
    };
    for (int s = 0; s < nsteps; s++) {
        const int ii = s - jj - kk;
        const bool act = lane_ok && ii >= 0 && ii < ci;
        T x[5];
        double xd[5], r0 = 0., u0 = 0., v0 = 0., w0 = 0., t0 = 0.;
        long l = 0;
        // the product waves form their B matrix while wave 0 solves (it does not depend on the
        // solution; with in-place sub-iterations the k products wait for the updated state)
        LuB<T> cb;
        if (wv > 0 && act && !(wv == 3 && kalias))
            cb = coef(ii, S.st[1][ii][L], S.st[2][ii][L], S.st[3][ii][L], S.st[4][ii][L]);
        if (wv == 0 && act) {
            const int oi = oi0 + ii;
            l = IDX(z, I(oi), j, k);
            // every LDS operand of the solve is read before the first use (one round trip);
            // neighbour terms in the reference order: lower 3dcart J,K,I; lower 3dfull and
            // upper K,J,I
            const T xal = S.st[5][ii][L];
            const T diag = S.st[9][ii][L];
#pragma unroll
            for (int e = 0; e < 5; e++) x[e] = S.st[10 + e][ii][L];
            T bn[3][5];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const int D = (LOWER && CART) ? (q == 0 ? 1 : q == 1 ? 2 : 0) : (q == 0 ? 2 : q == 1 ? 1 : 0);
                if (D == 0) {
#pragma unroll
                    for (int e = 0; e < 5; e++) bn[q][e] = ii == 0 ? S.hI[e][L] : S.px[0][e][L];
                } else if (D == 1) {
#pragma unroll
                    for (int e = 0; e < 5; e++) bn[q][e] = jj == 0 ? S.hJ[e][ii][kk] : S.px[1][e][L - 1];
                } else {
#pragma unroll
                    for (int e = 0; e < 5; e++) bn[q][e] = kk == 0 ? S.hK[e][ii][jj] : S.px[2][e][L - 8];
                }
            }
            if (newton) {
                // the Newton update works on the double state: the staged copy (FP64 sweeps)
                // or global memory (mixed precision)
                if (F64) {
                    r0 = (double)S.st[0][ii][L]; u0 = (double)S.st[1][ii][L]; v0 = (double)S.st[2][ii][L];
                    w0 = (double)S.st[3][ii][L]; t0 = (double)S.st[4][ii][L];
                } else {
                    r0 = rop[l]; u0 = rop[l + nd]; v0 = rop[l + 2 * nd]; w0 = rop[l + 3 * nd]; t0 = rop[l + 4 * nd];
                }
            }
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const int D = (LOWER && CART) ? (q == 0 ? 1 : q == 1 ? 2 : 0) : (q == 0 ? 2 : q == 1 ? 1 : 0);
                const int o = D == 0 ? oi : D == 1 ? oj : ok;
                if (o == 0) continue;
                T* b = bn[q];
                if (FSTEP && (D == 0 ? ii == 0 : D == 1 ? jj == 0 : kk == 0)) {
                    // a product of the previous block in D: wait for it if it was still pending
                    // when the block was staged, then give the slot back
                    const long nb = D == 0 ? bi - 1 + (long)nbi * (bj + (long)nbj * bk)
                                  : D == 1 ? bi + (long)nbi * (bj - 1 + (long)nbj * bk)
                                           : bi + (long)nbi * (bj + (long)nbj * (bk - 1));
                    const long base = LuFace<CI>::at(nb, 0, D, D == 0 ? L : D == 1 ? ii * 8 + kk : ii * 8 + jj);
                    double v[5];
#pragma unroll
                    for (int e = 0; e < 5; e++) v[e] = (double)b[e];
                    if (lu_any_pending(v, 5)) {
#pragma unroll
                        for (int e = 0; e < 5; e++) v[e] = lu_ldc(a.fx + base + (long)e * LuFace<CI>::N);
                        lu_settle(v, 5, a.fx + base, LuFace<CI>::N, a.sync + 1);
#pragma unroll
                        for (int e = 0; e < 5; e++) b[e] = (T)v[e];
                    }
#pragma unroll
                    for (int e = 0; e < 5; e++) lu_stc(a.fx + base + (long)e * LuFace<CI>::N, __longlong_as_double(kLuPending));
                }
#pragma unroll
                for (int e = 0; e < 5; e++) x[e] = x[e] + b[e] * xal;
            }
#pragma unroll
            for (int e = 0; e < 5; e++) {
                x[e] = x[e] * diag; xd[e] = (double)x[e]; dw[l + e * nd] = xd[e]; S.xs[e][L] = x[e];
            }
            if (newton && kalias) newton_update(l, xd, r0, u0, v0, w0, t0);
        }
        __syncthreads();
        if (wv == 0 && act && newton && !kalias) newton_update(l, xd, r0, u0, v0, w0, t0);
        if (wv > 0 && act) {
            // products of the cell just solved for its three consumers
            T b[5];
#pragma unroll
            for (int e = 0; e < 5; e++) x[e] = S.xs[e][L];
            if (wv == 3 && kalias) cb = coef(ii, S.wk[0][L], S.wk[1][L], S.wk[2][L], S.wk[3][L]);
            if (wv == 1) lu_apply<CART, 0, T>(cb, x, b);
            else if (wv == 2) lu_apply<CART, 1, T>(cb, x, b);
            else lu_apply<CART, 2, T>(cb, x, b);
#pragma unroll
            for (int e = 0; e < 5; e++) S.px[wv - 1][e][L] = b[e];
            if (FLOW) {
                // products whose consumer is in the next block: hand them over
                long blk = -1;
                int face = wv - 1, f = 0;
                if (wv == 1 && ii == ci - 1 && bi + 1 < nbi) { blk = bi + (long)nbi * (bj + (long)nbj * bk); f = L; }
                if (wv == 2 && jj == 7 && bj + 1 < nbj) { blk = bi + (long)nbi * (bj + (long)nbj * bk); f = ii * 8 + kk; }
                if (wv == 3 && kk == 7 && bk + 1 < nbk) { blk = bi + (long)nbi * (bj + (long)nbj * bk); f = ii * 8 + jj; }
                if (blk >= 0) {
#pragma unroll
                    for (int e = 0; e < 5; e++) lu_stc(a.fx + LuFace<CI>::at(blk, e, face, f), (double)b[e]);
                }
            }
        }
        __syncthreads();
    }
    (void)NV;
}

template <bool CART, bool LOWER, int CI, typename T = double>
__global__ __launch_bounds__(256) void k_lu_block(KZ z, LuArgs a, int diag, int bk_lo, int nbi, int nbj) {
    const int bj = blockIdx.x, bk = bk_lo + blockIdx.y, bi = diag - bj - bk;
    if (bi < 0 || bi >= nbi || bj >= nbj) return;
    __shared__ LuLds<CART, CI, T> S;
    lu_block<CART, LOWER, CI, T>(z, a, bi, bj, bk, S);
}

// Dataflow schedule: one launch per sweep.  Workgroups take blocks from a counter in diagonal
// order (a block only waits on blocks handed out before it, so there is no deadlock whatever the
// dispatch order), stage their own fields at once, read the products of their halo cells from
// the face exchange as the neighbour blocks hand them over, and hand their own face products on.
// Launch diagonals are no longer barriers: a block starts as soon as its three neighbours are done.
template <bool CART, bool LOWER, int CI, typename T = double, bool STEP = false>
__global__ __launch_bounds__(256) void k_lu_flow(KZ z, LuArgs a, int nbi, int nbj, int nbk) {
    __shared__ LuLds<CART, CI, T> S;
    __shared__ int tk;
    for (;;) {
        if (threadIdx.x == 0) tk = atomicAdd(a.sync, 1);
        __syncthreads();
        const int t = tk;
        __syncthreads();
        if (t >= a.flow_n) break;
        const LuWork w = a.flow_work[t];
        lu_block<CART, LOWER, CI, T, true, STEP>(z, a, w.bi, w.bj, w.bk, S, nbi, nbj, nbk);
        __syncthreads();
    }
}

// Dataflow sweeps of several zones in one launch: the blocks of every zone's LU box in one list
// ordered by block diagonal, handed out from one counter.  A block only waits on blocks of its own
// zone on earlier diagonals, which were handed out before it; each zone has its own face exchange
// and error word.
template <bool CART, bool LOWER, int CI>
__global__ __launch_bounds__(256) void k_lu_flow_list(const LuZone* __restrict__ zones,
                                                      const LuWork* __restrict__ work, int nwork, int* counter) {
    __shared__ LuLds<CART, CI, double> S;
    __shared__ int tk;
    for (;;) {
        if (threadIdx.x == 0) tk = atomicAdd(counter, 1);
        __syncthreads();
        const int t = tk;
        __syncthreads();
        if (t >= nwork) break;
        const LuWork w = work[t];
        const KZ z = zones[w.zone].z;
        const LuArgs a = zones[w.zone].a;
        const int nbi = (a.i1 - a.i0 + CI) / CI, nbj = (a.j1 - a.j0 + 8) / 8, nbk = (a.k1 - a.k0 + 8) / 8;
        lu_block<CART, LOWER, CI, double, true>(z, a, w.bi, w.bj, w.bk, S, nbi, nbj, nbk);
        __syncthreads();
    }
}

// Several zones in one launch per block diagonal: block b of the launch takes its zone and its
// block coordinates from the work list (zones are independent: each keeps its own diagonal
// order, diagonal d of every zone goes into launch d).
template <bool CART, bool LOWER, int CI, typename T = double>
__global__ __launch_bounds__(256) void k_lu_block_list(const LuZone* __restrict__ zones,
                                                       const LuWork* __restrict__ work) {
    const LuWork w = work[blockIdx.x];
    __shared__ LuLds<CART, CI, T> S;
    const KZ z = zones[w.zone].z;
    const LuArgs a = zones[w.zone].a;
    lu_block<CART, LOWER, CI, T>(z, a, w.bi, w.bj, w.bk, S);
}

// ---------------------------------------------------------------------------------------
// Pipelined schedule: one persistent wave per 8 x 8 tile of (j, k) lines, each wave sweeping
// its full i-lines in the skewed order above.  A tile needs, at its step s, cells that the
// tiles before it in j and k finish at their step s + 7, so each tile simply trails its two
// producers by about eight steps and the whole sweep takes about ni + nj + nk steps instead
// of one launch per block diagonal.
//
// Cross-tile values travel through a small exchange buffer and validate themselves: before a
// sweep the buffer is filled with a signalling-NaN pattern that arithmetic never produces; the
// lanes on a tile's last j (k) line store their results there with agent-scope stores (seen by
// every XCD), and the consumer tile re-reads them with agent-scope loads until they are no longer
// the pattern.  Every other result is an ordinary store: only 2 of the 15 lines of a tile cross
// to another tile, so the write-through traffic is small.  Every wait is bounded: past the
// bound, or once any wave has given up, the error word is set and the kernel runs to its end
// with whatever it read, so the grid always drains; the host then reports the failure.
// Tiles are taken from a counter in dependency order (by j-tile + k-tile), so a tile only
// waits on tiles that started before it: no deadlock whatever the dispatch order.

// ticket -> tile, enumerating the tiles by diagonal jt + kt
__device__ __forceinline__ void lu_tile_of(int t, int ntj, int ntk, int& jt, int& kt) {
    for (int d = 0; d < ntj + ntk - 1; d++) {
        int lo = d - (ntj - 1) > 0 ? d - (ntj - 1) : 0, hi = d < ntk - 1 ? d : ntk - 1;
        int cnt = hi - lo + 1;
        if (t < cnt) { kt = lo + t; jt = d - kt; return; }
        t -= cnt;
    }
    jt = kt = 0;
}

// Everything one step of a lane reads from global memory, fetched P steps ahead so that the
// memory latency overlaps the arithmetic of the steps in between.  The neighbour values of
// another tile may still be the pending pattern when fetched; they are re-read at use.
template <bool CART>
struct LuSlot {
    double w[5], cf[5], in[5];            // own state, coefficients, right-hand side
    double rj[5], wj[4], cj;              // j neighbour in the previous tile: result, state, coe
    double rk[5], wk[4], ck;              // k neighbour in the previous tile
    double mo[CART ? 1 : 18];             // 3dfull: own faces (D, component, low/high)
    double mj[CART ? 1 : 6], mk[CART ? 1 : 6];
};



// Exchange layout: for the j boundary of tile row tj, value e of cell (oi, ok) (sweep offsets)
// at xj[e * NXJ + (tj * nk + ok) * ni + oi]; for the k boundary, value e of (oi, oj) at
// xk[e * NXK + (tk * nj + oj) * ni + oi], e = 0..4 the result, 5..8 the Newton-updated
// u, v, w, T of an upper sweep in place.
long lu_exchange_size(int ni, int nj, int nk) {
    long ntj = (nj + 7) / 8, ntk = (nk + 7) / 8;
    return 5 * ntj * nk * (long)ni + 9 * ntk * nj * (long)ni;
}

template <bool CART, bool LOWER, int P>
__global__ __launch_bounds__(64) void k_lu_pipe(KZ z, LuArgs a, const double* __restrict__ in,
                                                double* __restrict__ res, int ntj, int ntk) {
    __shared__ double ex[2][2][5][64];
    __shared__ int s_tile;
    int* ticket = a.sync;
    int* err = a.sync + 1;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1);
    __syncthreads();
    int jt, kt;
    lu_tile_of(s_tile, ntj, ntk, jt, kt);
    const int L = threadIdx.x, jj = L & 7, kk = L >> 3;
    const int ni = a.i1 - a.i0 + 1, nj = a.j1 - a.j0 + 1, nk = a.k1 - a.k0 + 1;
    const int oj = jt * 8 + jj, ok = kt * 8 + kk;
    const bool lane_ok = oj < nj && ok < nk;
    const int j = LOWER ? a.j0 + oj : a.j1 - oj, k = LOWER ? a.k0 + ok : a.k1 - ok;
    const int sgn = LOWER ? -1 : 1;
    const long nd = z.nd, ndm = z.ndm;
    const LUC<> c = lu_const(z);
    const double signe = LOWER ? 0.5 : -0.5;
    const double* __restrict__ rop = a.rop;
    const bool kalias = !LOWER && a.rop_k != a.rop;
    const double* kstate = kalias ? a.out : (LOWER ? rop : a.rop_k);
    const bool has_j = jj == 0 && oj > 0, has_k = kk == 0 && ok > 0;
    // this lane's line feeds the next tile in j (k)
    const bool put_j = jj == 7 && oj + 1 < nj, put_k = kk == 7 && ok + 1 < nk;
    const long NXJ = (long)ntj * nk * ni, NXK = (long)ntk * nj * ni;
    double* const xj = a.xch;
    double* const xk = a.xch + 5 * NXJ;
    const long xjin = ((long)(jt - 1) * nk + ok) * ni, xjout = ((long)jt * nk + ok) * ni;
    const long xkin = ((long)(kt - 1) * nj + oj) * ni, xkout = ((long)kt * nj + oj) * ni;
    const double tci = CART ? z.ti[0] : 0., tcj = CART ? z.tj[0] : 0., tck = CART ? z.tk[0] : 0.;
    const double* tm[3] = {z.ti, z.tj, z.tk};
    const long minc[3] = {z.mi, z.mj, z.mk};

    auto fetch = [&](LuSlot<CART>& q, int s) {
        const int ii = s - jj - kk;
        if (!(lane_ok && ii >= 0 && ii < ni)) return;
        const int i = LOWER ? a.i0 + ii : a.i1 - ii;
        const long l = IDX(z, i, j, k);
#pragma unroll
        for (int e = 0; e < 5; e++) { q.w[e] = rop[l + e * nd]; q.cf[e] = z.coe[l + e * nd]; q.in[e] = in[l + e * nd]; }
        if (!CART) {
            const long lt = MTR(z, i, j, k);
#pragma unroll
            for (int D = 0; D < 3; D++)
#pragma unroll
                for (int cc = 0; cc < 3; cc++) {
                    q.mo[6 * D + 2 * cc] = tm[D][lt + cc * ndm];
                    q.mo[6 * D + 2 * cc + 1] = tm[D][lt + minc[D] + cc * ndm];
                }
        }
        if (has_j) {
            const long l1 = IDX(z, i, j + sgn, k);
#pragma unroll
            for (int e = 0; e < 5; e++) q.rj[e] = lu_ldc(xj + xjin + ii + e * NXJ);
#pragma unroll
            for (int e = 0; e < 4; e++) q.wj[e] = rop[l1 + (e + 1) * nd];
            q.cj = z.coe[l1 + 2 * nd];
            if (!CART) {
                const long lt1 = MTR(z, i, j + sgn, k);
#pragma unroll
                for (int cc = 0; cc < 3; cc++) { q.mj[2 * cc] = z.tj[lt1 + cc * ndm]; q.mj[2 * cc + 1] = z.tj[lt1 + z.mj + cc * ndm]; }
            }
        }
        if (has_k) {
            const long l1 = IDX(z, i, j, k + sgn);
#pragma unroll
            for (int e = 0; e < 5; e++) q.rk[e] = lu_ldc(xk + xkin + ii + e * NXK);
#pragma unroll
            for (int e = 0; e < 4; e++) q.wk[e] = kalias ? lu_ldc(xk + xkin + ii + (5 + e) * NXK) : kstate[l1 + (e + 1) * nd];
            q.ck = z.coe[l1 + 3 * nd];
            if (!CART) {
                const long lt1 = MTR(z, i, j, k + sgn);
#pragma unroll
                for (int cc = 0; cc < 3; cc++) { q.mk[2 * cc] = z.tk[lt1 + cc * ndm]; q.mk[2 * cc + 1] = z.tk[lt1 + z.mk + cc * ndm]; }
            }
        }
    };

    double pi[5] = {0., 0., 0., 0., 0.};
    auto body = [&](LuSlot<CART>& q, int s) {
        const int ii = s - jj - kk;
        const int rd = (s + 1) & 1, wr = s & 1;
        if (!(lane_ok && ii >= 0 && ii < ni)) return;
        const int i = LOWER ? a.i0 + ii : a.i1 - ii;
        const long l = IDX(z, i, j, k);
        const double r0 = q.w[0], u0 = q.w[1], v0 = q.w[2], w0 = q.w[3], t0 = q.w[4];
        const double cf4 = q.cf[4];
        const double xal = q.cf[0] * signe;
        double x[5], b[5];
#pragma unroll
        for (int e = 0; e < 5; e++) x[e] = LOWER ? q.in[e] : q.in[e] * cf4;
#pragma unroll
        for (int qq = 0; qq < 3; qq++) {
            const int D = (LOWER && CART) ? (qq == 0 ? 1 : qq == 1 ? 2 : 0) : (qq == 0 ? 2 : qq == 1 ? 1 : 0);
            const int o = D == 0 ? ii : D == 1 ? oj : ok;
            if (o == 0) continue;
            if (D == 0) {
#pragma unroll
                for (int e = 0; e < 5; e++) b[e] = pi[e];
            } else if (D == 1) {
                if (jj == 0) {
                    lu_settle(q.rj, 5, xj + xjin + ii, NXJ, err);
                    double tx = 0., ty = 0., tz = 0.;
                    if (CART) tx = tcj;
                    else { tx = 0.5 * (q.mj[1] + q.mj[0]); ty = 0.5 * (q.mj[3] + q.mj[2]); tz = 0.5 * (q.mj[5] + q.mj[4]); }
                    lu_prod<CART, 1>(c, q.wj[0], q.wj[1], q.wj[2], q.wj[3], q.cj * signe, tx, ty, tz, q.rj, b);
                } else {
#pragma unroll
                    for (int e = 0; e < 5; e++) b[e] = ex[rd][0][e][L - 1];
                }
            } else {
                if (kk == 0) {
                    lu_settle(q.rk, 5, xk + xkin + ii, NXK, err);
                    if (kalias) lu_settle(q.wk, 4, xk + xkin + ii + 5 * NXK, NXK, err);
                    double tx = 0., ty = 0., tz = 0.;
                    if (CART) tx = tck;
                    else { tx = 0.5 * (q.mk[1] + q.mk[0]); ty = 0.5 * (q.mk[3] + q.mk[2]); tz = 0.5 * (q.mk[5] + q.mk[4]); }
                    lu_prod<CART, 2>(c, q.wk[0], q.wk[1], q.wk[2], q.wk[3], q.ck * signe, tx, ty, tz, q.rk, b);
                } else {
#pragma unroll
                    for (int e = 0; e < 5; e++) b[e] = ex[rd][1][e][L - 8];
                }
            }
#pragma unroll
            for (int e = 0; e < 5; e++) x[e] = x[e] + b[e] * xal;
        }
        const double diag = 1. / cf4;
#pragma unroll
        for (int e = 0; e < 5; e++) {
            x[e] = x[e] * diag;
            res[l + e * nd] = x[e];
            if (put_j) lu_stc(xj + xjout + ii + e * NXJ, x[e]);
            if (put_k) lu_stc(xk + xkout + ii + e * NXK, x[e]);
        }
        double uk = u0, vk = v0, wk = w0, tk = t0;
        if (!LOWER && a.newton) {
            double* O = a.out;
            const double cvinv = 1. / z.cv, cvinv2 = 0.5 * cvinv;
            const double r1 = r0 + x[0];
            const double r_1 = 1. / r1;
            const double u1 = (r0 * u0 + x[1]) * r_1, v1 = (r0 * v0 + x[2]) * r_1, w1 = (r0 * w0 + x[3]) * r_1;
            const double roe_old = r0 * (z.cv * t0 + 0.5 * (u0 * u0 + v0 * v0 + w0 * w0));
            const double t1 = (roe_old + x[4]) * r_1 * cvinv - cvinv2 * (u1 * u1 + v1 * v1 + w1 * w1);
            O[l] = r1; O[l + nd] = u1; O[l + 2 * nd] = v1; O[l + 3 * nd] = w1; O[l + 4 * nd] = t1;
            if (kalias) {
                uk = u1; vk = v1; wk = w1; tk = t1;
                if (put_k) {
                    lu_stc(xk + xkout + ii + 5 * NXK, u1); lu_stc(xk + xkout + ii + 6 * NXK, v1);
                    lu_stc(xk + xkout + ii + 7 * NXK, w1); lu_stc(xk + xkout + ii + 8 * NXK, t1);
                }
            }
        }
        double tx = 0., ty = 0., tz = 0.;
        if (CART) tx = tci;
        else { tx = 0.5 * (q.mo[1] + q.mo[0]); ty = 0.5 * (q.mo[3] + q.mo[2]); tz = 0.5 * (q.mo[5] + q.mo[4]); }
        lu_prod<CART, 0>(c, u0, v0, w0, t0, q.cf[1] * signe, tx, ty, tz, x, pi);
        if (CART) tx = tcj;
        else { tx = 0.5 * (q.mo[7] + q.mo[6]); ty = 0.5 * (q.mo[9] + q.mo[8]); tz = 0.5 * (q.mo[11] + q.mo[10]); }
        lu_prod<CART, 1>(c, u0, v0, w0, t0, q.cf[2] * signe, tx, ty, tz, x, b);
#pragma unroll
        for (int e = 0; e < 5; e++) ex[wr][0][e][L] = b[e];
        if (CART) tx = tck;
        else { tx = 0.5 * (q.mo[13] + q.mo[12]); ty = 0.5 * (q.mo[15] + q.mo[14]); tz = 0.5 * (q.mo[17] + q.mo[16]); }
        lu_prod<CART, 2>(c, uk, vk, wk, tk, q.cf[3] * signe, tx, ty, tz, x, b);
#pragma unroll
        for (int e = 0; e < 5; e++) ex[wr][1][e][L] = b[e];
    };

    const int nsteps = ni + 14;
    LuSlot<CART> sl[P];
#pragma unroll
    for (int q = 0; q < P; q++) fetch(sl[q], q);
    for (int s0 = 0; s0 < nsteps; s0 += P) {
#pragma unroll
        for (int q = 0; q < P; q++) {
            const int s = s0 + q;
            if (s < nsteps) {
                body(sl[q], s);
                fetch(sl[q], s + P);
            }
            // the next step reads what the neighbouring lanes of this wave wrote to LDS
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
    }
}

// ---------------------------------------------------------------------------------------
// Streaming schedule (LU_STREAM, Cartesian zones).  One wave per pencil of 8 x 8 (j, k) lines,
// sweeping the whole i extent of the LU box as a skewed wavefront: lane (jj, kk) handles cell
// ii = s - jj - kk of its line at step s of the lower sweep, and ii = R - 1 - s - jj - kk at step
// s of the upper sweep (R = ni + 14), so the cell it needs from the line before it in j (k) was
// finished one step earlier by lane L -/+ 1 (L -/+ 8).
//
//  * The wave solves its 64 cells (the reference's neighbour terms and division) and forms the
//    three products of each for its consumers in i, j and k: the i product stays in the lane's
//    registers, the j product moves one lane by a DPP row shift, the k product eight lanes by
//    ds_bpermute.  No barrier: the wave runs the pencil alone.
//  * Inputs come from a skewed, pencil-major copy made before the sweeps (k_lu_pack): row r of a
//    pencil holds, for each lane, the cell r - jj - kk of its line (right-hand side, state, the
//    coefficients: 16 doubles), so a step reads whole 512-byte rows, one per array, and every byte
//    fetched is used.  The wave loads the rows kLuStAhead steps ahead into registers.  The lower
//    sweep writes dW* back into the right-hand-side rows; the upper sweep reads it from there and
//    writes dW and the Newton-updated state into the rows it has read, which k_lu_unpack copies
//    to the zone arrays.
//  * Pencils hand their j-face (k-face) products to the next pencil in sweep order through a
//    global exchange, one value per (pencil, lane, variable, cell), with agent-scope stores; the
//    consumer lanes load them two steps ahead and re-read values that are still the pending
//    pattern (bounded waits, error word).  The exchange is filled with the pending pattern before
//    each sweep (one fill kernel: cheaper than agent-scope stores giving every slot back, whose
//    write-through holds up the loads issued after them).
//  * Pencils are taken from a counter in diagonal order (reversed for the upper sweep): a pencil
//    only waits on pencils handed out before it, so there is no deadlock whatever the dispatch
//    order.
// Every cell runs the expressions of k_lu_block in the same order, so the results are the
// reference's bit for bit.

// doubles per row and lane of the skewed copy: right-hand side (dW* after the lower sweep) 0-4,
// state 5-9, coe1..4 / 2 10-13 (the sweep's sign products up to the sign, formed by the same
// multiplications), 1 / coe5 14, coe5 15; on 3dfull zones the mean face normals of the i, j and
// k products 16-24 (lu_normal)
template <bool CART> constexpr int lu_skf() { return CART ? 16 : 25; }
__device__ __forceinline__ int lu_skf_of(const KZ& z) { return z.typ == 2 ? 16 : 25; }
// 1: the B matrices of a step are formed one step ahead (while the step before waits), 0: at the
// step (as fast at 256^3, and fewer instructions)
#ifndef FM_LU_ST_COEF_AHEAD
#define FM_LU_ST_COEF_AHEAD 0
#endif
// rows in flight per wave
#ifndef FM_LU_ST_AHEAD
#define FM_LU_ST_AHEAD 4
#endif
constexpr int kLuStAhead = FM_LU_ST_AHEAD;
// steps of slack a consumer lane builds up once it had to wait for the previous pencil
#ifndef FM_LU_ST_SLACK
#define FM_LU_ST_SLACK 1
#endif
constexpr int kLuStSlack = FM_LU_ST_SLACK;
// s_sleep arguments (units of 64 cycles) of the wait for the slack step and of the re-reads
#ifndef FM_LU_ST_SLEEP_SLACK
#define FM_LU_ST_SLEEP_SLACK 4
#endif
#ifndef FM_LU_ST_SLEEP_RETRY
#define FM_LU_ST_SLEEP_RETRY 2
#endif
// steps ahead at which a lane loads the products of the previous pencil (divides kLuStAhead)
constexpr int kLuStHalo = 2;

__device__ __forceinline__ long lu_sk_row(int t, int r, int R, int nfld) { return ((long)t * R + r) * nfld * 64; }

long lu_skew_size(bool cart, int ni, int nj, int nk) {
    long nt = (long)((nj + 7) / 8) * ((nk + 7) / 8);
    return nt * (ni + 14) * (cart ? 16 : 25) * 64;
}

// arrays the pack copies, in launch order: the right-hand side, the state (unless a.w_ready), the
// coefficients, the normals (3dfull, unless a.n_ready)

// exchange slot of value e of cell ii of lane l8 on the j (face 0) or k (face 1) side of the
// consuming pencil t.  The eight lanes of a face handle cells with the same ii + l8 at one step
// (in either sweep), so the slots are ordered by ii + l8, then value, then lane: a step reads and
// writes 64 contiguous bytes per value.
__device__ __forceinline__ long lu_st_slot(int face, int t, int l8, int e, int ii, int ntile, int ni) {
    return ((((long)face * ntile + t) * (ni + 7) + ii + l8) * 5 + e) * 8 + l8;
}

__host__ __device__ long lu_stream_exchange_size(int ni, int nj, int nk) {
    long nt = (long)((nj + 7) / 8) * ((nk + 7) / 8);
    return 2 * nt * 40 * (long)(ni + 7);
}

// the value of the lane before this one in j in sweep order: L - 1 (lower) or L + 1 (upper), by
// a row shift inside rows of 16 lanes (lanes on the first line of the pencil get a value they do
// not use: they take the j neighbour from the exchange)
template <bool LOWER>
__device__ __forceinline__ double lu_from_jprev(double v) {
    constexpr int ctl = LOWER ? 0x111 : 0x101;   // row_shr:1 / row_shl:1
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, ctl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), ctl, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// the value of the lane before this one in k in sweep order: L - 8 (lower) or L + 8 (upper);
// lanes on the first k line read their own value (they take the k neighbour from the exchange)
template <bool LOWER>
__device__ __forceinline__ double lu_from_kprev(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int src = (LOWER ? (lane >= 8 ? lane - 8 : lane) : (lane < 56 ? lane + 8 : lane)) << 2;
    const int lo = __builtin_amdgcn_ds_bpermute(src, (int)b);
    const int hi = __builtin_amdgcn_ds_bpermute(src, (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The skewed copy: workgroup (t, y, f) writes array f of rows [64 y, 64 y + 64) of pencil t.  Line
// L of the pencil contributes its cells r0 - d .. r0 - d + 63 (d = jj + kk) to those rows: a
// contiguous run per line, staged through LDS so that the row writes (along the lanes) are
// contiguous too; every thread issues all its loads before it waits for one.
__device__ __forceinline__ void lu_pack_rows(const KZ& z, const LuArgs& a, int ntj, int t, int r0, int f,
                                             double (*T)[65]) {
    const int jt = t % ntj, kt = t / ntj;
    const int ni = a.i1 - a.i0 + 1, nj = a.j1 - a.j0 + 1, nk = a.k1 - a.k0 + 1, R = ni + 14;
    const int nfld = lu_skf_of(z);
    const long nd = z.nd;
    const double* __restrict__ src = f < 5 ? a.dw + f * nd : f < 10 ? a.rop + (f - 5) * nd
                                   : z.coe + (f < 14 ? f - 10 : 4) * nd;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // wave w reads lines w, w + 4, ...: lane x takes the cell of row r0 + x
    double v[16];
#pragma unroll
    for (int m = 0; m < 16; m++) {
        const int L = w + 4 * m, oj = jt * 8 + (L & 7), ok = kt * 8 + (L >> 3);
        const int i = r0 + lane - (L & 7) - (L >> 3);
        const bool in_box = oj < nj && ok < nk && i >= 0 && i < ni;
        if (f < 16) {
            v[m] = in_box ? src[IDX(z, a.i0 + i, a.j0 + oj, a.k0 + ok)] : 0.;
        } else {
            // 3dfull: component f of the mean face normal of the product in direction D
            double n3[3] = {0., 0., 0.};
            if (in_box) {
                const long lt = MTR(z, a.i0 + i, a.j0 + oj, a.k0 + ok);
                const int D = (f - 16) / 3;
                if (D == 0) lu_normal<false, 0>(z, lt, n3[0], n3[1], n3[2]);
                else if (D == 1) lu_normal<false, 1>(z, lt, n3[0], n3[1], n3[2]);
                else lu_normal<false, 2>(z, lt, n3[0], n3[1], n3[2]);
            }
            const int cpt = (f - 16) % 3;
            v[m] = cpt == 0 ? n3[0] : cpt == 1 ? n3[1] : n3[2];
        }
    }
#pragma unroll
    for (int m = 0; m < 16; m++) {
        double x = v[m];
        if (f >= 10 && f < 14) x = x * 0.5;
        else if (f == 14) x = 1. / x;
        T[w + 4 * m][lane] = x;
    }
    __syncthreads();
    double* __restrict__ dst = a.skw + f * 64 + lane;
#pragma unroll
    for (int m = 0; m < 16; m++) {
        const int rr = w + 4 * m, r = r0 + rr;
        if (r < R) dst[lu_sk_row(t, r, R, nfld)] = T[lane][rr];
    }
}


// Pencil t and row block y of workgroup (x, y) of a (pencil, row block) grid.
// Consecutive workgroups go to the eight XCDs in turn; this order gives each XCD a contiguous
// range of (t, y) with y fastest (XCD c the c-th of eight near-equal parts), so that the row
// blocks of a pencil run one after another on one XCD and the cache lines a block's runs share
// with the next block's (the runs of a line start at r0 - jj - kk, not on a line boundary) are
// served from that XCD's L2.
#ifndef FM_LU_XCD_ORDER
#define FM_LU_XCD_ORDER 1
#endif
__device__ __forceinline__ void lu_copy_tile(int& t, int& y) {
    const int nx = gridDim.x, ny = gridDim.y, n = nx * ny;
    if (!FM_LU_XCD_ORDER) {
        t = blockIdx.x; y = blockIdx.y;
        return;
    }
    // XCD c holds the ids id = c (mod 8): n / 8 of them, one more for the first n % 8 XCDs
    const int id = blockIdx.x + nx * blockIdx.y, c = id % 8, per = n / 8, rem = n % 8;
    const int w = c * per + (c < rem ? c : rem) + id / 8;
    t = w / ny; y = w % ny;
}

__global__ __launch_bounds__(256) void k_lu_pack(KZ z, LuArgs a, int ntj) {
    __shared__ double T[64][65];
    // with a.w_ready the state rows hold the state already (arrays 5-9 skipped)
    const int f = a.w_ready && blockIdx.z >= 5 ? blockIdx.z + 5 : blockIdx.z;
    int t, y;
    lu_copy_tile(t, y);
    lu_pack_rows(z, a, ntj, t, y * 64, f, T);
}

// The skewed copies of several zones: workgroup (w, y, c) packs the c-th array its zone needs
// (as in k_lu_pack: the state rows are skipped with a.w_ready, the normal rows with a.n_ready) of
// rows [64 y, 64 y + 64) of pencil work[w].
__global__ __launch_bounds__(256) void k_lu_pack_list(const LuZone* __restrict__ zones, const LuWork* __restrict__ work) {
    __shared__ double T[64][65];
    const LuWork w = work[blockIdx.x];
    const KZ& z = zones[w.zone].z;
    const LuArgs& a = zones[w.zone].a;
    const int ntj = (a.j1 - a.j0 + 8) / 8, R = a.i1 - a.i0 + 15, r0 = blockIdx.y * 64;
    const int c = blockIdx.z;
    if (r0 >= R || c >= lu_pack_count(z.typ == 2, a)) return;
    lu_pack_rows(z, a, ntj, w.bj + ntj * w.bk, r0, a.w_ready && c >= 5 ? c + 5 : c, T);
}

// The results back to the zone arrays: workgroup (t, y, f) copies array f (dW 0-4, the updated
// state 5-9) of rows [64 y, 64 y + 64) of pencil t, through LDS so that the row reads (along the
// lanes) and the zone writes (along i) are contiguous.
__device__ __forceinline__ void lu_unpack_rows(const KZ& z, const LuArgs& a, int ntj, int t, int r0, int f,
                                               double (*T)[65]) {
    const int jt = t % ntj, kt = t / ntj;
    const int ni = a.i1 - a.i0 + 1, nj = a.j1 - a.j0 + 1, nk = a.k1 - a.k0 + 1, R = ni + 14;
    const long nd = z.nd;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nfld = lu_skf_of(z);
    const double* __restrict__ src = a.skw + f * 64 + lane;
    double v[16];
#pragma unroll
    for (int m = 0; m < 16; m++) {
        const int r = r0 + w + 4 * m;
        v[m] = r < R ? src[lu_sk_row(t, r, R, nfld)] : 0.;
    }
#pragma unroll
    for (int m = 0; m < 16; m++) T[lane][w + 4 * m] = v[m];
    __syncthreads();
    double* __restrict__ dst = f < 5 ? a.dw + f * nd : a.out + (f - 5) * nd;
#pragma unroll
    for (int m = 0; m < 16; m++) {
        const int L = w + 4 * m, oj = jt * 8 + (L & 7), ok = kt * 8 + (L >> 3);
        const int i = r0 + lane - (L & 7) - (L >> 3);
        if (oj < nj && ok < nk && i >= 0 && i < ni && r0 + lane < R)
            dst[IDX(z, a.i0 + i, a.j0 + oj, a.k0 + ok)] = T[L][lane];
    }
}


__global__ __launch_bounds__(256) void k_lu_unpack(KZ z, LuArgs a, int ntj, int f0) {
    __shared__ double T[64][65];
    int t, y;
    lu_copy_tile(t, y);
    lu_unpack_rows(z, a, ntj, t, y * 64, f0 + blockIdx.z, T);
}

// The results of several zones back: workgroup (w, y, c) copies array f0 + c (dW 0-4, the updated
// state 5-9) of rows [64 y, 64 y + 64) of pencil work[w]; as in k_lu_unpack, dW stays in the rows
// when a.keep_dw is off.
__global__ __launch_bounds__(256) void k_lu_unpack_list(const LuZone* __restrict__ zones, const LuWork* __restrict__ work) {
    __shared__ double T[64][65];
    const LuWork w = work[blockIdx.x];
    const KZ& z = zones[w.zone].z;
    const LuArgs& a = zones[w.zone].a;
    const int ntj = (a.j1 - a.j0 + 8) / 8, R = a.i1 - a.i0 + 15, r0 = blockIdx.y * 64;
    const int f = lu_unpack_first(a) + (int)blockIdx.z;
    if (r0 >= R || f >= lu_unpack_end(a)) return;
    lu_unpack_rows(z, a, ntj, w.bj + ntj * w.bk, r0, f, T);
}

// The pencil exchanges of several zones filled with the pending pattern (and the shared pencil
// counter reset): workgroup (n, c) fills chunk c of zone n's exchange.
__global__ __launch_bounds__(256) void k_lu_fill_list(const LuZone* __restrict__ zones, int* counter) {
    const LuArgs& a = zones[blockIdx.x].a;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *counter = 0;
    const long n = lu_stream_exchange_size(a.i1 - a.i0 + 1, a.j1 - a.j0 + 1, a.k1 - a.k0 + 1);
    for (long e = (long)blockIdx.y * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.y * blockDim.x)
        a.stx[e] = __longlong_as_double(kLuPending);
}

// One pencil (jt, kt) of a sweep, in sweep-order tile coordinates.
template <bool CART, bool LOWER>
__device__ __forceinline__ void lu_pencil(const KZ& z, const LuArgs& a, int jt, int kt, int ntj, int ntk) {
    constexpr int NF = lu_skf<CART>();
    const int L = threadIdx.x, jj = L & 7, kk = L >> 3;
    const int ni = a.i1 - a.i0 + 1, nj = a.j1 - a.j0 + 1, nk = a.k1 - a.k0 + 1;
    const int R = ni + 14;
    const int ntile = ntj * ntk;
    const LUC<> c = lu_const<double>(z);
    double* __restrict__ X = a.stx;
    double* __restrict__ SK = a.skw;
    int* err = a.sync + 1;
    const bool kalias = !LOWER && a.rop_k != a.rop;
    const bool newton = !LOWER && a.newton;
    const double tci = z.ti[0], tcj = z.tj[0], tck = z.tk[0];
    const double cvinv = 1. / z.cv, cvinv2 = 0.5 * cvinv;
    // first and last line of a pencil in sweep order
    const int first = LOWER ? 0 : 7, last = LOWER ? 7 : 0;
    {
        const int tileid = jt + ntj * kt;
        const int oj = jt * 8 + jj, ok = kt * 8 + kk;
        const bool lane_ok = oj < nj && ok < nk;
        // the lane's neighbours in j and k before it in sweep order, inside the LU box
        const bool hjn = LOWER ? oj > 0 : oj + 1 < nj;
        const bool hkn = LOWER ? ok > 0 : ok + 1 < nk;
        // the lanes of the first line in j (k) take the products of the previous pencil; those of
        // the last line hand theirs to the next pencil
        const bool get_j = lane_ok && jj == first && hjn;
        const bool get_k = lane_ok && kk == first && hkn;
        const bool put_j = lane_ok && jj == last && (LOWER ? oj + 1 < nj : oj > 0);
        const bool put_k = lane_ok && kk == last && (LOWER ? ok + 1 < nk : ok > 0);
        const int tj_next = put_j ? (LOWER ? jt + 1 : jt - 1) + ntj * kt : 0;
        const int tk_next = put_k ? jt + ntj * (LOWER ? kt + 1 : kt - 1) : 0;
        // value e of cell ii on the lane's face: hjp + (ii * 5 + e) * 8 (hkp)
        double* hjp = X + lu_st_slot(0, tileid, kk, 0, 0, ntile, ni);
        double* hkp = X + lu_st_slot(1, tileid, jj, 0, 0, ntile, ni);
        // the lane's cell at step s
        auto cell = [&](int s) { return (LOWER ? s : R - 1 - s) - jj - kk; };

        // the rows of steps s .. s + kLuStAhead - 1 in flight (slot s % kLuStAhead); the lower sweep
        // reads neither the density nor coe5
        double in[kLuStAhead][NF];
        auto rows = [&](int s, int q) {
            const int sc = s < R ? s : R - 1;
            const double* __restrict__ p = SK + lu_sk_row(tileid, LOWER ? sc : R - 1 - sc, R, NF) + L;
#pragma unroll
            for (int f = 0; f < NF; f++)
                if (!(LOWER && (f == 5 || f == 15))) in[q][f] = p[f * 64];
        };
        // halo products of step s, kLuStHalo steps ahead (buffer s % kLuStHalo), loaded by the
        // lanes that take them (the exchange is read by those lanes only: no other lane touches it)
        double hj[kLuStHalo][5], hk[kLuStHalo][5];
        auto hload = [&](int s, int bsel) {
            const int ii = cell(s);
            const bool v = s < R && ii >= 0 && ii < ni;
            if (get_j && v) {
#pragma unroll
                for (int e = 0; e < 5; e++) hj[bsel][e] = lu_ldc(hjp + (ii * 5 + e) * 8);
            }
            if (get_k && v) {
#pragma unroll
                for (int e = 0; e < 5; e++) hk[bsel][e] = lu_ldc(hkp + (ii * 5 + e) * 8);
            }
        };
        // the products of the previous step: i (own next cell), and the j and k products of the
        // lanes before, already moved
        double pi[5], bj[5], bk[5];
#pragma unroll
        for (int e = 0; e < 5; e++) { pi[e] = 0.; bj[e] = 0.; bk[e] = 0.; }

#pragma unroll
        for (int q = 0; q < kLuStAhead; q++) rows(q, q);
#pragma unroll
        for (int q = 0; q < kLuStHalo; q++) hload(q, q);
        // the B matrices of the three products of a step, formed before it waits on its
        // neighbours (they do not depend on the solution; with in-place sub-iterations the k
        // matrix takes the updated state and is formed with the product)
        LuB<double> B0, B1, B2;
        auto coefs = [&](int qn) {
            const double u = in[qn][6], v = in[qn][7], w = in[qn][8], t = in[qn][9];
            const double ci = LOWER ? in[qn][11] : -in[qn][11], cj = LOWER ? in[qn][12] : -in[qn][12];
            const double ck = LOWER ? in[qn][13] : -in[qn][13];
            if (CART) {
                B0 = lu_coef<true, 0, double>(c, u, v, w, t, ci, tci, 0., 0.);
                B1 = lu_coef<true, 1, double>(c, u, v, w, t, cj, tcj, 0., 0.);
                if (!(newton && kalias)) B2 = lu_coef<true, 2, double>(c, u, v, w, t, ck, tck, 0., 0.);
            }
            // (3dfull: the full 5 x 5 matrices are formed at their step, registers being short)
        };
#if FM_LU_ST_COEF_AHEAD
        coefs(0);
#endif
        for (int s0 = 0; s0 < R; s0 += kLuStAhead) {
#pragma unroll
            for (int q = 0; q < kLuStAhead; q++) {
                const int s = s0 + q;
                const int ii = cell(s);
                const bool act = s < R && lane_ok && ii >= 0 && ii < ni;
                const int hb = q % kLuStHalo;
#if !FM_LU_ST_COEF_AHEAD
                coefs(q);
#endif
                // the inputs of the step (lanes without a cell read zeros or stale values, never
                // used)
                double x[5], wv[5];
#pragma unroll
                for (int e = 0; e < 5; e++) {
                    x[e] = in[q][e];
                    wv[e] = LOWER && e == 0 ? 0. : in[q][5 + e];
                }
                // the sweep's divisor, sign products (coe * 0.5 negated for the upper sweep equals
                // coe * -0.5) and scaled right-hand side
                const double xal = LOWER ? in[q][10] : -in[q][10], diag = in[q][14];
                const double ck = LOWER ? in[q][13] : -in[q][13];
                if (!LOWER) {
#pragma unroll
                    for (int e = 0; e < 5; e++) x[e] = x[e] * in[q][15];
                }
                // halo products of the previous pencils: while any lane still holds the pending
                // pattern, the wave re-reads its values (bounded)
                const bool uj = get_j && act, uk = get_k && act;
                {
                    auto pend = [&]() {
                        bool p = false;
#pragma unroll
                        for (int e = 0; e < 5; e++)
                            p = p | (uj & (__double_as_longlong(hj[hb][e]) == kLuPending)) |
                                (uk & (__double_as_longlong(hk[hb][e]) == kLuPending));
                        return p;
                    };
                    if (__builtin_amdgcn_ballot_w64(pend())) {
                        int spins = 0;
                        // the previous pencil is about to pass this step: wait until it has also
                        // handed over the cell kLuStSlack steps ahead (it writes its cells in
                        // order), so that the next steps find their values in place instead of
                        // waiting one round trip each
                        {
                            const int iia = LOWER ? min(ii + kLuStSlack, ni - 1) : max(ii - kLuStSlack, 0);
                            const double* pa = (uj ? hjp : hkp) + (iia * 5 + 4) * 8;
                            const bool wa = uj || uk;
                            double va = wa ? lu_ldc(pa) : 0.;
                            while (__builtin_amdgcn_ballot_w64(wa && __double_as_longlong(va) == kLuPending)) {
                                if (++spins > kLuSpinMax ||
                                    ((spins & 63) == 1 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
                                    break;
                                __builtin_amdgcn_s_sleep(FM_LU_ST_SLEEP_SLACK);
                                if (wa) va = lu_ldc(pa);
                            }
#pragma unroll
                            for (int d = 1; d < kLuStHalo; d++) hload(s + d, (hb + d) % kLuStHalo);
                        }
                        do {
                            if (++spins > kLuSpinMax ||
                                ((spins & 63) == 1 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                                if (L == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                break;
                            }
                            __builtin_amdgcn_s_sleep(FM_LU_ST_SLEEP_RETRY);
                            hload(s, hb);
                            // every access of the wave done: the steps after this rare path keep
                            // exact memory counters
                            __builtin_amdgcn_s_waitcnt(0);
                        } while (__builtin_amdgcn_ballot_w64(pend()));
                    }
                }
#pragma unroll
                for (int e = 0; e < 5; e++) {
                    bj[e] = jj == first ? hj[hb][e] : bj[e];
                    bk[e] = kk == first ? hk[hb][e] : bk[e];
                }
                const bool hi = LOWER ? ii > 0 : ii + 1 < ni;
                // neighbour terms in the reference order: lower 3dcart J, K, I; lower 3dfull and
                // upper K, J, I
#pragma unroll
                for (int qq = 0; qq < 3; qq++) {
                    const int D = (LOWER && CART) ? (qq == 0 ? 1 : qq == 1 ? 2 : 0) : (qq == 0 ? 2 : qq == 1 ? 1 : 0);
                    const bool has = D == 0 ? hi : D == 1 ? hjn : hkn;
#pragma unroll
                    for (int e = 0; e < 5; e++) {
                        const double bb = D == 0 ? pi[e] : D == 1 ? bj[e] : bk[e];
                        const double xn = x[e] + bb * xal;
                        x[e] = has ? xn : x[e];
                    }
                }
#pragma unroll
                for (int e = 0; e < 5; e++) x[e] = x[e] * diag;
                const double uu = wv[1], vv = wv[2], ww = wv[3], tt = wv[4];
                // the i and j products (the j product moves one lane by DPP)
                double pk[5], pj[5];
                if (CART) {
                    lu_apply<true, 0, double>(B0, x, pi);
                    lu_apply<true, 1, double>(B1, x, pj);
                } else {
                    const double ci = LOWER ? in[q][11] : -in[q][11], cj = LOWER ? in[q][12] : -in[q][12];
                    lu_prod<false, 0, double>(c, uu, vv, ww, tt, ci, in[q][16], in[q][17], in[q][18], x, pi);
                    lu_prod<false, 1, double>(c, uu, vv, ww, tt, cj, in[q][19], in[q][20], in[q][21], x, pj);
                }
#pragma unroll
                for (int e = 0; e < 5; e++) bj[e] = lu_from_jprev<LOWER>(pj[e]);
                double o[5] = {0., 0., 0., 0., 0.};
                if (newton) {
                    // mjr_newton.for: the Newton-updated state; in place (kalias) the k products
                    // see it
                    const double r0 = wv[0];
                    const double r1 = r0 + x[0];
                    const double r_1 = 1. / r1;
                    const double u1 = (r0 * uu + x[1]) * r_1, v1 = (r0 * vv + x[2]) * r_1, w1 = (r0 * ww + x[3]) * r_1;
                    const double roe_old = r0 * (z.cv * tt + 0.5 * (uu * uu + vv * vv + ww * ww));
                    const double t1 = (roe_old + x[4]) * r_1 * cvinv - cvinv2 * (u1 * u1 + v1 * v1 + w1 * w1);
                    o[0] = r1; o[1] = u1; o[2] = v1; o[3] = w1; o[4] = t1;
                }
                // the k product (moves eight lanes by ds_bpermute)
                if (newton && kalias) {
                    if (CART) lu_prod<true, 2, double>(c, o[1], o[2], o[3], o[4], ck, tck, 0., 0., x, pk);
                    else lu_prod<false, 2, double>(c, o[1], o[2], o[3], o[4], ck, in[q][22], in[q][23], in[q][24], x, pk);
                } else if (CART) {
                    lu_apply<true, 2, double>(B2, x, pk);
                } else {
                    lu_prod<false, 2, double>(c, uu, vv, ww, tt, ck, in[q][22], in[q][23], in[q][24], x, pk);
                }
#pragma unroll
                for (int e = 0; e < 5; e++) bk[e] = lu_from_kprev<LOWER>(pk[e], L);
                if (act) {
                    // lower: dW* into the right-hand-side row the upper sweep reads; upper: dW and
                    // the updated state into the row just read
                    double* __restrict__ p = SK + lu_sk_row(tileid, LOWER ? s : R - 1 - s, R, NF) + L;
#pragma unroll
                    for (int e = 0; e < 5; e++) p[e * 64] = x[e];
                    if (newton) {
#pragma unroll
                        for (int e = 0; e < 5; e++) p[(5 + e) * 64] = o[e];
                    }
                }
                // the face products for the next pencils
                if (put_j && act) {
#pragma unroll
                    for (int e = 0; e < 5; e++) lu_stc(X + lu_st_slot(0, tj_next, kk, e, ii, ntile, ni), pj[e]);
                }
                if (put_k && act) {
#pragma unroll
                    for (int e = 0; e < 5; e++) lu_stc(X + lu_st_slot(1, tk_next, jj, e, ii, ntile, ni), pk[e]);
                }
                hload(s + kLuStHalo, hb);
                // the row kLuStAhead steps ahead into the slot just read
                rows(s + kLuStAhead, q);
#if FM_LU_ST_COEF_AHEAD
                coefs((q + 1) % kLuStAhead);
#endif
            }
        }
    }
}


template <bool CART, bool LOWER>
__global__ __launch_bounds__(64) void k_lu_stream(KZ z, LuArgs a, int ntj, int ntk) {
    const int ntile = ntj * ntk;
    for (;;) {
        int tk0 = 0;
        if (threadIdx.x == 0) tk0 = atomicAdd(a.sync, 1);
        const int t = __builtin_amdgcn_readfirstlane(tk0);
        if (t >= ntile) break;
        int jt, kt;
        lu_tile_of(t, ntj, ntk, jt, kt);
        if (!LOWER) { jt = ntj - 1 - jt; kt = ntk - 1 - kt; }
        lu_pencil<CART, LOWER>(z, a, jt, kt, ntj, ntk);
    }
}

// Streaming sweeps of several zones in one launch: the pencils of every zone in one list ordered by
// pencil diagonal (jt + kt), handed out from one counter; the upper sweep takes each zone's
// pencils mirrored, so in either sweep a pencil only waits on pencils of its zone handed out before
// it.  Each zone has its own rows, exchange and error word.
template <bool CART, bool LOWER>
__global__ __launch_bounds__(64) void k_lu_stream_list(const LuZone* __restrict__ zones,
                                                       const LuWork* __restrict__ work, int nwork, int* counter) {
    for (;;) {
        int tk0 = 0;
        if (threadIdx.x == 0) tk0 = atomicAdd(counter, 1);
        const int t = __builtin_amdgcn_readfirstlane(tk0);
        if (t >= nwork) break;
        const LuWork w = work[t];
        const KZ& z = zones[w.zone].z;
        const LuArgs& a = zones[w.zone].a;
        const int ntj = (a.j1 - a.j0 + 8) / 8, ntk = (a.k1 - a.k0 + 8) / 8;
        int jt = w.bj, kt = w.bk;
        if (!LOWER) { jt = ntj - 1 - jt; kt = ntk - 1 - kt; }
        lu_pencil<CART, LOWER>(z, a, jt, kt, ntj, ntk);
    }
}

// fill n doubles with a 64-bit pattern and zero the tile counter
__global__ void k_fill_lin(double* __restrict__ p, long n, long long bits, int* sync) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (sync && t == 0) sync[0] = 0;
    if (t < n) p[t] = __longlong_as_double(bits);
}


// Copy the Newton-updated LU box back (in-place sub-iterations).
__global__ void k_copy_box(KZ z, const double* __restrict__ src, double* __restrict__ dst, Box b) {
    int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int k = b.k0 + blockIdx.y;
    if (t >= ni * nj) return;
    int i = b.i0 + t % ni, j = b.j0 + t / ni;
    long nd = z.nd, l = IDX(z, i, j, k);
#pragma unroll
    for (int e = 0; e < 5; e++) dst[l + e * nd] = src[l + e * nd];
}

namespace {
template <bool CART, bool LOWER>
void sweep(const KZ& z, const LuArgs& a, hipStream_t st) {
    constexpr int CI = lu_ci<CART>();
    int ni = a.i1 - a.i0 + 1, nj = a.j1 - a.j0 + 1, nk = a.k1 - a.k0 + 1;
    int nbi = (ni + CI - 1) / CI, nbj = (nj + 7) / 8, nbk = (nk + 7) / 8;
    int ndiag = nbi + nbj + nbk - 2;
    for (int d = 0; d < ndiag; d++) {
        int lo = d - (nbi - 1) - (nbj - 1);
        if (lo < 0) lo = 0;
        int hi = d < nbk - 1 ? d : nbk - 1;
        int jhi = d < nbj - 1 ? d : nbj - 1;
        dim3 g(jhi + 1, hi - lo + 1);
        if (a.fp32) hipLaunchKernelGGL((k_lu_block<CART, LOWER, CI, float>), g, dim3(256), 0, st, z, a, d, lo, nbi, nbj);
        else hipLaunchKernelGGL((k_lu_block<CART, LOWER, CI>), g, dim3(256), 0, st, z, a, d, lo, nbi, nbj);
    }
}
}  // namespace

namespace {

template <bool CART>
void flow(const KZ& z, const LuArgs& a, hipStream_t st) {
    constexpr int CI = lu_ci<CART>();
    int ni = a.i1 - a.i0 + 1, nj = a.j1 - a.j0 + 1, nk = a.k1 - a.k0 + 1;
    int nbi = (ni + CI - 1) / CI, nbj = (nj + 7) / 8, nbk = (nk + 7) / 8;
    long nx = lu_flow_exchange_size(CART, ni, nj, nk);
    // as many workgroups as can be resident at once (measured: 512 on a Cartesian zone beat 1024
    // and 2048, whose extra workgroups only queue behind the resident ones)
    int cus = 0, dev = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    const void* kfn = a.fp32 ? (const void*)k_lu_flow<CART, true, CI, float> : (const void*)k_lu_flow<CART, true, CI>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kfn, 256, 0) != hipSuccess || per < 1) per = 1;
    long cap = (long)per * cus;
    if (cap > FM_LU_FLOW_GRID) cap = FM_LU_FLOW_GRID;
    if (a.flow_grid > 0 && a.flow_grid < cap) cap = a.flow_grid;
    unsigned grid = (unsigned)(a.flow_n < cap ? a.flow_n : cap);
    // the exchange is pending before a sweep (filled when allocated, and every slot a sweep writes
    // is read and given back by its consumer): only the block counter is reset
    (void)nx;
    hipLaunchKernelGGL(k_fill_lin, dim3(1), dim3(64), 0, st, a.fx, 0L, kLuPending, a.sync);
    if (a.fp32) hipLaunchKernelGGL((k_lu_flow<CART, true, CI, float>), dim3(grid), dim3(256), 0, st, z, a, nbi, nbj, nbk);
    else if (a.flow_step) hipLaunchKernelGGL((k_lu_flow<CART, true, CI, double, true>), dim3(grid), dim3(256), 0, st, z, a, nbi, nbj, nbk);
    else hipLaunchKernelGGL((k_lu_flow<CART, true, CI>), dim3(grid), dim3(256), 0, st, z, a, nbi, nbj, nbk);
    hipLaunchKernelGGL(k_fill_lin, dim3(1), dim3(64), 0, st, a.fx, 0L, kLuPending, a.sync);
    if (a.fp32) hipLaunchKernelGGL((k_lu_flow<CART, false, CI, float>), dim3(grid), dim3(256), 0, st, z, a, nbi, nbj, nbk);
    else if (a.flow_step) hipLaunchKernelGGL((k_lu_flow<CART, false, CI, double, true>), dim3(grid), dim3(256), 0, st, z, a, nbi, nbj, nbk);
    else hipLaunchKernelGGL((k_lu_flow<CART, false, CI>), dim3(grid), dim3(256), 0, st, z, a, nbi, nbj, nbk);
}

void fill_exchange(const LuArgs& a, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_lin, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a.xch, n, kLuPending, a.sync);
}

// Streaming sweeps of a zone: the skewed copy of the inputs, then for each sweep the exchange filled
// with the pending pattern (and the tile counter reset) and one launch with as many workgroups as
// can be resident at once (at most one per pencil).
template <bool CART>
void stream(const KZ& z, const LuArgs& a, hipStream_t st) {
    int ni = a.i1 - a.i0 + 1, nj = a.j1 - a.j0 + 1, nk = a.k1 - a.k0 + 1;
    int ntj = (nj + 7) / 8, ntk = (nk + 7) / 8;
    const long nx = lu_stream_exchange_size(ni, nj, nk);
    hipLaunchKernelGGL(k_lu_pack, dim3((unsigned)(ntj * ntk), (unsigned)((ni + 14 + 63) / 64), (unsigned)lu_pack_count(CART, a)),
                       dim3(256), 0, st, z, a, ntj);
    int cus = 0, dev = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_lu_stream<CART, true>, 64, 0) != hipSuccess || per < 1)
        per = 1;
    long cap = (long)per * cus;
    unsigned grid = (unsigned)((long)ntj * ntk < cap ? (long)ntj * ntk : cap);
    hipLaunchKernelGGL(k_fill_lin, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, st, a.stx, nx, kLuPending, a.sync);
    hipLaunchKernelGGL((k_lu_stream<CART, true>), dim3(grid), dim3(64), 0, st, z, a, ntj, ntk);
    hipLaunchKernelGGL(k_fill_lin, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, st, a.stx, nx, kLuPending, a.sync);
    hipLaunchKernelGGL((k_lu_stream<CART, false>), dim3(grid), dim3(64), 0, st, z, a, ntj, ntk);
    // dW back only if asked (inside fm_compute it is copied once, after the last time step)
    const int f0 = lu_unpack_first(a), f1 = lu_unpack_end(a);
    hipLaunchKernelGGL(k_lu_unpack, dim3((unsigned)(ntj * ntk), (unsigned)((ni + 14 + 63) / 64), (unsigned)(f1 - f0)),
                       dim3(256), 0, st, z, a, ntj, f0);
}

template <bool CART>
void pipeline(const KZ& z, const LuArgs& a, hipStream_t st) {
    int ni = a.i1 - a.i0 + 1, nj = a.j1 - a.j0 + 1, nk = a.k1 - a.k0 + 1;
    int ntj = (nj + 7) / 8, ntk = (nk + 7) / 8;
    long nx = lu_exchange_size(ni, nj, nk);
    // lower: RHS (drodm) -> dW* (dws)
    fill_exchange(a, nx, st);
    hipLaunchKernelGGL((k_lu_pipe<CART, true, CART ? 4 : 2>), dim3(ntj * ntk), dim3(64), 0, st, z, a, a.dw, a.dws, ntj, ntk);
    // upper: dW* -> dW (drodm), Newton update into out
    fill_exchange(a, nx, st);
    hipLaunchKernelGGL((k_lu_pipe<CART, false, CART ? 4 : 2>), dim3(ntj * ntk), dim3(64), 0, st, z, a, a.dws, a.dw, ntj, ntk);
}
}  // namespace

void fill_pending(double* x, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_lin, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, n, kLuPending, (int*)nullptr);
}


// Block lists of a set of zones of one mesh kind for the diagonal-batched sweeps: for every zone
// (LU box of boxes[n]) its CI x 8 x 8 blocks in sweep coordinates, grouped by diagonal;
// offs[d]..offs[d+1] are the blocks of launch d.
void lu_work_list(bool cart, const std::vector<Box>& boxes, std::vector<LuWork>& work, std::vector<int>& offs) {
    const int CI = cart ? lu_ci<true>() : lu_ci<false>();
    int nd = 0;
    std::vector<std::array<int, 3>> nb(boxes.size());
    for (size_t n = 0; n < boxes.size(); n++) {
        const Box& b = boxes[n];
        nb[n] = {(b.i1 - b.i0 + 1 + CI - 1) / CI, (b.j1 - b.j0 + 1 + 7) / 8, (b.k1 - b.k0 + 1 + 7) / 8};
        if (b.i1 < b.i0 || b.j1 < b.j0 || b.k1 < b.k0) nb[n] = {0, 0, 0};
        int d = nb[n][0] + nb[n][1] + nb[n][2] - 2;
        if (nb[n][0] > 0 && d > nd) nd = d;
    }
    work.clear(); offs.assign(1, 0);
    for (int d = 0; d < nd; d++) {
        for (size_t n = 0; n < boxes.size(); n++) {
            if (nb[n][0] == 0) continue;
            for (int bk = 0; bk < nb[n][2]; bk++)
                for (int bj = 0; bj < nb[n][1]; bj++) {
                    int bi = d - bj - bk;
                    if (bi >= 0 && bi < nb[n][0]) work.push_back(LuWork{(int)n, bi, bj, bk});
                }
        }
        offs.push_back((int)work.size());
    }
}

void launch_lu_list(bool cart, bool lower, bool fp32, const LuZone* zones, const LuWork* work,
                    const std::vector<int>& offs, hipStream_t st) {
    constexpr int CC = lu_ci<true>(), CF = lu_ci<false>();
    for (size_t d = 0; d + 1 < offs.size(); d++) {
        int n = offs[d + 1] - offs[d];
        if (n <= 0) continue;
        const LuWork* w = work + offs[d];
        if (fp32) {
            if (cart) {
                if (lower) hipLaunchKernelGGL((k_lu_block_list<true, true, CC, float>), dim3(n), dim3(256), 0, st, zones, w);
                else hipLaunchKernelGGL((k_lu_block_list<true, false, CC, float>), dim3(n), dim3(256), 0, st, zones, w);
            } else {
                if (lower) hipLaunchKernelGGL((k_lu_block_list<false, true, CF, float>), dim3(n), dim3(256), 0, st, zones, w);
                else hipLaunchKernelGGL((k_lu_block_list<false, false, CF, float>), dim3(n), dim3(256), 0, st, zones, w);
            }
        } else if (cart) {
            if (lower) hipLaunchKernelGGL((k_lu_block_list<true, true, CC>), dim3(n), dim3(256), 0, st, zones, w);
            else hipLaunchKernelGGL((k_lu_block_list<true, false, CC>), dim3(n), dim3(256), 0, st, zones, w);
        } else {
            if (lower) hipLaunchKernelGGL((k_lu_block_list<false, true, CF>), dim3(n), dim3(256), 0, st, zones, w);
            else hipLaunchKernelGGL((k_lu_block_list<false, false, CF>), dim3(n), dim3(256), 0, st, zones, w);
        }
    }
}

// pencils (zone, 0, jt, kt) of several LU boxes, in pencil-diagonal order
void lu_pencil_list(const std::vector<Box>& boxes, std::vector<LuWork>& work, int& maxrb) {
    int nd = 0;
    maxrb = 0;
    std::vector<std::array<int, 2>> nt(boxes.size());
    for (size_t n = 0; n < boxes.size(); n++) {
        const Box& b = boxes[n];
        const bool empty = b.i1 < b.i0 || b.j1 < b.j0 || b.k1 < b.k0;
        nt[n] = {empty ? 0 : (b.j1 - b.j0 + 1 + 7) / 8, empty ? 0 : (b.k1 - b.k0 + 1 + 7) / 8};
        if (!empty) {
            nd = std::max(nd, nt[n][0] + nt[n][1] - 1);
            maxrb = std::max(maxrb, (b.i1 - b.i0 + 1 + 14 + 63) / 64);
        }
    }
    work.clear();
    for (int d = 0; d < nd; d++)
        for (size_t n = 0; n < boxes.size(); n++)
            for (int kt = 0; kt < nt[n][1]; kt++) {
                const int jt = d - kt;
                if (jt >= 0 && jt < nt[n][0]) work.push_back(LuWork{(int)n, 0, jt, kt});
            }
}

// Streaming LU-SGS of several zones of one mesh kind: the skewed copies, then for each sweep the
// exchanges filled (and the pencil counter reset) and one launch over the pencil list, then the
// results back.
void launch_lu_stream_list(bool cart, const LuZone* zones, int nz, const LuWork* work, int nwork, int maxrb,
                           int npack, int nunpack, int* counter, hipStream_t st) {
    if (nwork <= 0 || nz <= 0) return;
    int cus = 0, dev = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    const void* kfn = cart ? (const void*)k_lu_stream_list<true, true> : (const void*)k_lu_stream_list<false, true>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kfn, 64, 0) != hipSuccess || per < 1) per = 1;
    const long cap = (long)per * cus;
    const unsigned grid = (unsigned)(nwork < cap ? nwork : cap);
    if (npack > 0)
        hipLaunchKernelGGL(k_lu_pack_list, dim3((unsigned)nwork, (unsigned)maxrb, (unsigned)npack), dim3(256), 0, st,
                           zones, work);
    for (int lower = 1; lower >= 0; lower--) {
        hipLaunchKernelGGL(k_lu_fill_list, dim3((unsigned)nz, 64), dim3(256), 0, st, zones, counter);
        if (cart) {
            if (lower) hipLaunchKernelGGL((k_lu_stream_list<true, true>), dim3(grid), dim3(64), 0, st, zones, work, nwork, counter);
            else hipLaunchKernelGGL((k_lu_stream_list<true, false>), dim3(grid), dim3(64), 0, st, zones, work, nwork, counter);
        } else {
            if (lower) hipLaunchKernelGGL((k_lu_stream_list<false, true>), dim3(grid), dim3(64), 0, st, zones, work, nwork, counter);
            else hipLaunchKernelGGL((k_lu_stream_list<false, false>), dim3(grid), dim3(64), 0, st, zones, work, nwork, counter);
        }
    }
    if (nunpack > 0)
        hipLaunchKernelGGL(k_lu_unpack_list, dim3((unsigned)nwork, (unsigned)maxrb, (unsigned)nunpack), dim3(256), 0, st,
                           zones, work);
}

void launch_lu_flow_list(bool cart, const LuZone* zones, const LuWork* work, int nwork, int* counter,
                         hipStream_t st) {
    if (nwork <= 0) return;
    constexpr int CC = lu_ci<true>(), CF = lu_ci<false>();
    int cus = 0, dev = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    const void* kfn = cart ? (const void*)k_lu_flow_list<true, true, CC> : (const void*)k_lu_flow_list<false, true, CF>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kfn, 256, 0) != hipSuccess || per < 1) per = 1;
    long cap = (long)per * cus;
    const unsigned grid = (unsigned)(nwork < cap ? nwork : cap);
    for (int lower = 1; lower >= 0; lower--) {
        hipLaunchKernelGGL(k_fill_lin, dim3(1), dim3(64), 0, st, (double*)nullptr, 0L, kLuPending, counter);
        if (cart) {
            if (lower) hipLaunchKernelGGL((k_lu_flow_list<true, true, CC>), dim3(grid), dim3(256), 0, st, zones, work, nwork, counter);
            else hipLaunchKernelGGL((k_lu_flow_list<true, false, CC>), dim3(grid), dim3(256), 0, st, zones, work, nwork, counter);
        } else {
            if (lower) hipLaunchKernelGGL((k_lu_flow_list<false, true, CF>), dim3(grid), dim3(256), 0, st, zones, work, nwork, counter);
            else hipLaunchKernelGGL((k_lu_flow_list<false, false, CF>), dim3(grid), dim3(256), 0, st, zones, work, nwork, counter);
        }
    }
}

long lu_flow_exchange_size(bool cart, int ni, int nj, int nk) {
    const int CI = cart ? lu_ci<true>() : lu_ci<false>();
    long nb = (long)((ni + CI - 1) / CI) * ((nj + 7) / 8) * ((nk + 7) / 8);
    return nb * 5 * (64 + 16 * CI);
}

bool lu_streams(const KZ& z, const LuArgs& a) {
    (void)z;
    return a.schedule == LU_STREAM && !a.fp32 && a.skw;
}

// dW of the last streaming LU-SGS, left in its skewed rows, back to the zone array
void launch_lu_unpack_dw(const KZ& z, const LuArgs& a, hipStream_t st) {
    if (!lu_streams(z, a) || a.i1 < a.i0 || a.j1 < a.j0 || a.k1 < a.k0) return;
    int ni = a.i1 - a.i0 + 1, nj = a.j1 - a.j0 + 1, nk = a.k1 - a.k0 + 1;
    int ntj = (nj + 7) / 8, ntk = (nk + 7) / 8;
    hipLaunchKernelGGL(k_lu_unpack, dim3((unsigned)(ntj * ntk), (unsigned)((ni + 14 + 63) / 64), 5u),
                       dim3(256), 0, st, z, a, ntj, 0);
}

void launch_lu(const KZ& z, const LuArgs& a, hipStream_t st) {
    if (a.i1 < a.i0 || a.j1 < a.j0 || a.k1 < a.k0) return;
    if (lu_streams(z, a)) {
        if (z.typ == 2) stream<true>(z, a, st); else stream<false>(z, a, st);
        return;
    }
    if (a.schedule == LU_DATAFLOW || a.schedule == LU_STREAM) {
        if (z.typ == 2) flow<true>(z, a, st); else flow<false>(z, a, st);
        return;
    }
    if (a.schedule == LU_PIPELINE) {
        if (z.typ == 2) pipeline<true>(z, a, st); else pipeline<false>(z, a, st);
        return;
    }
    if (z.typ == 2) { sweep<true, true>(z, a, st); sweep<true, false>(z, a, st); }
    else { sweep<false, true>(z, a, st); sweep<false, false>(z, a, st); }
}

// error word of the pipelined sweeps (a bounded wait gave up); clears it
int lu_pipeline_failed(const LuArgs& a, hipStream_t st) {
    if ((a.schedule != LU_PIPELINE && a.schedule != LU_DATAFLOW && a.schedule != LU_STREAM) || !a.sync) return 0;
    int h = 0;
    if (hipMemcpyAsync(&h, a.sync + 1, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
    if (hipStreamSynchronize(st) != hipSuccess) return -1;
    if (h) {
        int zero = 0;
        (void)hipMemcpy(a.sync + 1, &zero, sizeof(int), hipMemcpyHostToDevice);
    }
    return h;
}

void launch_copy_box(const KZ& z, const double* src, double* dst, Box b, hipStream_t st) {
    int n0 = b.i1 - b.i0 + 1, n1 = b.j1 - b.j0 + 1, n2 = b.k1 - b.k0 + 1;
    if (n0 <= 0 || n1 <= 0 || n2 <= 0) return;
    hipLaunchKernelGGL(k_copy_box, dim3((n0 * n1 + 255) / 256, n2), dim3(256), 0, st, z, src, dst, b);
}

}  // namespace fm
