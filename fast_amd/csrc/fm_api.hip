// C ABI of the fastmi library: contexts, zones, and the sub-iteration / time-step drivers.
// Host-side orchestration follows the reference drivers (Compute/computePT.cpp,
// Compute/gsdr3.cpp, Compute/lhs.cpp, BC/BCzone.cpp) for the implemented options; all
// numerics run in the HIP kernels of fm_kernels.hip.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <cstdarg>
#include <mutex>
#include <iterator>
#include <string>
#include <vector>
#include "fastmi.h"
#include "fm_kernels.hpp"

using namespace fm;

// param_solver.h slots
enum { NIJK = 0, NIJK_MTR = 5, NIJK_XYZ = 10, IJKV = 20, ITYPZONE = 25, IFLOW = 27, ILES = 28,
       ITYPCP = 29, KFLUDOM = 33, LALE = 34, NEQ = 36, NEQ_IJ = 37, NEQ_K = 38, NEQ_COE = 39,
       NEQ_VENT = 40,
       NDIMDX = 41, NDIMDX_XYZ = 42, NDIMDX_MTR = 43, DTLOC = 46, RK = 52, LEVEL = 53,
       EXPLOC = 54, NSSITER = 58, CACHEBLCKI = 59, CACHEBLCKJ = 60, CACHEBLCKK = 61, SLOPE = 64, PT_BC = 70, NB_RELAX = 71, IMPLICITSOLVER = 74,
       LU_MATCH = 75, SENSORTYPE = 85, BC_TYPE = 0, BC_IDIR = 1, BC_FEN = 2, BC_NBDATA = 8 };
enum { R_DTC = 0, R_GAMMA = 1, R_CVINF = 2, R_ROINF = 3, R_PINF = 4, R_VINF = 5, R_TINF = 6,
       R_VXINF = 19, R_VYINF = 20, R_VZINF = 21,
       R_PRANDT = 10, R_XMUL0 = 11, R_TEMP0 = 12, R_CS = 13, R_CFL = 15, R_PSIROE = 34,
       R_WIG_AMP = 72 };

static thread_local std::string g_err;
static int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap; va_start(ap, fmt); vsnprintf(buf, sizeof buf, fmt, ap); va_end(ap);
    g_err = buf;
    return code;
}
#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return fail(FM_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); } while (0)

struct BcRec {
    int type, idir, fen[6];
    int kind;              // kernel kind (BcArgs::kind) when applied to a state
    double state[5];       // conservative reference state (BCzone.cpp DEFAULT_STATE)
    double* pext;          // outpres: device copy of the per-face-point pressure
    int inc_bc[3];         // outpres: indbci strides (INCREMENT_BC.h)
};

struct Zone {
    std::vector<int> p;
    std::vector<double> pr;
    std::vector<BcRec> bcs;
    KZ k;
    double* ro[3];           // current N, N-1, N+1
    double* ro_scratch;      // Newton target when the LU sweep works in place
    double* lu_dws;          // dW* between the pipelined LU sweeps
    int* lu_sync;            // tile counter and error word of the pipelined LU sweeps
    double* lu_xch;          // tile-boundary exchange of the pipelined LU sweeps (on first use)
    double* adj_buf;         // face contributions of the transposed product (on first use)
    double* lu_fx;           // dataflow LU schedule: face exchange and block order (on first use)
    double* lu_stx;          // streaming LU schedule: pencil exchange (on first use)
    double* lu_skw;          // streaming LU schedule: skewed copy of the sweep inputs (on first use)
    int dw_pending;          // dW of the last streaming LU-SGS is still in its skewed rows (dw_la)
    LuArgs dw_la;
    // the state level whose LU-box values the state rows of the skewed copy hold (written there by
    // the last upper sweep; nullptr: none), and that box
    const double* w_rows;
    Box w_box;
    // 3dfull: the box whose mean face normals the normal rows hold (n_known)
    int n_known;
    Box n_box;
    LuWork* flow_work; int flow_n; Box flow_box;
    double* norm_part; double* norm_out; int norm_blocks;
    double h_norms[10]; int have_norms, have_cfl;
    // residuals of the sub-iterations of the last verified step, one slot of 10 doubles (sum of
    // squares, max |R| per equation) per sub-iteration; rec_lo..rec_hi are the recorded ones
    double* norm_hist; int rec_lo, rec_hi;
    long n_metric;
};

struct fm_ctx {
    int device;
    int lu_schedule;         // LU_PIPELINE or LU_BLOCKS
    int res_schedule;        // RES_CELLS, RES_FACES, RES_KCOL or RES_AUTO (per zone)
    hipStream_t stream;
    std::vector<Zone*> zones;
    hipEvent_t e0, e1;
    // injection transfers between zones (all raccords, flattened)
    TransferPoint* tr_pts;
    long tr_n;
    // per raccord: first point, point count, and whether its transpose must run point by point
    // (a donor or receiver cell repeated, or a receiver that is also a donor)
    std::vector<long> tr_off, tr_cnt;
    std::vector<char> tr_serial;
    // a receiver cell filled by more than one point: the transfers then run raccord by raccord,
    // in order, and a raccord that repeats a receiver runs its points in order (the last wins)
    bool tr_in_order = false;
    std::vector<char> tr_rdup;
    // diagonal-batched LU-SGS over several zones: cached block lists (per set of LU boxes)
    // and the per-zone argument table, staged through two pinned host buffers
    struct LuList { bool cart; bool stream; std::vector<Box> boxes; LuWork* work; std::vector<int> offs; int maxrb; };
    std::vector<LuList> lu_lists;
    LuZone* lu_tab; int lu_tab_cap;
    LuZone* lu_host[2]; hipEvent_t lu_ev[2]; int lu_slot;
    // the edge and corner table of several zones, outside graphs (staged as the LU table)
    CoinsZone* co_tab; int co_cap;
    CoinsZone* co_host[2]; hipEvent_t co_ev[2]; int co_slot;
    int* flow_ctr;           // block counter of the batched dataflow sweeps
    // several zones: each zone's right-hand side and boundary conditions run on a stream of its
    // own (forked from and joined back to `stream`); `cur` is the stream the zone work goes to
    hipStream_t cur;
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> side_ev;
    hipEvent_t fork_ev;
    int lu_batch;            // batch the LU of several zones (default on)
    int lu_fp32;             // mixed precision: LU-SGS sweep arithmetic in FP32 (default off)
    int res_kcol_n;          // tile shapes: k-column length of RES_KCOL (0: built-in)
    int lu_flow_grid;        //   workgroups of a dataflow sweep (0: all that are resident)
    // time steps captured as HIP graphs (FM_OPT_GRAPH): one per distinct configuration
    // inside fm_compute: the streaming LU-SGS leaves dW in its skewed rows (copied back once at
    // the end: within a time step nothing reads it before the next residual overwrites it)
    int defer_dw;
    struct GraphEntry {
        std::vector<const void*> ptrs;  // every zone's three level pointers
        std::vector<int> dw_pending;    // per zone, after the time step
        std::vector<const double*> w_rows;        // per zone, before (part of the key) and after
        std::vector<const double*> w_rows_after;
        int first_it, verif, n0, n1;
        long epoch;
        int state;                      // 1: ran once eagerly, 2: captured
        hipGraphExec_t exec;
        // device tables a captured step reads: LU (key 2 sub-iteration + mesh kind), edges and
        // corners (key -1 - sub-iteration)
        std::vector<std::pair<long, void*>> tabs;
    };
    int graph_on;
    long graph_epoch;                   // bumped by anything a captured kernel argument depends on
    std::vector<GraphEntry*> graphs;
};

extern "C" const char* fm_last_error(void) { return g_err.c_str(); }

// Device scratch of fm_wall_effort_noalloc: the per-face fields and contributions of one face
// window (33 x faces doubles) and their 12 sums, in a file-scope device array so that the entry
// point allocates nothing.  Each device holds its own copy until the process ends; host code
// reaches it only through the address hipGetSymbolAddress returns for that device
// (g_scratch_ptr), resolved once per device by fm_ctx_create.
constexpr long kScratchFaces = 65536;
constexpr long kScratchDoubles = 33 * kScratchFaces + 16;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mutex;

static int resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(FM_E_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return FM_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return fail(FM_E_HIP, "scratch address: %s", hipGetErrorString(e));
    g_scratch_ptr[device] = (double*)ptr;
    return FM_OK;
}

extern "C" int fm_ctx_create(fm_ctx** out, int device) {
    if (!out) return fail(FM_E_ARG, "null output");
    HIPCHK(hipSetDevice(device));
    fm_ctx* c = new fm_ctx();
    c->device = device;
    c->lu_schedule = LU_STREAM;
    c->res_schedule = RES_AUTO;
    c->lu_batch = 1;
    c->lu_fp32 = 0;
    c->graph_on = 1;
    c->cur = nullptr;
    c->fork_ev = nullptr;
    c->defer_dw = 0;
    c->res_kcol_n = 0;
    c->lu_flow_grid = 0;
    if (int rc = resolve_scratch(device)) { delete c; return rc; }
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return fail(FM_E_HIP, "stream: %s", hipGetErrorString(e)); }
    c->cur = c->stream;
    HIPCHK(hipEventCreate(&c->e0));
    HIPCHK(hipEventCreate(&c->e1));
    *out = c;
    return FM_OK;
}

static void free_zone(Zone* z) {
    double* bufs[] = {z->ro[0], z->ro[1], z->ro[2], z->ro_scratch, z->k.xmut, z->k.drodm,
                      z->k.coe, z->k.wig, z->k.ti, (double*)z->k.x, (double*)z->k.y,
                      (double*)z->k.z, z->norm_part, z->norm_out, z->norm_hist};
    for (double* b : bufs) if (b) (void)hipFree(b);
    for (BcRec& b : z->bcs) if (b.pext) (void)hipFree(b.pext);
    if (z->lu_dws) (void)hipFree(z->lu_dws);
    if (z->lu_sync) (void)hipFree(z->lu_sync);
    if (z->lu_xch) (void)hipFree(z->lu_xch);
    if (z->adj_buf) (void)hipFree(z->adj_buf);
    if (z->lu_fx) (void)hipFree(z->lu_fx);
    if (z->lu_stx) (void)hipFree(z->lu_stx);
    if (z->lu_skw) (void)hipFree(z->lu_skw);
    if (z->flow_work) (void)hipFree(z->flow_work);
    delete z;
}

extern "C" int fm_ctx_destroy(fm_ctx* c) {
    if (!c) return FM_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (Zone* z : c->zones) free_zone(z);
    if (c->tr_pts) (void)hipFree(c->tr_pts);
    for (auto& l : c->lu_lists) if (l.work) (void)hipFree(l.work);
    for (auto* g : c->graphs) {
        if (g->state == 2) (void)hipGraphExecDestroy(g->exec);
        for (auto& t : g->tabs) (void)hipFree(t.second);
        delete g;
    }
    if (c->lu_tab) (void)hipFree(c->lu_tab);
    if (c->co_tab) (void)hipFree(c->co_tab);
    if (c->flow_ctr) (void)hipFree(c->flow_ctr);
    for (hipStream_t s2 : c->side) (void)hipStreamDestroy(s2);
    for (hipEvent_t e2 : c->side_ev) (void)hipEventDestroy(e2);
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    for (int q = 0; q < 2; q++) {
        if (c->lu_host[q]) (void)hipHostFree(c->lu_host[q]);
        if (c->lu_ev[q]) (void)hipEventDestroy(c->lu_ev[q]);
        if (c->co_host[q]) (void)hipHostFree(c->co_host[q]);
        if (c->co_ev[q]) (void)hipEventDestroy(c->co_ev[q]);
    }
    (void)hipEventDestroy(c->e0);
    (void)hipEventDestroy(c->e1);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return FM_OK;
}

static void fill_kz_params(Zone* z) {
    const int* p = z->p.data();
    const double* r = z->pr.data();
    KZ& k = z->k;
    k.dtc = r[R_DTC]; k.gam = r[R_GAMMA]; k.cv = r[R_CVINF]; k.roinf = r[R_ROINF];
    k.pinf = r[R_PINF]; k.vinf = r[R_VINF]; k.tinf = r[R_TINF]; k.prandt = r[R_PRANDT];
    k.xmul0 = r[R_XMUL0]; k.temp0 = r[R_TEMP0]; k.cs = r[R_CS]; k.cfl = r[R_CFL];
    k.psiroe = r[R_PSIROE]; k.wig_amp = r[R_WIG_AMP];
    k.kflu = p[KFLUDOM]; k.slope = p[SLOPE]; k.iflow = p[IFLOW]; k.itypcp = p[ITYPCP];
    k.dtloc = p[DTLOC]; k.rk = p[RK]; k.nssiter = p[NSSITER]; k.sensortype = p[SENSORTYPE];
}

static int dmalloc(double** p, long n) {
    if (n <= 0) n = 1;
    if (hipMalloc((void**)p, sizeof(double) * n) != hipSuccess)
        return fail(FM_E_NOMEM, "hipMalloc of %ld doubles failed", n);
    if (hipMemset(*p, 0, sizeof(double) * n) != hipSuccess) return fail(FM_E_HIP, "hipMemset");
    return FM_OK;
}

extern "C" int fm_zone_add(fm_ctx* c, const int* p, int n_int, const double* pr, int n_real,
                           const double* x, const double* y, const double* zc, int* zone_id) {
    if (!c || !p || !pr || !x || !y || !zc || n_int < 137 || n_real < 73)
        return fail(FM_E_ARG, "fm_zone_add: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    int typ = p[ITYPZONE];
    if (typ != 0 && typ != 2) return fail(FM_E_UNSUPPORTED, "zone type %d (3dhomo/2d) not implemented", typ);
    if (p[NEQ] != 5 || p[IFLOW] < 1 || p[IFLOW] > 2 || p[ILES] != 0)
        return fail(FM_E_UNSUPPORTED, "only 5-equation Euler / laminar NS is implemented");
    if (p[LALE] != 0) return fail(FM_E_UNSUPPORTED, "moving meshes are not implemented");
    if (p[EXPLOC] != 0) return fail(FM_E_UNSUPPORTED, "explicit local time stepping is not implemented");
    if (p[ITYPCP] <= 1 && (p[IMPLICITSOLVER] != 0 || p[NB_RELAX] != 1 || p[LU_MATCH] != 0))
        return fail(FM_E_UNSUPPORTED, "implicit solver must be LU-SGS without relaxation or overlap");
    if (p[ITYPCP] == 2 && p[DTLOC] == 1) return fail(FM_E_UNSUPPORTED, "explicit scheme with local time step");
    int kf = p[KFLUDOM], sl = p[SLOPE];
    if (!((kf == 5 && sl >= 1 && sl <= 3) || ((kf == 1 || kf == 2) && sl == 2)))
        return fail(FM_E_UNSUPPORTED, "scheme %d / slope %d not implemented", kf, sl);
    if (p[NIJK + 3] != 2 || p[NIJK + 4] != 2) return fail(FM_E_UNSUPPORTED, "two ghost layers required");
    // cache blocking (CACHEBLCKI/J/K): any j and k block sizes give the reference's per-cell sum
    // order; one block in i (the reference default 2048 covers any i extent below 4096)
    if (p[CACHEBLCKI] < 1 || p[CACHEBLCKJ] < 1 || p[CACHEBLCKK] < 1)
        return fail(FM_E_ARG, "cache block sizes must be positive");
    if (p[IJKV] / p[CACHEBLCKI] > 1)
        return fail(FM_E_UNSUPPORTED, "more than one cache block in i (ni %d, CACHEBLCKI %d)", p[IJKV],
                    p[CACHEBLCKI]);
    Zone* z = new Zone();
    memset(&z->k, 0, sizeof(KZ));
    z->p.assign(p, p + n_int);
    z->pr.assign(pr, pr + n_real);
    z->ro[0] = z->ro[1] = z->ro[2] = nullptr; z->ro_scratch = nullptr;
    z->lu_dws = nullptr; z->lu_sync = nullptr; z->lu_xch = nullptr; z->adj_buf = nullptr;
    z->lu_fx = nullptr; z->lu_stx = nullptr; z->lu_skw = nullptr; z->dw_pending = 0; z->w_rows = nullptr; z->n_known = 0; z->flow_work = nullptr; z->flow_n = 0; z->flow_box = Box{0, -1, 0, -1, 0, -1};
    z->norm_part = z->norm_out = z->norm_hist = nullptr; z->have_norms = 0; z->have_cfl = 0;
    z->rec_lo = 1; z->rec_hi = 0;
    KZ& k = z->k;
    k.ni = p[IJKV]; k.nj = p[IJKV + 1]; k.nk = p[IJKV + 2];
    k.nit = p[NIJK]; k.njt = p[NIJK + 1]; k.nkt = p[NIJK + 2];
    k.nf = p[NIJK + 3]; k.nfk = p[NIJK + 4];
    k.nd = p[NDIMDX]; k.incj = p[NIJK]; k.inck = (long)p[NIJK] * p[NIJK + 1];
    k.typ = typ;
    k.mi = p[NIJK_MTR]; k.mj = p[NIJK_MTR + 1]; k.mk = p[NIJK_MTR + 2];
    k.mf = p[NIJK_MTR + 3]; k.mfk = p[NIJK_MTR + 4]; k.ndm = p[NDIMDX_MTR];
    k.xi = p[NIJK_XYZ]; k.xj = p[NIJK_XYZ + 1]; k.xf = p[NIJK_XYZ + 3]; k.xfk = p[NIJK_XYZ + 4];
    // cache blocks of the flux pass for one thread owning the zone (WORK_DISTRIBUTION_BEGIN.for:
    // the CACHEBLCK* slots as they are; crsdm_scater.for:42-44: max(1, n / size) blocks per
    // direction, the last one taking the remainder)
    k.cbj = p[CACHEBLCKJ]; k.cbk = p[CACHEBLCKK];
    k.nbj = k.nj / k.cbj > 1 ? k.nj / k.cbj : 1;
    k.nbk = k.nk / k.cbk > 1 ? k.nk / k.cbk : 1;
    fill_kz_params(z);
    // boundary-condition table (FastC._BCcompact layout)
    int pt_bcs = p[PT_BC];
    int nb = pt_bcs > 0 ? p[pt_bcs] : 0;
    for (int n = 0; n < nb; n++) {
        int pt = p[pt_bcs + 1 + n];
        BcRec b; b.type = p[pt + BC_TYPE]; b.idir = p[pt + BC_IDIR]; b.pext = nullptr;
        for (int q = 0; q < 6; q++) b.fen[q] = p[pt + BC_FEN + q];
        int t = b.type, nbdata = p[pt + BC_NBDATA];
        const int* size_data = p + pt + BC_NBDATA + 1;
        const double* data = nbdata ? pr + p[pt_bcs + 1 + n + nb] : nullptr;
        if (b.idir < 1 || b.idir > 6) { free_zone(z); return fail(FM_E_ARG, "BC direction %d", b.idir); }
        if (t == 0) b.kind = 0;
        else if (t == 11) b.kind = 2;
        else if ((t == 4 || t == 6) && p[IFLOW] > 1) b.kind = 1;
        else if (t == 3 || t == 5 || (t == 4 && p[IFLOW] == 1)) b.kind = 3;
        else if (t == 6) b.kind = -1;    // viscous wall in an Euler zone: BCzone.cpp does nothing
        else if (t == 1 || (t == 13 && nbdata == 0)) b.kind = 4;
        else if (t == 10) b.kind = 5;
        else if (t == 16 && (nbdata == 1 || nbdata == 2)) b.kind = 6;
        else {
            free_zone(z);
            return fail(FM_E_UNSUPPORTED, "boundary condition type %d (with %d data arrays) not implemented", t, nbdata);
        }
        // reference state: free stream unless the BC carries its own data
        if (nbdata == 0) {
            b.state[0] = pr[R_ROINF]; b.state[1] = pr[R_VXINF] * pr[R_ROINF];
            b.state[2] = pr[R_VYINF] * pr[R_ROINF]; b.state[3] = pr[R_VZINF] * pr[R_ROINF];
            b.state[4] = pr[R_PINF] / (pr[R_GAMMA] - 1.) +
                         0.5 * (pr[R_VXINF] * b.state[1] + pr[R_VYINF] * b.state[2] + pr[R_VZINF] * b.state[3]);
        } else {
            for (int q = 0; q < 5; q++) b.state[q] = 0.;
        }
        if (b.kind == 6) {
            const int* fen = b.fen;
            if (b.idir <= 2) { b.inc_bc[0] = fen[3] - fen[2] + 1; b.inc_bc[1] = fen[2]; b.inc_bc[2] = fen[4]; }
            else if (b.idir <= 4) { b.inc_bc[0] = fen[1] - fen[0] + 1; b.inc_bc[1] = fen[0]; b.inc_bc[2] = fen[4]; }
            else { b.inc_bc[0] = fen[1] - fen[0] + 1; b.inc_bc[1] = fen[0]; b.inc_bc[2] = fen[2]; }
            // one exit pressure per face point (a leading 1-value array is the mesh-motion
            // coefficient and is skipped, as BCzone.cpp does)
            const double* pe = size_data[0] == 1 ? data + 1 : data;
            long npe = size_data[0] == 1 ? size_data[1] : size_data[0];
            long need = (long)(b.fen[1] - b.fen[0] + 1) * (b.fen[3] - b.fen[2] + 1) * (b.fen[5] - b.fen[4] + 1);
            if (npe < need || (pe - pr) + npe > n_real) {
                free_zone(z);
                return fail(FM_E_ARG, "BCOutpres: %ld pressure values for %ld face points", npe, need);
            }
            double* dp = nullptr;
            int rc2 = dmalloc(&dp, npe);
            if (rc2 != FM_OK) { free_zone(z); return rc2; }
            b.pext = dp;
            z->bcs.push_back(b);
            if (hipMemcpy(dp, pe, sizeof(double) * npe, hipMemcpyHostToDevice) != hipSuccess) {
                free_zone(z);
                return fail(FM_E_HIP, "BCOutpres data upload");
            }
            continue;
        }
        z->bcs.push_back(b);
    }
    long nd = k.nd, nxyz = p[NDIMDX_XYZ];
    long ndm = k.ndm, nmtr = ndm * (2 * p[NEQ_IJ] + p[NEQ_K] + 1);
    z->n_metric = nmtr;
    int rc = FM_OK;
    double *xd, *yd, *zd, *mtr;
    if ((rc = dmalloc(&z->ro[0], 5 * nd)) || (rc = dmalloc(&z->ro[1], 5 * nd)) ||
        (rc = dmalloc(&z->ro[2], 5 * nd)) || (rc = dmalloc(&z->ro_scratch, 5 * nd)) ||
        (p[ITYPCP] <= 1 && (rc = dmalloc(&z->lu_dws, 5 * nd))) ||
        (rc = dmalloc(&k.xmut, nd)) || (rc = dmalloc(&k.drodm, 5 * nd)) ||
        (rc = dmalloc(&k.coe, (long)p[NEQ_COE] * nd)) || (rc = dmalloc(&k.wig, 3 * nd)) ||
        (rc = dmalloc(&mtr, nmtr)) || (rc = dmalloc(&xd, nxyz)) || (rc = dmalloc(&yd, nxyz)) ||
        (rc = dmalloc(&zd, nxyz))) {
        free_zone(z);
        return rc;
    }
    if (hipMalloc((void**)&z->lu_sync, 4 * sizeof(int)) != hipSuccess ||
        hipMemset(z->lu_sync, 0, 4 * sizeof(int)) != hipSuccess) {
        free_zone(z);
        return fail(FM_E_NOMEM, "hipMalloc of the LU counters failed");
    }
    k.ti = mtr; k.tj = mtr + ndm * p[NEQ_IJ]; k.tk = mtr + 2 * ndm * p[NEQ_IJ];
    k.vol = mtr + ndm * (2 * p[NEQ_IJ] + p[NEQ_K]);
    k.x = xd; k.y = yd; k.z = zd;
    z->norm_blocks = 512;
    if ((rc = dmalloc(&z->norm_part, 10 * z->norm_blocks)) || (rc = dmalloc(&z->norm_out, 16)) ||
        (rc = dmalloc(&z->norm_hist, 10 * (long)(k.nssiter > 0 ? k.nssiter : 1)))) {
        free_zone(z);
        return rc;
    }
    HIPCHK(hipMemcpy(xd, x, sizeof(double) * nxyz, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(yd, y, sizeof(double) * nxyz, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(zd, zc, sizeof(double) * nxyz, hipMemcpyHostToDevice));
    c->zones.push_back(z);
    c->graph_epoch++;
    if (zone_id) *zone_id = (int)c->zones.size() - 1;
    return FM_OK;
}

static Zone* get_zone(fm_ctx* c, int zone) {
    if (!c || zone < 0 || zone >= (int)c->zones.size()) return nullptr;
    return c->zones[zone];
}

extern "C" int fm_zone_set_param_real(fm_ctx* c, int zone, const double* pr, int n_real) {
    Zone* z = get_zone(c, zone);
    if (!z || !pr || n_real < 73) return fail(FM_E_ARG, "fm_zone_set_param_real: bad arguments");
    z->pr.assign(pr, pr + n_real);
    fill_kz_params(z);
    c->graph_epoch++;        // captured steps hold the old values in their kernel arguments
    return FM_OK;
}

extern "C" int fm_zone_metric(fm_ctx* c, int zone) {
    Zone* z = get_zone(c, zone);
    if (!z) return fail(FM_E_ARG, "bad zone %d", zone);
    HIPCHK(hipSetDevice(c->device));
    launch_metric(z->k, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return FM_OK;
}

static void resolve_dw(fm_ctx* c, Zone* z);

static double* field_ptr(Zone* z, int field, long* n) {
    long nd = z->k.nd;
    switch (field) {
    case FM_FIELD_RO: *n = 5 * nd; return z->ro[0];
    case FM_FIELD_RO_M1: *n = 5 * nd; return z->ro[1];
    case FM_FIELD_RO_P1: *n = 5 * nd; return z->ro[2];
    case FM_FIELD_MUT: *n = nd; return z->k.xmut;
    case FM_FIELD_DRODM: *n = 5 * nd; return z->k.drodm;
    case FM_FIELD_COE: *n = (long)z->p[NEQ_COE] * nd; return z->k.coe;
    case FM_FIELD_METRIC: *n = z->n_metric; return z->k.ti;
    case FM_FIELD_WIG: *n = 3 * nd; return z->k.wig;
    }
    return nullptr;
}

extern "C" int fm_zone_set_field(fm_ctx* c, int zone, int field, const double* host, long n) {
    Zone* z = get_zone(c, zone);
    if (!z || !host) return fail(FM_E_ARG, "fm_zone_set_field: bad arguments");
    long sz; double* d = field_ptr(z, field, &sz);
    if (!d || n != sz) return fail(FM_E_ARG, "field %d: size %ld expected %ld", field, n, sz);
    z->w_rows = nullptr;
    if (field == FM_FIELD_DRODM) z->dw_pending = 0;
    if (field == FM_FIELD_METRIC) {
        z->n_known = 0;
        c->graph_epoch++;     // captured steps hold whether the normal rows were in place
    }
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(d, host, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FM_OK;
}

static int check_lu(fm_ctx* c);

extern "C" int fm_zone_get_field(fm_ctx* c, int zone, int field, double* host, long n) {
    Zone* z = get_zone(c, zone);
    if (!z || !host) return fail(FM_E_ARG, "fm_zone_get_field: bad arguments");
    long sz; double* d = field_ptr(z, field, &sz);
    if (!d || n != sz) return fail(FM_E_ARG, "field %d: size %ld expected %ld", field, n, sz);
    HIPCHK(hipSetDevice(c->device));
    if (field == FM_FIELD_DRODM) resolve_dw(c, z);
    HIPCHK(hipStreamSynchronize(c->stream));
    // a sweep whose bounded wait gave up leaves wrong values: report it before handing them out
    if (int rc = check_lu(c)) return rc;
    HIPCHK(hipMemcpy(host, d, sizeof(double) * n, hipMemcpyDeviceToHost));
    return FM_OK;
}

// ---------------------------------------------------------------------------------------
// BC windows: indice_cl_sdm.for for a thread box equal to the zone, npass = 0
static bool cl_window(const Zone* z, int idir, const int* fen, int w[6]) {
    const KZ& k = z->k;
    int sdm[6] = {1, k.ni, 1, k.nj, 1, k.nk};
    if (!(sdm[0] <= fen[1] && sdm[1] >= fen[0] && sdm[2] <= fen[3] && sdm[3] >= fen[2] &&
          sdm[4] <= fen[5] && sdm[5] >= fen[4])) return false;
    for (int a = 0; a < 3; a++) {
        w[2 * a] = fen[2 * a] > sdm[2 * a] ? fen[2 * a] : sdm[2 * a];
        w[2 * a + 1] = fen[2 * a + 1] < sdm[2 * a + 1] ? fen[2 * a + 1] : sdm[2 * a + 1];
    }
    int f = k.nf, fk = k.nfk;
    switch (idir) {
    case 1: w[1] = w[0] - 1; w[0] = w[0] - f; break;
    case 2: w[1] = w[0] + f; w[0] = w[0] + 1; break;
    case 3: w[3] = w[2] - 1; w[2] = w[2] - f; break;
    case 4: w[3] = w[2] + f; w[2] = w[2] + 1; break;
    case 5: w[5] = w[4] - 1; w[4] = w[4] - fk; break;
    default: w[5] = w[4] + fk; w[4] = w[4] + 1; break;
    }
    return true;
}

// BCzone.cpp dispatch (lrhs = 1: RHS of the implicit step, returns the LU range;
// lrhs = 0: ghost-cell fill of a state).
static void bczone(fm_ctx* c, Zone* z, int lrhs, double* rop, int shift_lu[6], int* lcorner_out) {
    const KZ& k = z->k;
    int ind_dm[6] = {1, k.ni, 1, k.nj, 1, k.nk};
    int lskip[6] = {1, 1, 1, 1, 1, 1}, ificmax[6] = {0, 0, 0, 0, 0, 0};
    int lcorner = 1;
    for (const BcRec& b : z->bcs) {
        int ip = b.idir - 1, t = b.type;
        if (lrhs == 1 && ((t >= 3 && t <= 7) || t == 12 || t == 18 || t == 30)) {
            lskip[ip] = 0; lcorner = 0;
            if ((t >= 3 && t <= 7) || t == 12) ificmax[ip] = 1;
        }
    }
    if (lrhs == 1) {
        for (int a = 0; a < 3; a++) {
            shift_lu[2 * a] = ind_dm[2 * a] - (1 - lskip[2 * a]) * ificmax[2 * a];
            shift_lu[2 * a + 1] = ind_dm[2 * a + 1] + (1 - lskip[2 * a + 1]) * ificmax[2 * a + 1];
        }
        for (int idir = 0; idir < 6; idir++) {
            if (lskip[idir] != 0) continue;
            int fen[6];
            for (int l = 0; l < 6; l++) fen[l] = ind_dm[l];
            if (idir % 2 == 0) { fen[idir] = 1; fen[idir + 1] = 1; } else fen[idir - 1] = fen[idir];
            BcArgs a; a.idir = idir + 1; a.kind = 0; a.lrhs = 1;
            if (cl_window(z, idir + 1, fen, a.w)) launch_bc_face(k, rop, a, c->cur);
        }
    }
    for (const BcRec& b : z->bcs) {
        if (!(lskip[b.idir - 1] == 0 || lrhs == 0)) continue;
        int t = b.type;
        if (lrhs == 1 && (t == 0 || t == 1 || t == 2 || t == 10 || t == 11 || t == 14 || t == 16 ||
                          t == 17 || t == 20 || t == 21)) t = 0;
        BcArgs a; a.idir = b.idir; a.lrhs = lrhs;
        a.kind = t == 0 ? 0 : b.kind;
        if (a.kind < 0) continue;
        a.tijk = b.idir <= 2 ? k.ti : b.idir <= 4 ? k.tj : k.tk;
        a.neq_mtr = b.idir <= 4 ? z->p[NEQ_IJ] : z->p[NEQ_K];
        for (int q = 0; q < 5; q++) a.state[q] = b.state[q];
        a.pext = b.pext;
        for (int q = 0; q < 3; q++) a.inc_bc[q] = b.inc_bc[q];
        if (cl_window(z, b.idir, b.fen, a.w)) launch_bc_face(k, rop, a, c->cur);
    }
    if (lcorner_out) *lcorner_out = lcorner;
}

static void coins(fm_ctx* c, Zone* z, double* rop, const int* ind) {
    Box L;
    const KZ& k = z->k;
    L.i0 = ind[0] > 1 ? ind[0] : 1; L.j0 = ind[2] > 1 ? ind[2] : 1; L.k0 = ind[4] > 1 ? ind[4] : 1;
    L.i1 = ind[1] < k.ni ? ind[1] : k.ni; L.j1 = ind[3] < k.nj ? ind[3] : k.nj; L.k1 = ind[5] < k.nk ? ind[5] : k.nk;
    launch_coins(k, rop, L, c->cur);
}

// The pipelined LU schedule's exchange buffer, sized for the largest LU box of the zone (one
// ghost layer past each face, where walls extend it).
static int ensure_exchange(fm_ctx* c, Zone* z) {
    if (c->lu_schedule != LU_PIPELINE || z->lu_xch) return FM_OK;
    long n = lu_exchange_size(z->k.ni + 2, z->k.nj + 2, z->k.nk + 2);
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMalloc(&z->lu_xch, sizeof(double) * n));
    return FM_OK;
}

// Dataflow LU schedule: the face exchange (sized for the largest LU box, interior plus a wall
// layer per side) and the block order of the LU box b (rebuilt when the box changes).
// The residual schedule of a zone: the context's, or with RES_AUTO six faces per cell on
// Cartesian zones and columns up k on curvilinear ones (the faster of the two on each mesh kind,
// measured on TGV 256^3 and the 257 x 129 x 33 channel)
static int res_schedule_of(const fm_ctx* c, const Zone* z) {
    if (c->res_schedule != RES_AUTO) return c->res_schedule;
    return z->k.typ == 2 ? RES_CELLS : RES_KCOL;
}

// Streaming LU schedule of a zone: the pencil exchange, sized for the largest LU box and
// filled with the pending pattern once (the sweeps leave it that way).
static bool uses_stream(const fm_ctx* c, const Zone* z) {
    (void)z;
    return c->lu_schedule == LU_STREAM && !c->lu_fp32;
}

static int ensure_stream(fm_ctx* c, Zone* z, LuArgs& la) {
    la.stx = nullptr;
    la.skw = nullptr;
    if (!uses_stream(c, z)) return FM_OK;
    if (!z->lu_stx) {
        const KZ& k = z->k;
        long n = lu_stream_exchange_size(k.ni + 2, k.nj + 2, k.nk + 2);
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMalloc(&z->lu_stx, sizeof(double) * n));
        fill_pending(z->lu_stx, n, c->cur);
        HIPCHK(hipGetLastError());
    }
    if (!z->lu_skw) {
        const KZ& k = z->k;
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMalloc(&z->lu_skw, sizeof(double) * lu_skew_size(k.typ == 2, k.ni + 2, k.nj + 2, k.nk + 2)));
    }
    la.stx = z->lu_stx;
    la.skw = z->lu_skw;
    return FM_OK;
}

// The dataflow face exchange of a zone, sized for the largest LU box (interior plus a wall layer
// per side) and filled with the pending pattern once: the sweeps give back every slot they read.
static int ensure_flow_exchange(fm_ctx* c, Zone* z) {
    if (z->lu_fx) return FM_OK;
    const KZ& k = z->k;
    long n = lu_flow_exchange_size(k.typ == 2, k.ni + 2, k.nj + 2, k.nk + 2);
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMalloc(&z->lu_fx, sizeof(double) * n));
    fill_pending(z->lu_fx, n, c->cur);
    HIPCHK(hipGetLastError());
    return FM_OK;
}

static int ensure_dataflow(fm_ctx* c, Zone* z, const Box& b, LuArgs& la) {
    // the dataflow schedule, and what the streaming schedule falls back to (curvilinear zones,
    // FP32 sweeps)
    if (c->lu_schedule != LU_DATAFLOW && !(c->lu_schedule == LU_STREAM && !uses_stream(c, z))) return FM_OK;
    const KZ& k = z->k;
    if (int rc = ensure_flow_exchange(c, z)) return rc;
    const Box& f = z->flow_box;
    if (!z->flow_work || f.i0 != b.i0 || f.i1 != b.i1 || f.j0 != b.j0 || f.j1 != b.j1 || f.k0 != b.k0 || f.k1 != b.k1) {
        std::vector<LuWork> work;
        std::vector<int> offs;
        lu_work_list(k.typ == 2, std::vector<Box>{b}, work, offs);
        HIPCHK(hipDeviceSynchronize());
        if (z->flow_work) { HIPCHK(hipFree(z->flow_work)); z->flow_work = nullptr; }
        HIPCHK(hipMalloc(&z->flow_work, sizeof(LuWork) * (work.empty() ? 1 : work.size())));
        if (!work.empty())
            HIPCHK(hipMemcpy(z->flow_work, work.data(), sizeof(LuWork) * work.size(), hipMemcpyHostToDevice));
        z->flow_n = (int)work.size();
        z->flow_box = b;
    }
    la.fx = z->lu_fx; la.flow_work = z->flow_work; la.flow_n = z->flow_n;
    // a box a few blocks wide in j and k is latency-bound: its blocks start before their halo is
    // complete and take the products step by step; wider boxes keep every resident workgroup busy
    // with blocks whose halo is complete (measured on 256 x {8..256}^2 boxes)
    const long nbjk = (long)((b.j1 - b.j0 + 1 + 7) / 8) * ((b.k1 - b.k0 + 1 + 7) / 8);
    la.flow_step = nbjk < 256;
    return FM_OK;
}

// Physical boundary conditions then edges and corners of one zone (gsdr3.cpp:436-439, bcs.cpp).
static void fill_ghosts(fm_ctx* c, Zone* z, double* rop) {
    int shift[6], sdm[6] = {1, z->k.ni, 1, z->k.nj, 1, z->k.nk};
    bczone(c, z, 0, rop, shift, nullptr);
    coins(c, z, rop, sdm);
}

// gsdr3.cpp:152-173: the level that receives the transfers and boundary conditions
static double* level_cl(const Zone* z, int nitcfg) {
    if (nitcfg == 1 || z->k.itypcp <= 1 || nitcfg % 2 == 1) return z->ro[2];
    return z->ro[0];
}

// Injection transfers of every raccord on the given field of every zone
// (transfert_multiblock.cpp:9-14 -> setInterpTransfersFast).
static int transfer(fm_ctx* c, double* const* fields) {
    if (c->tr_n == 0) return FM_OK;
    TransferArgs a;
    memset(&a, 0, sizeof a);
    for (size_t n = 0; n < c->zones.size(); n++) { a.f[n] = fields[n]; a.nd[n] = c->zones[n]->k.nd; }
    if (!c->tr_in_order) {
        launch_transfer(c->tr_pts, c->tr_n, a, c->stream);
    } else {
        for (size_t r = 0; r < c->tr_off.size(); r++)
            launch_transfer_part(c->tr_pts + c->tr_off[r], c->tr_cnt[r], a, c->tr_rdup[r], c->stream);
    }
    HIPCHK(hipGetLastError());
    return FM_OK;
}

// Roe wall-flux correction windows in BC order (correct_flux.for:52-133): wall types 3 (slip:
// bflwallslip.for) and 4, 6, 7, 12 (bflwall.for); the window's normal range is the face index.
static int wall_corrections(const Zone* z, RhsArgs& a) {
    a.ncorr = 0;
    a.ck_vent = z->p[NEQ_VENT] != 3 ? 0. : 1.;
    a.c_ale = 1. * (z->p[LALE] < 1 ? z->p[LALE] : 1);
    for (const BcRec& b : z->bcs) {
        int t = b.type;
        bool slip = t == 3, wall = t == 4 || t == 6 || t == 7 || t == 12;
        if (!slip && !wall) continue;
        int w[6];
        if (!cl_window(z, b.idir, b.fen, w)) continue;
        switch (b.idir) {
        case 1: w[0] = 1; w[1] = 1; break;
        case 2: w[1] = w[0]; break;
        case 3: w[2] = 1; w[3] = 1; break;
        case 4: w[3] = w[2]; break;
        case 5: w[4] = 1; w[5] = 1; break;
        default: w[5] = w[4]; break;
        }
        if (a.ncorr == kMaxWallCorr)
            return fail(FM_E_UNSUPPORTED, "more than %d Roe wall windows", kMaxWallCorr);
        WallCorr& c = a.corr[a.ncorr++];
        c.idir = b.idir; c.slip = slip;
        c.neq_mtr = b.idir <= 4 ? z->p[NEQ_IJ] : z->p[NEQ_K];
        for (int q = 0; q < 6; q++) c.c[q] = w[q];
    }
    return FM_OK;
}

static bool same_box(const Box& a, const Box& b) {
    return a.i0 == b.i0 && a.i1 == b.i1 && a.j0 == b.j0 && a.j1 == b.j1 && a.k0 == b.k0 && a.k1 == b.k1;
}
static bool lu_interior(const KZ& k, const LuArgs& la) {
    return la.i0 >= 1 && la.i1 <= k.ni && la.j0 >= 1 && la.j1 <= k.nj && la.k0 >= 1 && la.k1 <= k.nk;
}

// The copies a streaming LU-SGS of zone z can skip: dW stays in the rows inside fm_compute (copied
// back when drodm is read); the state rows already hold rop if the last upper sweep wrote them for
// this box and the box is interior (the ghost fill does not touch it); 3dfull: the normal rows,
// written by an earlier pack of the same box.  la carries the zone's rows (la.skw) here.
static void lu_stream_flags(const fm_ctx* c, const Zone* z, LuArgs& la) {
    const KZ& k = z->k;
    const Box lbx = {la.i0, la.i1, la.j0, la.j1, la.k0, la.k1};
    const bool st = lu_streams(k, la);
    la.keep_dw = !c->defer_dw;
    la.w_ready = st && lu_interior(k, la) && z->w_rows == la.rop && same_box(z->w_box, lbx);
    la.n_ready = st && k.typ != 2 && z->n_known && same_box(z->n_box, lbx);
}

// What the rows of zone z hold after its LU-SGS with la (any schedule: the others leave none).
static void lu_stream_after(Zone* z, const LuArgs& la) {
    const KZ& k = z->k;
    const Box lbx = {la.i0, la.i1, la.j0, la.j1, la.k0, la.k1};
    const bool st = lu_streams(k, la);
    if (!la.keep_dw && la.newton && st) { z->dw_pending = 1; z->dw_la = la; }
    if (st && k.typ != 2) { z->n_known = 1; z->n_box = lbx; }
    if (la.newton && lu_interior(k, la) && st) {
        z->w_rows = la.out == z->ro_scratch ? z->ro[2] : la.out;
        z->w_box = lbx;
    } else {
        z->w_rows = nullptr;
    }
}

// One sub-iteration of one zone (gsdr3.cpp:152-444 with navier_stokes_struct.for and lhs.cpp).
static int substep(fm_ctx* c, Zone* z, int nitcfg, int first_it, int lssiter_verif, int lexit_lu, int stop,
                   LuArgs* defer = nullptr) {
    KZ& k = z->k;
    hipStream_t st = c->cur;
    int implicit = k.itypcp <= 1;
    double* ro = z->ro[0]; double* ro_m1 = z->ro[1]; double* ro_p1 = z->ro[2];
    double *ro_ssiter, *ro_cl;
    if (nitcfg == 1) { ro_ssiter = ro; ro_cl = ro_p1; }
    else if (implicit) { ro_ssiter = ro_p1; ro_cl = ro_p1; }
    else if (nitcfg % 2 == 0) { ro_ssiter = ro_p1; ro_cl = ro; }
    else { ro_ssiter = ro; ro_cl = ro_p1; }
    int sdm[6] = {1, k.ni, 1, k.nj, 1, k.nk};
    int lamin = k.iflow == 2;
    if (nitcfg == 1) {
        if (implicit) {
            Box b = {1 - k.nf, k.ni + k.nf, 1 - k.nf, k.nj + k.nf, 1 - k.nfk, k.nk + k.nfk};
            launch_coe(k, ro_ssiter, first_it, k.dtloc ? 2 : 1, b, lamin, st);
        } else {
            Box b = {0, k.ni + 1, 0, k.nj + 1, 0, k.nk + 1};
            launch_coe(k, ro_ssiter, first_it, 0, b, lamin, st);
        }
    }
    // cptst3.for:114-119: CFL monitor at the first sub-iteration of a verification step
    if (nitcfg == 1 && lssiter_verif) {
        launch_cfl(k, ro_ssiter, z->norm_part, z->norm_blocks, z->norm_out + 10, st);
        z->have_cfl = 1;
    }
    if (k.kflu == 2 && nitcfg == 1) launch_wiggle(k, ro_ssiter, st);
    RhsArgs a;
    memset(&a, 0, sizeof a);
    a.schedule = res_schedule_of(c, z);
    a.kcol_n = c->res_kcol_n;
    if (k.kflu == 5) {
        int rc = wall_corrections(z, a);
        if (rc) return rc;
    }
    a.rop = ro_ssiter; a.rop_n = ro; a.rop_n1 = ro_m1;
    a.implicit = implicit; a.nitcfg = nitcfg;
    a.zero_init = (implicit || nitcfg == 1);
    a.coefH = 0.;
    if (k.rk == 3) { if (nitcfg == 2) a.coefH = -0.6830127018922193; else if (nitcfg == 3) a.coefH = -4. / 3.; }
    else if (k.rk == 2) { if (nitcfg == 2) a.coefH = -1.; }
    a.coefW = 1.;
    if (k.rk == 2 && nitcfg == 2) a.coefW = 0.5;
    if (k.rk == 3) a.coefW = nitcfg == 1 ? 0.5 : nitcfg == 2 ? 0.9106836025229591 : 0.3660254037844387;
    if (first_it == 0 || k.dtloc == 1) { a.c1 = -1.; a.c2 = 1.; a.c3 = 0.; }
    else { a.c1 = -1.5; a.c2 = 2.; a.c3 = -0.5; }
    if (nitcfg % 2 == 1) { a.rk_out = ro_p1; a.rk_in = ro; }
    else { a.rk_out = ro; a.rk_in = ro_p1; }
    if (launch_residual(k, a, st)) return fail(FM_E_UNSUPPORTED, "no residual kernel for this scheme");
    HIPCHK(hipGetLastError());
    z->dw_pending = 0;
    if (!implicit) z->w_rows = nullptr;       // the stage update rewrites the state
    if (stop == 1) return FM_OK;
    if (implicit) {
        int shift_lu[6], lcorner = 1;
        bczone(c, z, 1, k.drodm, shift_lu, &lcorner);
        if (lcorner == 0) coins(c, z, k.drodm, shift_lu);
        if (lssiter_verif == 1) {
            launch_norms(k, k.drodm, z->norm_part, z->norm_blocks, z->norm_hist + 10 * (nitcfg - 1), st);
        }
        if (stop == 2) return FM_OK;
        if (lexit_lu == 0) {
            LuArgs la;
            la.i0 = shift_lu[0]; la.i1 = shift_lu[1]; la.j0 = shift_lu[2]; la.j1 = shift_lu[3];
            la.k0 = shift_lu[4]; la.k1 = shift_lu[5];
            la.dw = k.drodm; la.newton = 1;
            la.dws = z->lu_dws; la.sync = z->lu_sync; la.schedule = c->lu_schedule; la.fp32 = c->lu_fp32;
    la.flow_grid = c->lu_flow_grid;
            if (int rc = ensure_exchange(c, z)) return rc;
            la.xch = z->lu_xch;
            la.fx = nullptr; la.flow_work = nullptr; la.flow_n = 0; la.flow_step = 0;
            {
                Box lb = {la.i0, la.i1, la.j0, la.j1, la.k0, la.k1};
                if (int rc = ensure_dataflow(c, z, lb, la)) return rc;
                // a batched zone runs the dataflow list: no streaming buffers for it
                if (defer) { la.stx = nullptr; la.skw = nullptr; }
                else if (int rc = ensure_stream(c, z, la)) return rc;
            }
            la.rop = ro_ssiter;
            if (ro_ssiter == ro_p1) {
                // in place (sub-iterations >= 2): the K terms of the upper sweep see the state
                // already updated on plane k+1 (invlu_u.for + mjr_newton per k plane)
                la.rop_k = z->ro_scratch; la.out = z->ro_scratch;
            } else {
                la.rop_k = ro_ssiter; la.out = ro_p1;
            }
            // (a batched zone: lu_batched sets the copies it skips, with its rows)
            lu_stream_flags(c, z, la);
            if (defer) { *defer = la; return FM_OK; }
            launch_lu(k, la, st);
            HIPCHK(hipGetLastError());
            lu_stream_after(z, la);
            if (la.out == z->ro_scratch) {
                // copy the updated cells back (the LU box, ghost layers it extends into included)
                Box lb = {la.i0, la.i1, la.j0, la.j1, la.k0, la.k1};
                launch_copy_box(k, z->ro_scratch, ro_p1, lb, st);
            }
        }
    }
    if (stop == 3) return FM_OK;
    if (lexit_lu == 0) fill_ghosts(c, z, ro_cl);
    HIPCHK(hipGetLastError());
    return FM_OK;
}

// A pipelined LU sweep whose bounded waits gave up has produced garbage: report it.
static int check_lu(fm_ctx* c) {
    for (Zone* z : c->zones) {
        LuArgs la;
        memset(&la, 0, sizeof la);
        la.schedule = LU_PIPELINE; la.sync = z->lu_sync;
        int f = lu_pipeline_failed(la, c->stream);
        if (f < 0) return fail(FM_E_HIP, "reading the LU error word failed");
        if (f > 0) {
            // a sweep that gave up may leave handed-over values in the exchanges
            const KZ& k = z->k;
            if (z->lu_stx)
                fill_pending(z->lu_stx, lu_stream_exchange_size(k.ni + 2, k.nj + 2, k.nk + 2), c->stream);
            if (z->lu_fx)
                fill_pending(z->lu_fx, lu_flow_exchange_size(k.typ == 2, k.ni + 2, k.nj + 2, k.nk + 2), c->stream);
            (void)hipStreamSynchronize(c->stream);
            return fail(FM_E_HIP, "LU-SGS sweep timed out waiting for a neighbour tile or block");
        }
    }
    return FM_OK;
}

extern "C" int fm_zone_substep(fm_ctx* c, int zone, int nitcfg, int nitrun, int first_it,
                               int lssiter_verif, int lexit_lu, int stop) {
    (void)nitrun;
    Zone* z = get_zone(c, zone);
    if (!z) return fail(FM_E_ARG, "bad zone %d", zone);
    if (nitcfg < 1 || nitcfg > z->k.nssiter) return fail(FM_E_ARG, "nitcfg %d out of range", nitcfg);
    HIPCHK(hipSetDevice(c->device));
    int rc = substep(c, z, nitcfg, first_it, lssiter_verif, lexit_lu, stop);
    if (rc) return rc;
    if (lssiter_verif == 1 && z->k.itypcp <= 1 && stop != 1) {
        // a later sub-iteration extends the record of the step, an earlier one starts a new one
        if (z->have_norms && nitcfg == z->rec_hi + 1) z->rec_hi = nitcfg;
        else { z->rec_lo = z->rec_hi = nitcfg; }
        z->have_norms = 1;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return check_lu(c);
}

// compute_effort.cpp:200-258 -> bceffort.for for one window of one zone: the per-face fields
// (21 x nfaces) and the 12 sums, summed here in the reference's loop order (k, j, i).
extern "C" int fm_wall_effort(fm_ctx* c, int zone, int level, int idir, const int* window,
                              const double* pos, double* effort, double* fields) {
    Zone* z = get_zone(c, zone);
    if (!z || level < 0 || level > 2 || idir < 1 || idir > 6 || !window || !effort)
        return fail(FM_E_ARG, "fm_wall_effort: bad arguments");
    const KZ& k = z->k;
    int d = (idir - 1) / 2;
    int lo[3] = {1, 1, 1}, hi[3] = {k.ni, k.nj, k.nk};
    hi[d] = hi[d] + 1;                       // face indices run to n + 1 in the normal direction
    for (int a = 0; a < 3; a++)
        if (window[2 * a] < lo[a] || window[2 * a + 1] > hi[a] || window[2 * a] > window[2 * a + 1])
            return fail(FM_E_ARG, "fm_wall_effort: window outside the zone");
    if (window[2 * d] != window[2 * d + 1]) return fail(FM_E_ARG, "fm_wall_effort: window is not one face plane");
    HIPCHK(hipSetDevice(c->device));
    long nf = (long)(window[1] - window[0] + 1) * (window[3] - window[2] + 1) * (window[5] - window[4] + 1);
    double* d_out = nullptr;
    HIPCHK(hipMalloc(&d_out, sizeof(double) * 33 * nf));
    EffArgs a;
    a.idir = idir;
    for (int q = 0; q < 6; q++) a.w[q] = window[q];
    for (int q = 0; q < 3; q++) a.pos[q] = pos ? pos[q] : 0.;
    a.rop = z->ro[level]; a.out = d_out;
    if (launch_effort(k, a, c->stream)) { (void)hipFree(d_out); return fail(FM_E_UNSUPPORTED, "no effort kernel for this scheme"); }
    std::vector<double> h((size_t)33 * nf);
    hipError_t e = hipMemcpyAsync(h.data(), d_out, sizeof(double) * 33 * nf, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail(FM_E_HIP, "fm_wall_effort: %s", hipGetErrorString(e));
    for (int q = 0; q < 12; q++) effort[q] = 0.;
    for (long t = 0; t < nf; t++)
        for (int q = 0; q < 12; q++) effort[q] = effort[q] + h[(size_t)(21 + q) * nf + t];
    if (fields) memcpy(fields, h.data(), sizeof(double) * 21 * nf);
    return FM_OK;
}

// fm_wall_effort without any allocation: the kernels write into the device scratch array and a
// one-thread kernel forms the 12 sums in the same face order as fm_wall_effort's host loop, so
// both give the same bits.  Calls that overlap on one device share the array and race.
extern "C" int fm_wall_effort_noalloc(fm_ctx* c, int zone, int level, int idir, const int* window,
                                      const double* pos, double* effort, double* fields) {
    Zone* z = get_zone(c, zone);
    if (!z || level < 0 || level > 2 || idir < 1 || idir > 6 || !window || !effort)
        return fail(FM_E_ARG, "fm_wall_effort_noalloc: bad arguments");
    const KZ& k = z->k;
    int d = (idir - 1) / 2;
    int lo[3] = {1, 1, 1}, hi[3] = {k.ni, k.nj, k.nk};
    hi[d] = hi[d] + 1;
    for (int a = 0; a < 3; a++)
        if (window[2 * a] < lo[a] || window[2 * a + 1] > hi[a] || window[2 * a] > window[2 * a + 1])
            return fail(FM_E_ARG, "fm_wall_effort_noalloc: window outside the zone");
    if (window[2 * d] != window[2 * d + 1])
        return fail(FM_E_ARG, "fm_wall_effort_noalloc: window is not one face plane");
    long nf = (long)(window[1] - window[0] + 1) * (window[3] - window[2] + 1) * (window[5] - window[4] + 1);
    if (nf > kScratchFaces)
        return fail(FM_E_ARG, "fm_wall_effort_noalloc: %ld faces, at most %ld", nf, kScratchFaces);
    if (c->device < 0 || c->device >= kMaxDevices || !g_scratch_ptr[c->device])
        return fail(FM_E_HIP, "fm_wall_effort_noalloc: no scratch on device %d", c->device);
    HIPCHK(hipSetDevice(c->device));
    double* buf = g_scratch_ptr[c->device];
    double* sums = buf + 33 * nf;
    EffArgs a;
    a.idir = idir;
    for (int q = 0; q < 6; q++) a.w[q] = window[q];
    for (int q = 0; q < 3; q++) a.pos[q] = pos ? pos[q] : 0.;
    a.rop = z->ro[level]; a.out = buf;
    if (launch_effort(k, a, c->stream)) return fail(FM_E_UNSUPPORTED, "no effort kernel for this scheme");
    launch_effort_sum(buf, nf, sums, c->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(effort, sums, sizeof(double) * 12, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && fields)
        e = hipMemcpyAsync(fields, buf, sizeof(double) * 21 * nf, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(FM_E_HIP, "fm_wall_effort_noalloc: %s", hipGetErrorString(e));
    return FM_OK;
}

extern "C" int fm_dpj_dpw(fm_ctx* c, int zone, int level, int idir, const int* window, double cosa,
                          double sina, double surfinv, double* dcd, double* dcl) {
    Zone* z = get_zone(c, zone);
    if (!z || level < 0 || level > 2 || idir < 1 || idir > 6 || !window || !dcd || !dcl)
        return fail(FM_E_ARG, "fm_dpj_dpw: bad arguments");
    const KZ& k = z->k;
    if (k.kflu != 1 || k.slope != 2 || k.iflow != 1)
        return fail(FM_E_UNSUPPORTED, "fm_dpj_dpw: AUSM+(P) o3 Euler zones only (the reference's option 213)");
    if (k.nf < 2 || k.nfk < 2) return fail(FM_E_UNSUPPORTED, "fm_dpj_dpw: the face stencil needs 2 ghost layers");
    int d = (idir - 1) / 2;
    int lo[3] = {1, 1, 1}, hi[3] = {k.ni, k.nj, k.nk};
    hi[d] = hi[d] + 1;
    for (int a = 0; a < 3; a++)
        if (window[2 * a] < lo[a] || window[2 * a + 1] > hi[a] || window[2 * a] > window[2 * a + 1])
            return fail(FM_E_ARG, "fm_dpj_dpw: window outside the zone");
    if (window[2 * d] != window[2 * d + 1]) return fail(FM_E_ARG, "fm_dpj_dpw: window is not one face plane");
    int face = idir % 2 == 1 ? 1 : (d == 0 ? k.ni : d == 1 ? k.nj : k.nk) + 1;
    if (window[2 * d] != face) return fail(FM_E_ARG, "fm_dpj_dpw: window is not on the zone face of idir %d", idir);
    HIPCHK(hipSetDevice(c->device));
    size_t bytes = sizeof(double) * 5 * (size_t)k.nd;
    double* buf = nullptr;
    HIPCHK(hipMalloc(&buf, 2 * bytes));
    hipError_t e = hipMemcpyAsync(buf, dcd, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(buf + 5 * k.nd, dcl, bytes, hipMemcpyHostToDevice, c->stream);
    DpjArgs a;
    a.idir = idir;
    for (int q = 0; q < 6; q++) a.w[q] = window[q];
    a.cosa = cosa; a.sina = sina; a.surfinv = surfinv;
    a.rop = z->ro[level]; a.dcd = buf; a.dcl = buf + 5 * k.nd;
    if (e == hipSuccess) { launch_dpj_dpw(k, a, c->stream); e = hipGetLastError(); }
    if (e == hipSuccess) e = hipMemcpyAsync(dcd, buf, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dcl, buf + 5 * k.nd, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(FM_E_HIP, "fm_dpj_dpw: %s", hipGetErrorString(e));
    return FM_OK;
}

// RHS arguments that make the residual kernels write the face balance alone into drodm: the
// explicit path with coefW = 0 (its stage state goes to the scratch level and is discarded)
static int balance_args(Zone* z, const double* rop, RhsArgs& a, int schedule) {
    memset(&a, 0, sizeof a);
    a.schedule = schedule;
    if (z->k.kflu == 5) {
        int rc = wall_corrections(z, a);
        if (rc) return rc;
    }
    a.rop = rop; a.rop_n = rop; a.rop_n1 = rop;
    a.implicit = 0; a.nitcfg = 1; a.zero_init = 1; a.coefH = 0.; a.coefW = 0.;
    a.rk_in = rop; a.rk_out = z->ro_scratch;
    return FM_OK;
}

extern "C" int fm_flux_balance(fm_ctx* c, int zone, int level, double* drodm) {
    Zone* z = get_zone(c, zone);
    if (!z || level < 0 || level > 2 || !drodm) return fail(FM_E_ARG, "fm_flux_balance: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    KZ& k = z->k;
    hipStream_t st = c->stream;
    const double* rop = z->ro[level];
    // viscosity (and the coe array the update reads) of this state, then the SENSEUR sensor
    Box b = {0, k.ni + 1, 0, k.nj + 1, 0, k.nk + 1};
    launch_coe(k, rop, 0, 0, b, k.iflow == 2, st);
    if (k.kflu == 2) launch_wiggle(k, rop, st);
    RhsArgs a;
    if (int rc = balance_args(z, rop, a, res_schedule_of(c, z))) return rc;
    a.kcol_n = c->res_kcol_n;
    z->dw_pending = 0;
    HIPCHK(hipMemsetAsync(k.drodm, 0, sizeof(double) * 5 * (size_t)k.nd, st));
    if (launch_residual(k, a, st)) return fail(FM_E_UNSUPPORTED, "no residual kernel for this scheme");
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(drodm, k.drodm, sizeof(double) * 5 * (size_t)k.nd, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return FM_OK;
}

extern "C" int fm_residual_vjp(fm_ctx* c, int zone, int level, const double* lam, double* out) {
    Zone* z = get_zone(c, zone);
    if (!z || level < 0 || level > 2 || !lam || !out) return fail(FM_E_ARG, "fm_residual_vjp: bad arguments");
    KZ& k = z->k;
    if (k.kflu != 1 || k.slope != 2 || k.iflow != 1)
        return fail(FM_E_UNSUPPORTED, "fm_residual_vjp: AUSM+(P) o3 Euler zones only");
    if (k.nf < 2 || k.nfk < 2) return fail(FM_E_UNSUPPORTED, "fm_residual_vjp: the face stencil needs 2 ghost layers");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    if (!z->adj_buf) {
        if (int rc = dmalloc(&z->adj_buf, vjp_buffer_doubles(k))) return rc;
    }
    size_t bytes = sizeof(double) * 5 * (size_t)k.nd;
    double* d = nullptr;
    HIPCHK(hipMalloc(&d, 2 * bytes));
    VjpArgs a;
    a.rop = z->ro[level]; a.lam = d; a.out = d + 5 * k.nd; a.buf = z->adj_buf;
    hipError_t e = hipMemcpyAsync(d, lam, bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(d + 5 * k.nd, 0, bytes, st);
    if (e == hipSuccess) { launch_residual_vjp(k, a, st); e = hipGetLastError(); }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d + 5 * k.nd, bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(FM_E_HIP, "fm_residual_vjp: %s", hipGetErrorString(e));
    return FM_OK;
}

extern "C" int fm_residual_vjp_zones(fm_ctx* c, int level, const double* const* lam, double* const* out) {
    if (!c || level < 0 || level > 2 || !lam || !out) return fail(FM_E_ARG, "fm_residual_vjp_zones: bad arguments");
    int nz = (int)c->zones.size();
    if (nz > kMaxTransferZones) return fail(FM_E_UNSUPPORTED, "fm_residual_vjp_zones: more than %d zones", kMaxTransferZones);
    size_t total = 0;
    for (int n = 0; n < nz; n++) {
        KZ& k = c->zones[n]->k;
        if (!lam[n] || !out[n]) return fail(FM_E_ARG, "fm_residual_vjp_zones: zone %d: no array", n);
        if (k.kflu != 1 || k.slope != 2 || k.iflow != 1)
            return fail(FM_E_UNSUPPORTED, "fm_residual_vjp_zones: zone %d: AUSM+(P) o3 Euler zones only", n);
        if (k.nf < 2 || k.nfk < 2)
            return fail(FM_E_UNSUPPORTED, "fm_residual_vjp_zones: zone %d: the face stencil needs 2 ghost layers", n);
        total += 5 * (size_t)k.nd;
    }
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    for (Zone* z : c->zones)
        if (!z->adj_buf)
            if (int rc = dmalloc(&z->adj_buf, vjp_buffer_doubles(z->k))) return rc;
    double* d = nullptr;
    HIPCHK(hipMalloc(&d, 2 * total * sizeof(double)));
    TransferArgs g;
    memset(&g, 0, sizeof g);
    hipError_t e = hipSuccess;
    size_t o = 0;
    for (int n = 0; n < nz && e == hipSuccess; n++) {
        Zone* z = c->zones[n];
        size_t cnt = 5 * (size_t)z->k.nd;
        VjpArgs a;
        double* dl = d + o;
        a.rop = z->ro[level]; a.lam = dl; a.out = d + total + o; a.buf = z->adj_buf;
        g.f[n] = a.out; g.nd[n] = z->k.nd;
        e = hipMemcpyAsync(dl, lam[n], cnt * sizeof(double), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemsetAsync(a.out, 0, cnt * sizeof(double), st);
        if (e == hipSuccess) { launch_residual_vjp(z->k, a, st); e = hipGetLastError(); }
        o += cnt;
    }
    // the transfers' transpose, raccords in reverse order
    for (int r = (int)c->tr_off.size() - 1; r >= 0 && e == hipSuccess; r--) {
        launch_transfer_vjp(c->tr_pts + c->tr_off[r], c->tr_cnt[r], g, c->tr_serial[r], st);
        e = hipGetLastError();
    }
    o = 0;
    for (int n = 0; n < nz && e == hipSuccess; n++) {
        size_t cnt = 5 * (size_t)c->zones[n]->k.nd;
        e = hipMemcpyAsync(out[n], d + total + o, cnt * sizeof(double), hipMemcpyDeviceToHost, st);
        o += cnt;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipStreamSynchronize(st);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(FM_E_HIP, "fm_residual_vjp_zones: %s", hipGetErrorString(e));
    return FM_OK;
}

extern "C" int fm_transfers_set(fm_ctx* c, int nrac, const int* no_r, const int* no_d,
                                const int64_t* nb_pts, const int64_t* rcv_pts, const int64_t* dnr_pts) {
    if (!c || nrac < 0 || (nrac > 0 && (!no_r || !no_d || !nb_pts || !rcv_pts || !dnr_pts)))
        return fail(FM_E_ARG, "fm_transfers_set: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    if ((long)c->zones.size() > kMaxTransferZones && nrac > 0)
        return fail(FM_E_UNSUPPORTED, "transfers between more than %d zones", kMaxTransferZones);
    std::vector<TransferPoint> h;
    std::vector<long> roff, rcnt;
    std::vector<char> rserial, rdup;
    long off = 0;
    for (int r = 0; r < nrac; r++) {
        Zone* zr = get_zone(c, no_r[r]);
        Zone* zd = get_zone(c, no_d[r]);
        if (!zr || !zd) return fail(FM_E_ARG, "raccord %d: bad zone %d <- %d", r, no_r[r], no_d[r]);
        if (nb_pts[r] < 0) return fail(FM_E_ARG, "raccord %d: %lld points", r, (long long)nb_pts[r]);
        roff.push_back(off); rcnt.push_back((long)nb_pts[r]);
        for (int64_t q = 0; q < nb_pts[r]; q++, off++) {
            int64_t ir = rcv_pts[off], id = dnr_pts[off];
            if (ir < 0 || ir >= zr->k.nd || id < 0 || id >= zd->k.nd)
                return fail(FM_E_ARG, "raccord %d point %lld: index outside the zone", r, (long long)q);
            h.push_back(TransferPoint{no_r[r], no_d[r], (long)ir, (long)id});
        }
        std::vector<long> rv(rcv_pts + roff.back(), rcv_pts + off), dv(dnr_pts + roff.back(), dnr_pts + off);
        std::sort(rv.begin(), rv.end()); std::sort(dv.begin(), dv.end());
        bool dup_r = std::adjacent_find(rv.begin(), rv.end()) != rv.end();
        rdup.push_back(dup_r ? 1 : 0);
        bool serial = dup_r || std::adjacent_find(dv.begin(), dv.end()) != dv.end();
        if (!serial && no_r[r] == no_d[r]) {
            std::vector<long> both;
            std::set_intersection(rv.begin(), rv.end(), dv.begin(), dv.end(), std::back_inserter(both));
            serial = !both.empty();
        }
        rserial.push_back(serial ? 1 : 0);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->tr_pts) { HIPCHK(hipFree(c->tr_pts)); c->tr_pts = nullptr; }
    c->tr_n = 0;
    if (!h.empty()) {
        HIPCHK(hipMalloc(&c->tr_pts, h.size() * sizeof(TransferPoint)));
        HIPCHK(hipMemcpy(c->tr_pts, h.data(), h.size() * sizeof(TransferPoint), hipMemcpyHostToDevice));
        c->tr_n = (long)h.size();
    }
    c->tr_off = roff; c->tr_cnt = rcnt; c->tr_serial = rserial; c->tr_rdup = rdup;
    std::vector<std::pair<int, long>> recv;
    recv.reserve(h.size());
    for (const TransferPoint& q : h) recv.emplace_back(q.zr, q.rcv);
    std::sort(recv.begin(), recv.end());
    c->tr_in_order = std::adjacent_find(recv.begin(), recv.end()) != recv.end();
    c->graph_epoch++;
    return FM_OK;
}

extern "C" int fm_transfer(fm_ctx* c, int level) {
    if (!c || level < 0 || level > 2) return fail(FM_E_ARG, "fm_transfer: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    std::vector<double*> f;
    for (Zone* z : c->zones) f.push_back(z->ro[level]);
    int rc = transfer(c, f.data());
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return FM_OK;
}

extern "C" int fm_apply_bc(fm_ctx* c, int zone, int level, int nstep) {
    (void)nstep;
    Zone* z = get_zone(c, zone);
    if (!z || level < 0 || level > 2) return fail(FM_E_ARG, "fm_apply_bc: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    int shift[6], sdm[6] = {1, z->k.ni, 1, z->k.nj, 1, z->k.nk};
    bczone(c, z, 0, z->ro[level], shift, nullptr);
    coins(c, z, z->ro[level], sdm);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return FM_OK;
}

// LU-SGS of several zones, diagonal-batched: one launch per block diagonal for all the zones of
// one mesh kind (3dcart / 3dfull), lower sweeps then upper sweeps, then the copy-back of the
// zones updated in place.  Same blocks and the same per-block work as the per-zone sweeps.
static int lu_batched(fm_ctx* c, const std::vector<Zone*>& zs, const std::vector<LuArgs>& las,
                      fm_ctx::GraphEntry* ge = nullptr, int nstep = 0, bool capture = false) {
    hipStream_t st = c->stream;
    // FP64 streaming: one launch per sweep over the pencils of all zones; FP64 dataflow: one launch
    // per sweep over the blocks of all zones, each zone with its own face exchange; otherwise (blocks
    // schedule, FP32 sweeps) one launch per block diagonal
    const bool strm = c->lu_schedule == LU_STREAM && !c->lu_fp32;
    const bool flow = c->lu_schedule == LU_DATAFLOW && !c->lu_fp32;
    if ((flow || strm) && !c->flow_ctr) {
        if (capture) return fail(FM_E_HIP, "graph capture: dataflow counter missing");
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipMalloc(&c->flow_ctr, 4 * sizeof(int)));
        HIPCHK(hipMemset(c->flow_ctr, 0, 4 * sizeof(int)));
    }
    // each zone's LuArgs as launched, for the record of what its rows hold
    std::vector<LuArgs> done(las);
    for (int cart = 0; cart < 2; cart++) {
        std::vector<Box> boxes;
        std::vector<LuZone> tab;
        int npack = 0, nunpack = 0;
        for (size_t n = 0; n < zs.size(); n++) {
            if ((zs[n]->k.typ == 2) != (cart == 1)) continue;
            LuArgs a = las[n];
            if (flow) {
                if (!zs[n]->lu_fx && capture) return fail(FM_E_HIP, "graph capture: face exchange missing");
                if (int rc = ensure_flow_exchange(c, zs[n])) return rc;
                a.fx = zs[n]->lu_fx; a.flow_step = 0; a.schedule = LU_DATAFLOW;
            }
            if (strm) {
                if ((!zs[n]->lu_skw || !zs[n]->lu_stx) && capture) return fail(FM_E_HIP, "graph capture: LU rows missing");
                if (int rc = ensure_stream(c, zs[n], a)) return rc;
                a.schedule = LU_STREAM;
                lu_stream_flags(c, zs[n], a);
                npack = std::max(npack, lu_pack_count(cart == 1, a));
                nunpack = std::max(nunpack, lu_unpack_end(a) - lu_unpack_first(a));
            }
            boxes.push_back(Box{a.i0, a.i1, a.j0, a.j1, a.k0, a.k1});
            tab.push_back(LuZone{zs[n]->k, a});
            done[n] = a;
        }
        if (tab.empty()) continue;
        fm_ctx::LuList* L = nullptr;
        for (auto& l : c->lu_lists) {
            if (l.cart != (cart == 1) || l.stream != strm || l.boxes.size() != boxes.size()) continue;
            bool same = true;
            for (size_t q = 0; q < boxes.size() && same; q++)
                same = !memcmp(&l.boxes[q], &boxes[q], sizeof(Box));
            if (same) { L = &l; break; }
        }
        if (!L) {
            fm_ctx::LuList nl;
            nl.cart = cart == 1; nl.stream = strm; nl.boxes = boxes; nl.work = nullptr; nl.maxrb = 0;
            std::vector<LuWork> w;
            if (strm) {
                lu_pencil_list(boxes, w, nl.maxrb);
                nl.offs.assign(1, 0);
                nl.offs.push_back((int)w.size());
            } else {
                lu_work_list(cart == 1, boxes, w, nl.offs);
            }
            if (!w.empty()) {
                HIPCHK(hipMalloc(&nl.work, w.size() * sizeof(LuWork)));
                HIPCHK(hipMemcpy(nl.work, w.data(), w.size() * sizeof(LuWork), hipMemcpyHostToDevice));
            }
            c->lu_lists.push_back(nl);
            L = &c->lu_lists.back();
        }
        if (!L->work) continue;
        if (ge) {
            // a captured step reads a table of its own, written once before its capture
            long key = 2L * nstep + cart;
            LuZone* d = nullptr;
            for (auto& t : ge->tabs) if (t.first == key) d = (LuZone*)t.second;
            if (!d) {
                if (capture) return fail(FM_E_HIP, "graph capture: LU table missing");
                HIPCHK(hipMalloc(&d, tab.size() * sizeof(LuZone)));
                ge->tabs.push_back({key, d});
            }
            if (!capture) HIPCHK(hipMemcpy(d, tab.data(), tab.size() * sizeof(LuZone), hipMemcpyHostToDevice));
            if (strm) {
                launch_lu_stream_list(cart == 1, d, (int)tab.size(), L->work, L->offs.back(), L->maxrb, npack, nunpack,
                                      c->flow_ctr, st);
            } else if (flow) {
                launch_lu_flow_list(cart == 1, d, L->work, L->offs.back(), c->flow_ctr, st);
            } else {
                launch_lu_list(cart == 1, true, c->lu_fp32, d, L->work, L->offs, st);
                launch_lu_list(cart == 1, false, c->lu_fp32, d, L->work, L->offs, st);
            }
            HIPCHK(hipGetLastError());
            continue;
        }
        if ((int)tab.size() > c->lu_tab_cap) {
            HIPCHK(hipStreamSynchronize(st));
            if (c->lu_tab) HIPCHK(hipFree(c->lu_tab));
            for (int q = 0; q < 2; q++) {
                if (c->lu_host[q]) HIPCHK(hipHostFree(c->lu_host[q]));
                HIPCHK(hipHostMalloc(&c->lu_host[q], tab.size() * sizeof(LuZone)));
                if (!c->lu_ev[q]) HIPCHK(hipEventCreateWithFlags(&c->lu_ev[q], hipEventDisableTiming));
            }
            HIPCHK(hipMalloc(&c->lu_tab, tab.size() * sizeof(LuZone)));
            c->lu_tab_cap = (int)tab.size();
        }
        // the table goes through a pinned buffer that the previous copy from it has released
        int q = c->lu_slot; c->lu_slot ^= 1;
        HIPCHK(hipEventSynchronize(c->lu_ev[q]));
        memcpy(c->lu_host[q], tab.data(), tab.size() * sizeof(LuZone));
        HIPCHK(hipMemcpyAsync(c->lu_tab, c->lu_host[q], tab.size() * sizeof(LuZone), hipMemcpyHostToDevice, st));
        HIPCHK(hipEventRecord(c->lu_ev[q], st));
        if (strm) {
            launch_lu_stream_list(cart == 1, c->lu_tab, (int)tab.size(), L->work, L->offs.back(), L->maxrb, npack,
                                  nunpack, c->flow_ctr, st);
        } else if (flow) {
            launch_lu_flow_list(cart == 1, c->lu_tab, L->work, L->offs.back(), c->flow_ctr, st);
        } else {
            launch_lu_list(cart == 1, true, c->lu_fp32, c->lu_tab, L->work, L->offs, st);
            launch_lu_list(cart == 1, false, c->lu_fp32, c->lu_tab, L->work, L->offs, st);
        }
        HIPCHK(hipGetLastError());
    }
    for (size_t n = 0; n < zs.size(); n++) {
        const LuArgs& la = las[n];
        lu_stream_after(zs[n], done[n]);
        if (la.out == zs[n]->ro_scratch) {
            Box lb = {la.i0, la.i1, la.j0, la.j1, la.k0, la.k1};
            launch_copy_box(zs[n]->k, zs[n]->ro_scratch, zs[n]->ro[2], lb, st);
        }
    }
    HIPCHK(hipGetLastError());
    return FM_OK;
}

// Several zones: the per-zone work of a phase (right-hand side, boundary conditions) runs on up
// to kSideStreams streams forked from the solver stream and joined back to it, so that the
// kernels of small zones run side by side.  Graph capture records the fork and join as branches.
constexpr size_t kSideStreams = 8;

static int side_fork(fm_ctx* c, size_t ns, bool capture) {
    if (c->side.size() < ns || !c->fork_ev) {
        if (capture) return fail(FM_E_HIP, "graph capture: zone streams missing");
        while (c->side.size() < ns) {
            hipStream_t s2;
            hipEvent_t e2;
            HIPCHK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&e2, hipEventDisableTiming));
            c->side.push_back(s2); c->side_ev.push_back(e2);
        }
        if (!c->fork_ev) HIPCHK(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
    }
    HIPCHK(hipEventRecord(c->fork_ev, c->stream));
    for (size_t s = 0; s < ns; s++) HIPCHK(hipStreamWaitEvent(c->side[s], c->fork_ev, 0));
    return FM_OK;
}

static int side_join(fm_ctx* c, size_t ns) {
    c->cur = c->stream;
    for (size_t s = 0; s < ns; s++) {
        HIPCHK(hipEventRecord(c->side_ev[s], c->side[s]));
        HIPCHK(hipStreamWaitEvent(c->stream, c->side_ev[s], 0));
    }
    return FM_OK;
}

// The edges and corners of several zones in one launch, after their physical boundary conditions
// (per zone, gsdr3.cpp:436-439): a zone's edges and corners read only its own face ghosts.  A
// captured step reads a table of its own (written once before its capture); otherwise the table
// goes through a pinned buffer that the previous copy from it has released.
static int coins_batched(fm_ctx* c, const std::vector<CoinsZone>& co, int nblk, fm_ctx::GraphEntry* ge,
                         int nstep, bool capture) {
    if (co.empty()) return FM_OK;
    hipStream_t st = c->stream;
    const size_t bytes = co.size() * sizeof(CoinsZone);
    CoinsZone* d = nullptr;
    if (ge) {
        const long key = -1L - nstep;
        for (auto& t : ge->tabs) if (t.first == key) d = (CoinsZone*)t.second;
        if (!d) {
            if (capture) return fail(FM_E_HIP, "graph capture: corner table missing");
            HIPCHK(hipMalloc(&d, bytes));
            ge->tabs.push_back({key, d});
        }
        if (!capture) HIPCHK(hipMemcpy(d, co.data(), bytes, hipMemcpyHostToDevice));
    } else {
        if ((int)co.size() > c->co_cap) {
            HIPCHK(hipStreamSynchronize(st));
            if (c->co_tab) HIPCHK(hipFree(c->co_tab));
            for (int q = 0; q < 2; q++) {
                if (c->co_host[q]) HIPCHK(hipHostFree(c->co_host[q]));
                HIPCHK(hipHostMalloc(&c->co_host[q], bytes));
                if (!c->co_ev[q]) HIPCHK(hipEventCreateWithFlags(&c->co_ev[q], hipEventDisableTiming));
            }
            HIPCHK(hipMalloc(&c->co_tab, bytes));
            c->co_cap = (int)co.size();
        }
        int q = c->co_slot; c->co_slot ^= 1;
        HIPCHK(hipEventSynchronize(c->co_ev[q]));
        memcpy(c->co_host[q], co.data(), bytes);
        HIPCHK(hipMemcpyAsync(c->co_tab, c->co_host[q], bytes, hipMemcpyHostToDevice, st));
        HIPCHK(hipEventRecord(c->co_ev[q], st));
        d = c->co_tab;
    }
    launch_coins_list(d, (int)co.size(), nblk, st);
    HIPCHK(hipGetLastError());
    return FM_OK;
}

// Within a sub-iteration every zone first gets its right-hand side and LU-SGS update, then
// the transfers fill the ghost cells shared with other zones, then every zone's physical
// boundary conditions and corners are applied (gsdr3.cpp:205-444).
// One time step of every zone: sub-iterations n0..n1 (gsdr3 per sub-iteration).
static int time_step(fm_ctx* c, int first_it, int verif, int n0, int n1, fm_ctx::GraphEntry* ge, bool capture) {
    const size_t nz = c->zones.size();
    // several zones: their LU-SGS sweeps share one launch per block diagonal
    // several zones: the batched blocks schedule (one launch per diagonal for all zones) also
    // stands in for the dataflow schedule, whose per-zone launches and exchange fills cost more
    // than they save on many small zones
    const bool batch = nz > 1 && c->lu_batch &&
                       (c->lu_schedule == LU_BLOCKS || c->lu_schedule == LU_DATAFLOW || c->lu_schedule == LU_STREAM);
    std::vector<int> lexit(nz);
    std::vector<double*> cl(nz);
    const size_t ns = nz > 1 ? std::min(nz, kSideStreams) : 0;
    for (int nstep = n0; nstep <= n1; nstep++) {
        int any = 0;
        std::vector<Zone*> bz;
        std::vector<LuArgs> bl;
        if (ns) { if (int rc = side_fork(c, ns, capture)) return rc; }
        for (size_t n = 0; n < nz; n++) {
            Zone* z = c->zones[n];
            if (ns) c->cur = c->side[n % ns];
            int implicit = z->k.itypcp <= 1;
            int nssiter = z->k.nssiter;
            // souszones_list.cpp:44-84: a verification step computes residuals (and the
            // CFL monitor) at every sub-iteration and skips LU-SGS at the last one
            int lssiter_verif = verif, lexit_lu = 0;
            lexit[n] = -1;
            cl[n] = level_cl(z, nstep);
            if (implicit && nstep == nssiter) {
                if (!verif) continue;           // skip (computePT.cpp:581-582)
                lexit_lu = 1;
            }
            LuArgs la;
            la.i0 = 1; la.i1 = 0;
            int rc = substep(c, z, nstep, first_it, lssiter_verif, lexit_lu, 3, batch ? &la : nullptr);
            if (rc) { c->cur = c->stream; return rc; }
            if (batch && implicit && lexit_lu == 0 && la.i1 >= la.i0) { bz.push_back(z); bl.push_back(la); }
            lexit[n] = lexit_lu;
            any |= lexit_lu == 0;
        }
        if (ns) { if (int rc = side_join(c, ns)) return rc; }
        if (!bz.empty()) {
            int rc = lu_batched(c, bz, bl, ge, nstep, capture);
            if (rc) return rc;
        }
        if (!any) continue;
        int rc = transfer(c, cl.data());
        if (rc) return rc;
        if (ns) {
            // several zones: their physical boundary conditions side by side, then the edges and
            // corners of all of them in one launch
            if (int rc2 = side_fork(c, ns, capture)) return rc2;
            std::vector<CoinsZone> co;
            int nblk = 0;
            for (size_t n = 0; n < nz; n++) {
                if (lexit[n] != 0) continue;
                Zone* z = c->zones[n];
                c->cur = c->side[n % ns];
                int shift[6];
                bczone(c, z, 0, cl[n], shift, nullptr);
                const CoinsZone e = {z->k, cl[n], Box{1, z->k.ni, 1, z->k.nj, 1, z->k.nk}};
                co.push_back(e);
                nblk = std::max(nblk, coins_blocks(e.z, e.L));
            }
            if (int rc2 = side_join(c, ns)) return rc2;
            if (int rc2 = coins_batched(c, co, nblk, ge, nstep, capture)) return rc2;
        } else {
            for (size_t n = 0; n < nz; n++)
                if (lexit[n] == 0) fill_ghosts(c, c->zones[n], cl[n]);
        }
        HIPCHK(hipGetLastError());
    }
    return FM_OK;
}

// computePT.cpp:534-679 for the implemented options: nit_c steps; in each step the
// sub-iterations nstep_deb..nstep_fin of every zone; the last implicit sub-iteration only
// computes the residual (lexit_lu) and only on verification steps; time levels rotate.
// Within a sub-iteration every zone first gets its right-hand side and LU-SGS update, then
// the transfers fill the ghost cells shared with other zones, then every zone's physical
// boundary conditions and corners are applied (gsdr3.cpp:205-444).
//
// With FM_OPT_GRAPH a step is run eagerly the first time its configuration (level pointers,
// first_it, verification, sub-iteration range) appears, captured into a HIP graph the second
// time, and replayed from then on.  The levels rotate with period 3, so a run settles on a
// few graphs.
static int compute_steps(fm_ctx* c, int nitrun, int nstep_deb, int nstep_fin, int nit_c, int modulo_verif,
                         int first_it) {
    for (int it = 0; it < nit_c; it++) {
        int nitrun_loc = nitrun * nit_c + it;
        int verif = modulo_verif > 0 && (nitrun_loc % modulo_verif == 0);
        if (!c->graph_on) {
            int rc = time_step(c, first_it, verif, nstep_deb, nstep_fin, nullptr, false);
            if (rc) return rc;
        } else {
            std::vector<const void*> ptrs;
            std::vector<const double*> wr;
            for (Zone* z : c->zones) {
                for (int q = 0; q < 3; q++) ptrs.push_back(z->ro[q]);
                wr.push_back(z->w_rows);
            }
            fm_ctx::GraphEntry* ge = nullptr;
            for (auto* g : c->graphs)
                if (g->ptrs == ptrs && g->w_rows == wr && g->first_it == first_it && g->verif == verif &&
                    g->n0 == nstep_deb && g->n1 == nstep_fin && g->epoch == c->graph_epoch) { ge = g; break; }
            if (!ge) {
                ge = new fm_ctx::GraphEntry();
                ge->ptrs = ptrs; ge->w_rows = wr; ge->first_it = first_it; ge->verif = verif; ge->n0 = nstep_deb;
                ge->n1 = nstep_fin; ge->epoch = c->graph_epoch; ge->state = 0;
                c->graphs.push_back(ge);
            }
            if (ge->state == 0) {
                int rc = time_step(c, first_it, verif, nstep_deb, nstep_fin, ge, false);
                if (rc) return rc;
                ge->state = 1;
                ge->dw_pending.clear();
                ge->w_rows_after.clear();
                for (Zone* z : c->zones) { ge->dw_pending.push_back(z->dw_pending); ge->w_rows_after.push_back(z->w_rows); }
            } else {
                if (ge->state == 1) {
                    HIPCHK(hipStreamSynchronize(c->stream));
                    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                    int rc = time_step(c, first_it, verif, nstep_deb, nstep_fin, ge, true);
                    hipGraph_t g;
                    hipError_t e = hipStreamEndCapture(c->stream, &g);
                    if (rc) return rc;
                    if (e != hipSuccess) return fail(FM_E_HIP, "graph capture: %s", hipGetErrorString(e));
                    e = hipGraphInstantiate(&ge->exec, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    if (e != hipSuccess) return fail(FM_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
                    ge->state = 2;
                }
                HIPCHK(hipGraphLaunch(ge->exec, c->stream));
                // the host state a replayed step leaves is the one the eager step left
                for (size_t n = 0; n < c->zones.size() && n < ge->dw_pending.size(); n++) {
                    c->zones[n]->dw_pending = ge->dw_pending[n];
                    c->zones[n]->w_rows = ge->w_rows_after[n];
                }
            }
        }
        first_it = 1;
        for (Zone* z : c->zones) {
            // a verified step records the residual of every sub-iteration it ran
            if (verif && z->k.itypcp <= 1 && nstep_fin >= nstep_deb) {
                z->rec_lo = nstep_deb; z->rec_hi = nstep_fin; z->have_norms = 1;
            }
            double* m1 = z->ro[1]; z->ro[1] = z->ro[0]; z->ro[0] = z->ro[2]; z->ro[2] = m1;
        }
    }
    return FM_OK;
}

// dW of the last streaming LU-SGS inside fm_compute is still in its skewed rows: copy it to drodm
// before anything reads drodm (the next residual overwrites it instead)
static void resolve_dw(fm_ctx* c, Zone* z) {
    if (!z->dw_pending) return;
    launch_lu_unpack_dw(z->k, z->dw_la, c->stream);
    z->dw_pending = 0;
}

static int compute(fm_ctx* c, int nitrun, int nstep_deb, int nstep_fin, int nit_c, int modulo_verif, int first_it) {
    c->defer_dw = 1;
    int rc = compute_steps(c, nitrun, nstep_deb, nstep_fin, nit_c, modulo_verif, first_it);
    c->defer_dw = 0;
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    return FM_OK;
}

extern "C" int fm_compute(fm_ctx* c, int nitrun, int nstep_deb, int nstep_fin, int nit_c,
                          int modulo_verif, int first_it) {
    if (!c) return fail(FM_E_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    return compute(c, nitrun, nstep_deb, nstep_fin, nit_c, modulo_verif, first_it);
}

extern "C" int fm_compute_timed(fm_ctx* c, int nitrun, int nstep_deb, int nstep_fin, int nit_c,
                                int modulo_verif, int first_it, float* ms) {
    if (!c) return fail(FM_E_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipEventRecord(c->e0, c->stream));
    int rc = compute(c, nitrun, nstep_deb, nstep_fin, nit_c, modulo_verif, first_it);
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->e1, c->stream));
    HIPCHK(hipEventSynchronize(c->e1));
    if (ms) HIPCHK(hipEventElapsedTime(ms, c->e0, c->e1));
    return check_lu(c);
}

extern "C" int fm_sync(fm_ctx* c) {
    if (!c) return fail(FM_E_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return check_lu(c);
}

extern "C" int fm_cfl(fm_ctx* c, int zone, double* out) {
    Zone* z = get_zone(c, zone);
    if (!z || !out) return fail(FM_E_ARG, "fm_cfl: bad arguments");
    if (!z->have_cfl) return fail(FM_E_ARG, "no CFL computed yet (first sub-iteration of a verification step)");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    double h[3];
    HIPCHK(hipMemcpy(h, z->norm_out + 10, sizeof h, hipMemcpyDeviceToHost));
    // computePT.cpp:629-645: max, min and the mean over the interior cells
    out[0] = h[0]; out[1] = h[1];
    out[2] = h[2] / ((double)z->k.ni * z->k.nj * z->k.nk);
    return FM_OK;
}

extern "C" int fm_ctx_set_option(fm_ctx* c, int option, int value) {
    if (!c) return fail(FM_E_ARG, "null context");
    if (option == FM_OPT_LU_SCHEDULE) {
        if (value != FM_LU_BLOCKS && value != FM_LU_PIPELINE && value != FM_LU_DATAFLOW && value != FM_LU_STREAM)
            return fail(FM_E_ARG, "LU schedule %d", value);
        if (value == FM_LU_PIPELINE && c->lu_fp32)
            return fail(FM_E_UNSUPPORTED, "the pipelined LU-SGS schedule has no FP32 sweeps");
        c->lu_schedule = value == FM_LU_PIPELINE ? LU_PIPELINE : value == FM_LU_DATAFLOW ? LU_DATAFLOW
                       : value == FM_LU_STREAM ? LU_STREAM : LU_BLOCKS;
        c->graph_epoch++;
        return FM_OK;
    }
    if (option == FM_OPT_RESIDUAL_SCHEDULE) {
        if (value != FM_RES_CELLS && value != FM_RES_FACES && value != FM_RES_KCOL && value != FM_RES_AUTO)
            return fail(FM_E_ARG, "residual schedule %d", value);
        c->res_schedule = value == FM_RES_FACES ? RES_FACES : value == FM_RES_KCOL ? RES_KCOL
                        : value == FM_RES_AUTO ? RES_AUTO : RES_CELLS;
        c->graph_epoch++;
        return FM_OK;
    }
    if (option == FM_OPT_ZONE_BATCH) {
        if (value != 0 && value != 1) return fail(FM_E_ARG, "zone batch %d", value);
        c->lu_batch = value;
        c->graph_epoch++;
        return FM_OK;
    }
    if (option == FM_OPT_LU_PRECISION) {
        if (value != FM_LU_FP64 && value != FM_LU_FP32) return fail(FM_E_ARG, "LU precision %d", value);
        if (value == FM_LU_FP32 && c->lu_schedule == LU_PIPELINE)
            return fail(FM_E_UNSUPPORTED, "the pipelined LU-SGS schedule has no FP32 sweeps");
        c->lu_fp32 = value == FM_LU_FP32;
        c->graph_epoch++;
        return FM_OK;
    }
    if (option == FM_OPT_LU_GIVE_UP) {
        // failure-handling test hook: set every zone's LU error word as a wait that gave up would
        if (value != 1) return fail(FM_E_ARG, "LU give-up %d", value);
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->stream));
        int one = 1;
        for (Zone* z : c->zones)
            if (z->lu_sync) HIPCHK(hipMemcpy(z->lu_sync + 1, &one, sizeof(int), hipMemcpyHostToDevice));
        return FM_OK;
    }
    if (option == FM_OPT_RES_KCOL_N) {
        if (value < 0 || value > 64) return fail(FM_E_ARG, "k-column length %d (0..64)", value);
        c->res_kcol_n = value;
        c->graph_epoch++;
        return FM_OK;
    }
    if (option == FM_OPT_LU_FLOW_GRID) {
        if (value < 0) return fail(FM_E_ARG, "dataflow grid %d", value);
        c->lu_flow_grid = value;
        c->graph_epoch++;
        return FM_OK;
    }
    if (option == FM_OPT_GRAPH) {
        if (value != 0 && value != 1) return fail(FM_E_ARG, "graph %d", value);
        c->graph_on = value;
        c->graph_epoch++;
        return FM_OK;
    }
    return fail(FM_E_ARG, "unknown option %d", option);
}

extern "C" int fm_residual_norms(fm_ctx* c, int zone, double* out) {
    Zone* z = get_zone(c, zone);
    if (!z || !out) return fail(FM_E_ARG, "fm_residual_norms: bad arguments");
    if (!z->have_norms) return fail(FM_E_ARG, "no residual computed yet");
    double h[10];
    HIPCHK(hipStreamSynchronize(c->stream));
    if (int rc = check_lu(c)) return rc;
    HIPCHK(hipMemcpy(h, z->norm_hist + 10 * (z->rec_hi - 1), sizeof h, hipMemcpyDeviceToHost));
    double ncell = (double)z->k.ni * z->k.nj * z->k.nk;
    double dtc = z->k.dtc;
    for (int e = 0; e < 5; e++) {
        out[e] = sqrt(h[e] / ncell) / dtc;
        out[5 + e] = h[5 + e] / dtc;
    }
    return FM_OK;
}

// display_ss_iteration.cpp:96-137: L2 = sqrt(sum / cells) / DTC and Linf = max / DTC of every
// sub-iteration the last verified step recorded, in sub-iteration order.
extern "C" int fm_residual_history(fm_ctx* c, int zone, double* out, int nmax, int* nrec) {
    Zone* z = get_zone(c, zone);
    if (!z || !out || !nrec || nmax < 0) return fail(FM_E_ARG, "fm_residual_history: bad arguments");
    if (!z->have_norms) return fail(FM_E_ARG, "no residual computed yet");
    const int n = z->rec_hi - z->rec_lo + 1;
    *nrec = n;
    const int m = n < nmax ? n : nmax;
    if (m == 0) return FM_OK;
    std::vector<double> h((size_t)10 * m);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (int rc = check_lu(c)) return rc;
    HIPCHK(hipMemcpy(h.data(), z->norm_hist + 10 * (z->rec_lo - 1), sizeof(double) * 10 * m, hipMemcpyDeviceToHost));
    const double ncell = (double)z->k.ni * z->k.nj * z->k.nk, dtc = z->k.dtc;
    for (int r = 0; r < m; r++)
        for (int e = 0; e < 5; e++) {
            out[10 * r + e] = sqrt(h[10 * r + e] / ncell) / dtc;
            out[10 * r + 5 + e] = h[10 * r + 5 + e] / dtc;
        }
    return FM_OK;
}

extern "C" int fm_enstrophy_tke(fm_ctx* c, int zone, int level, double* out) {
    Zone* z = get_zone(c, zone);
    if (!z || !out || level < 0 || level > 2) return fail(FM_E_ARG, "fm_enstrophy_tke: bad arguments");
    int nblk = z->norm_blocks;
    launch_tke(z->k, z->ro[level], z->norm_part, nblk, c->stream);
    HIPCHK(hipGetLastError());
    std::vector<double> h(3 * nblk);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(h.data(), z->norm_part, sizeof(double) * 3 * nblk, hipMemcpyDeviceToHost));
    double sk = 0., se = 0., sv = 0.;
    for (int b = 0; b < nblk; b++) { sk += h[3 * b]; se += h[3 * b + 1]; sv += h[3 * b + 2]; }
    out[0] = sk; out[1] = sv; out[2] = se;
    return FM_OK;
}

// Phase timing (benchmark support): the kernels of one phase of sub-iteration 1, repeated.
extern "C" int fm_time_phase(fm_ctx* c, int zone, int phase, int reps, float* ms) {
    Zone* z = get_zone(c, zone);
    if (!z || reps < 1 || !ms) return fail(FM_E_ARG, "fm_time_phase: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    resolve_dw(c, z);
    KZ& k = z->k;
    hipStream_t st = c->stream;
    int implicit = k.itypcp <= 1;
    int lamin = k.iflow == 2;
    double* ro = z->ro[0];
    RhsArgs a;
    memset(&a, 0, sizeof a);
    a.schedule = res_schedule_of(c, z);
    a.kcol_n = c->res_kcol_n;
    if (k.kflu == 5) {
        int rc = wall_corrections(z, a);
        if (rc) return rc;
    }
    a.rop = ro; a.rop_n = ro; a.rop_n1 = z->ro[1]; a.implicit = implicit; a.nitcfg = 1;
    a.zero_init = 1; a.coefW = 0.5; a.rk_in = ro; a.rk_out = z->ro_scratch;
    a.c1 = -1.5; a.c2 = 2.; a.c3 = -0.5;
    LuArgs la;
    la.i0 = 1; la.i1 = k.ni; la.j0 = 1; la.j1 = k.nj; la.k0 = 1; la.k1 = k.nk;
    la.dw = k.drodm; la.rop = ro; la.rop_k = ro; la.out = z->ro_scratch; la.newton = 1;
    la.dws = z->lu_dws; la.sync = z->lu_sync; la.schedule = c->lu_schedule; la.fp32 = c->lu_fp32;
    la.flow_grid = c->lu_flow_grid;
    if (int rc = ensure_exchange(c, z)) return rc;
    la.xch = z->lu_xch;
    la.fx = nullptr; la.flow_work = nullptr; la.flow_n = 0; la.flow_step = 0;
    la.keep_dw = 1;
    la.w_ready = 0;
    la.n_ready = 0;
    z->w_rows = nullptr;                      // its sweeps write the scratch state into the rows
    z->n_known = 0;                           // and copy the normals of its own box
    c->graph_epoch++;                         // (captured steps hold whether they were in place)
    {
        Box lb = {la.i0, la.i1, la.j0, la.j1, la.k0, la.k1};
        if (int rc = ensure_dataflow(c, z, lb, la)) return rc;
        if (int rc = ensure_stream(c, z, la)) return rc;
    }
    Box b = {1 - k.nf, k.ni + k.nf, 1 - k.nf, k.nj + k.nf, 1 - k.nfk, k.nk + k.nfk};
    if (phase == FM_PHASE_VJP && (k.nf < 2 || k.nfk < 2))
        return fail(FM_E_UNSUPPORTED, "transposed product: the face stencil needs 2 ghost layers");
    if (phase == FM_PHASE_VJP && !z->adj_buf) {
        if (int rc = dmalloc(&z->adj_buf, vjp_buffer_doubles(k))) return rc;
    }
    HIPCHK(hipStreamSynchronize(st));
    // one untimed run first (first-launch costs stay out of the average)
    for (int r = -1; r < reps; r++) {
        if (r == 0) HIPCHK(hipEventRecord(c->e0, st));
        if (phase == FM_PHASE_COE) launch_coe(k, ro, 1, implicit ? (k.dtloc ? 2 : 1) : 0, b, lamin, st);
        else if (phase == FM_PHASE_RESIDUAL) {
            if (launch_residual(k, a, st)) return fail(FM_E_UNSUPPORTED, "no residual kernel");
        } else if (phase == FM_PHASE_LU) {
            if (!implicit) return fail(FM_E_ARG, "LU phase needs an implicit zone");
            launch_lu(k, la, st);
        } else if (phase == FM_PHASE_BC) {
            int shift[6], sdm[6] = {1, k.ni, 1, k.nj, 1, k.nk};
            bczone(c, z, 0, z->ro_scratch, shift, nullptr);
            coins(c, z, z->ro_scratch, sdm);
        } else if (phase == FM_PHASE_VJP) {
            // the transposed product with lam = the current drodm, into the scratch level
            VjpArgs va;
            va.rop = ro; va.lam = k.drodm; va.out = z->ro_scratch; va.buf = z->adj_buf;
            if (launch_residual_vjp(k, va, st)) return fail(FM_E_UNSUPPORTED, "no transposed product for this scheme");
        } else return fail(FM_E_ARG, "unknown phase %d", phase);
    }
    HIPCHK(hipEventRecord(c->e1, st));
    HIPCHK(hipEventSynchronize(c->e1));
    HIPCHK(hipGetLastError());
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, c->e0, c->e1));
    *ms = t / reps;
    return check_lu(c);
}
