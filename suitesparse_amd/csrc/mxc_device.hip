// mxc_device.hip -- device plan, launches and the C ABI of the MI355X
// supernodal Cholesky numeric path (see include/cholmod_mi355x.h).
//
// Schedule.  The supernodal elimination tree is cut into levels by height
// (leaves = 0).  For level l, in order:
//   1. every (descendant d, ancestor s) update with s in level l, as 64x64
//      MFMA tiles fused with the relative-map scatter-subtract into s
//      (one launch for the whole level);
//   2. for t = 0, 1, ...: POTRF + inverse of diagonal block t of every
//      supernode of the level that has one, TRSM of the rows below it through
//      the inverse (MFMA), and the in-supernode trailing update (MFMA).
// A descendant's update lands after all of its own updates are done because
// d sits in a lower level than s.  The reference instead walks supernodes one
// at a time (CHOLMOD/Supernodal/t_cholmod_super_numeric_worker.c:335) and, on
// its GPU path, ships each descendant over PCIe (CHOLMOD/GPU/t_cholmod_gpu.c).
//
// Summation order differs from the reference (atomic scatter-adds of
// independent descendants), so values agree to rounding, not bitwise.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <mutex>
#include <climits>
#include <cmath>
#include <map>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cholmod_mi355x.h"
#include "kernels_factor.h"
#include "kernels_solve.h"
#include "kernels_modify.h"
#include "symbolic_internal.h"

using namespace mxc;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";

static int set_hip_err(hipError_t e, const char *what, int line) {
    snprintf(g_err, sizeof(g_err), "%s failed at mxc_device.hip:%d: %s", what, line,
             hipGetErrorString(e));
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? MXC_OUT_OF_MEMORY
                                                                       : MXC_GPU_PROBLEM;
}
#define HIPTRY(x)                                              \
    do {                                                       \
        hipError_t e_ = (x);                                   \
        if (e_ != hipSuccess) return set_hip_err(e_, #x, __LINE__); \
    } while (0)

extern "C" const char *mxc_last_error(void) { return g_err; }

static int g_device = 0;

// HIP runtime calls that create or destroy runtime objects are serialized
// across host threads: graph capture and instantiation, graph launches, stream
// and event creation, device allocation and release (plan upload, work-buffer
// growth, factor release).  Each is short or rare (once per plan, per
// right-hand-side count or per growth), so no thread captures, launches a
// graph or frees memory while another does.  The host part of a plan build
// (pattern analysis, block lists, launch lists) runs outside the lock, so slot
// threads of a server build plans concurrently.  Recursive: releasing a
// factor also releases its FP64 fallback.
static std::recursive_mutex g_rt_mutex;

// Scratch of mxc_rcond_noalloc: the per-block (min, max, has_nan) partials of
// k_diag_minmax, which always runs RCOND_BLOCKS blocks (a grid-stride loop over
// the n diagonal entries, so the size does not depend on the factor).  Each GPU
// holds its own copy, from the first mxc_device_init on it until the process
// ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for that device (g_scratch).
static constexpr int RCOND_BLOCKS = 256;
static constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[3 * RCOND_BLOCKS];
static double *g_scratch[MAX_DEVICES];
static std::mutex g_scratch_mutex;           // calls sharing one GPU's array take turns

extern "C" int mxc_device_init(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        snprintf(g_err, sizeof(g_err), "no HIP device visible");
        return MXC_GPU_PROBLEM;
    }
    if (device < 0 || device >= count) return MXC_INVALID;
    HIPTRY(hipSetDevice(device));
    g_device = device;
    if (device < MAX_DEVICES) {
        std::lock_guard<std::mutex> lk(g_scratch_mutex);
        if (!g_scratch[device]) {
            void *p = nullptr;
            HIPTRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            g_scratch[device] = (double *)p;
        }
    }
    return MXC_OK;
}

extern "C" int mxc_device_info(int64_t *info) {
    hipDeviceProp_t p;
    HIPTRY(hipGetDeviceProperties(&p, g_device));
    info[0] = p.multiProcessorCount;
    info[1] = (int64_t)p.totalGlobalMem;
    info[2] = 950;   // built for gfx950 only
    info[3] = p.clockRate;
    return MXC_OK;
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
template <typename X>
struct DBuf {
    X *p = nullptr;
    size_t n = 0;
};

struct Launch {
    int kind;          // see enum below
    long long begin;   // first task
    long long count;   // number of tasks (grid size)
    int tile;          // k_tile: tile edge; sweep steps: number of diag workgroups
    int nfirst = 0;    // trailing launches with folded TRSMs: producer tiles at the head
};

// Tile-size buckets over (rows M, columns N, inner dimension K) of a block:
// 5 x 5 x 3 classes, one table per kind of block (descendant update, trailing
// update, TRSM).
static inline int tile_bucket(int M, int N, int K) {
    int mb = M <= 16 ? 0 : M <= 32 ? 1 : M <= 64 ? 2 : M <= 128 ? 3 : 4;
    int nb = N <= 8 ? 0 : N <= 16 ? 1 : N <= 32 ? 2 : N <= 64 ? 3 : 4;
    int kb = K <= 32 ? 0 : K <= 128 ? 1 : 2;
    return (mb * 5 + nb) * 3 + kb;
}
static constexpr int NBUCKET = 75;
static constexpr int NKIND = 3;     // 0 descendant update, 1 trailing, 2 TRSM
static inline int kind_of(int launch_kind) {
    return launch_kind == 0 ? 0 : (launch_kind == 3 || launch_kind == 9 || launch_kind == 14) ? 1 : 2;
}
// L_TRAIL2: the part of a panel's trailing update right of the next panel, on
// the second stream; L_FORK / L_JOIN: the event edges between the two streams;
// L_FCHAIN / L_BCHAIN: a level's chained sweep launch; L_TRAILD: a trailing
// launch whose (0, 0) tiles also factor the next step's diagonal blocks;
// L_FTINY / L_BTINY: a level's tiny supernodes, one lane each; L_FQUAD /
// L_BQUAD: its quarter-wave supernodes, sixteen lanes each; L_DIAG_TINY:
// diagonal blocks of at most TINY_DIAG columns, one lane each; L_TRSM_TINY:
// the panel TRSM of tiny supernodes, one lane each; L_SMALL: a level's small
// supernodes (FP64), each factored whole by one wave (k_small_level).
enum { L_UPD = 0, L_DIAG = 1, L_TRSM = 2, L_TRAIL = 3, L_FSTEP = 4, L_BSTEP = 5, L_FSMALL = 6, L_BSMALL = 7,
       L_DIAG_SMALL = 8, L_TRAIL2 = 9, L_FORK = 10, L_JOIN = 11, L_FCHAIN = 12, L_BCHAIN = 13,
       L_TRAILD = 14, L_FTINY = 15, L_BTINY = 16, L_DIAG_TINY = 17, L_TRSM_TINY = 18, L_FQUAD = 19,
       L_BQUAD = 20, L_SMALL = 21 };

// ---------------------------------------------------------------------------
// device memory pool: one hipMalloc, sub-allocated first-fit with coalescing
// (the Common-owned dev_mempool of CHOLMOD/GPU/cholmod_gpu.c:426-543, here
// holding whole factors instead of per-supernode buffers)
// ---------------------------------------------------------------------------
struct mxc_pool {
    int device = 0;
    char *base = nullptr;
    size_t cap = 0, used = 0, peak = 0;
    std::map<size_t, size_t> holes;        // offset -> bytes, no two adjacent
    std::map<size_t, size_t> live;         // offset -> bytes
    std::mutex mu;
};
static constexpr size_t POOL_ALIGN = 256;

static void *pool_take(mxc_pool *P, size_t bytes) {
    bytes = (bytes + POOL_ALIGN - 1) / POOL_ALIGN * POOL_ALIGN;
    std::lock_guard<std::mutex> lk(P->mu);
    for (auto it = P->holes.begin(); it != P->holes.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, rest = it->second - bytes;
        P->holes.erase(it);
        if (rest) P->holes[off + bytes] = rest;
        P->live[off] = bytes;
        P->used += bytes;
        P->peak = std::max(P->peak, P->used);
        return P->base + off;
    }
    return nullptr;
}

static void pool_give(mxc_pool *P, void *ptr) {
    std::lock_guard<std::mutex> lk(P->mu);
    const size_t off = (size_t)((char *)ptr - P->base);
    auto lv = P->live.find(off);
    if (lv == P->live.end()) return;
    size_t a = off, b = off + lv->second;
    P->used -= lv->second;
    P->live.erase(lv);
    auto nx = P->holes.lower_bound(a);
    if (nx != P->holes.end() && nx->first == b) {            // merge with the hole after
        b += nx->second;
        nx = P->holes.erase(nx);
    }
    if (nx != P->holes.begin()) {                           // merge with the hole before
        auto pv = std::prev(nx);
        if (pv->first + pv->second == a) {
            a = pv->first;
            P->holes.erase(pv);
        }
    }
    P->holes[a] = b - a;
}

struct StepBlocks {
    std::vector<DiagTask> diags;
    std::vector<UpdDesc> trsm, trail;
    std::vector<UpdDesc> trail2;   // right of the next panel (second stream)
    bool panel_end = false;        // trail holds the update of the next panel's columns
    std::vector<int64_t> diag_s, trail_s;   // supernode of each diag task / trail block
    std::vector<int64_t> trail2_s;          // supernode of each trail2 block
    std::vector<int64_t> trsm_s;            // supernode of each TRSM block
};

struct mxc_factor {
    int precision = MXC_FP64;
    int64_t n = 0, nsuper = 0, xsize = 0, nnz_a = 0, nnz_full = 0;
    int64_t wsize = 0;          // xsize + inverse blocks
    int nlevels = 0;
    int64_t npairs = 0, nupd_tiles = 0;
    double fl = 0, lnz = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;       // look-ahead: trailing updates right of the next panel
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool lookahead = true;
    bool split16 = true;                 // FP32 factors: split-FP16 update/trailing tiles
    bool chain = true;                   // solves: one chained launch per level
    bool diag_tails = true;              // trailing launches factor the next diagonal blocks
    bool fold_trsm = false;              // FP64: ... and apply the next step's TRSM (UpdDesc::fold)
    int *d_ready = nullptr;              // per diagonal task: published by a folding tail; + abort flag
    int64_t nready = 0;                  // diagonal tasks covered by d_ready
    bool merge_t2 = true;                // FP64: fold second-stream updates into later trailing launches
    int t2_per_cu = 3;                   // FP64 second-stream trailing launches: workgroups per CU (0: one per tile)
    int64_t small_max = 0;               // FP64: a level's small supernodes go to one k_small_level launch
                                         // when it has at most this many of them (0: never)
    int64_t nsmall = 0;                  // supernodes the plan factors that way
    bool deterministic = false;          // descendant updates in rounds: one pair per target per launch
    size_t bytes = 0;
    mxc_pool *pool = nullptr;            // device memory comes from this pool when set
    // device time of the last factorization / solve (event pairs on the factor's
    // stream around each call) and their running sums
    hipEvent_t ev_f0 = nullptr, ev_f1 = nullptr, ev_s0 = nullptr, ev_s1 = nullptr;
    double ms_factor = 0, ms_solve = 0, ms_factor_sum = 0, ms_solve_sum = 0;
    int64_t nfactor = 0, nsolve = 0;
    long long *d_seg = nullptr;          // refinement over segments: offsets and per-segment norms,
    double *d_segout = nullptr;          // grown on demand
    int64_t seg_cap = 0, segout_cap = 0;

    // host copies
    std::vector<int64_t> Super, Lpi, Lpx, Perm;
    std::vector<int> level_of;  // per supernode

    // device
    void *W = nullptr;                  // T[wsize]
    unsigned *d_Lh = nullptr;           // FP32 factors: split-FP16 pairs of L (xsize)
    int *d_Ls = nullptr;
    long long *d_pos = nullptr, *d_dpos = nullptr, *d_perm = nullptr;
    UpdDesc *d_desc = nullptr;
    Tile *d_tiles = nullptr;
    DiagTask *d_diag = nullptr;
    SmallTask *d_small = nullptr;
    SweepTask *d_stask = nullptr;
    int *d_fail = nullptr;            // [0] first failing column, [1..3] FP64 redo queue counters
                                      // (queued, refactored in FP64, finished redo workgroups)
    int *d_pinv = nullptr;            // inverse permutation (n), for the FP64 redo's A entries
    DiagTask *d_redo = nullptr;       // FP64 redo queue (one slot per diagonal task of the plan)
    double *d_adiag = nullptr;        // assembled diagonal of A (n)
    long long *d_rowptr = nullptr, *d_src = nullptr;
    int *d_col = nullptr;
    double *d_vals = nullptr, *d_Ax = nullptr;
    double *d_norm = nullptr;
    double *d_beta = nullptr;
    void *d_work = nullptr;             // T[n * work_nrhs] solve vector
    void *d_part = nullptr;             // T[nslots * 64 * work_nrhs] backward partial sums
    int *d_cnt = nullptr;               // chained sweeps: finished blocks per supernode (fwd, bwd), and
                                        // finished below workgroups per block (bwd)
    int64_t nbcnt = 0;                  // blocks of the supernodes that are not small
    int *d_abort = nullptr;             // chained sweeps: a wait gave up
    int64_t nslots = 0;
    double *d_b = nullptr, *d_x = nullptr, *d_r = nullptr;   // FP64 n * work_nrhs
    int64_t work_nrhs = 0;

    std::vector<Launch> factor_launches;   // per level: UPD, then (DIAG, TRSM, TRAIL)*
    std::vector<Launch> fwd_launches, bwd_launches;
    std::vector<Launch> fwd_chain, bwd_chain;   // chained sweeps: per level small + chain
    bool values_loaded = false;
    int64_t minor = 0;                     // L->minor of the last factorization
    int64_t nredo = 0;                     // diagonal blocks refactored in FP64 (last factorization)
    double rho = 0;                        // pivot-ratio threshold (0: exact semantics)
    // FP64 fallback factor of a low-precision factor whose refinement stalls,
    // or whose last factorization met a pivot <= 0 (escalated: the FP64
    // factor then decides the status and serves every solve until the next
    // factorization)
    mxc_factor *fallback = nullptr;
    bool escalated = false;
    std::vector<int64_t> Ls_host, Ap_host, Ai_host;
    std::vector<int64_t> IPerm;              // built on the first sparse right-hand-side solve
    // host copy of the factorization's blocks (re-flattened after autotuning)
    std::vector<std::vector<UpdDesc>> blk_upd;        // per level
    std::vector<std::vector<StepBlocks>> blk_steps;   // per level, per 64-column step
    size_t plan_bytes = 0;                            // bytes of d_desc + d_tiles + d_diag + d_small

    // tile edge per (M, N) bucket: 32 up to 64 x 32, 128 for M > 128 and N > 32,
    // else 64 (mxc_set_tile overrides an entry before the plan is built)
    int tune[NKIND][NBUCKET];

    // hipGraph replay: the launch sequence depends only on the pattern, so it is
    // captured once and replayed for every refactorization / solve
    bool use_graphs = true;
    hipGraph_t g_factor = nullptr;
    hipGraphExec_t x_factor = nullptr;
    std::map<int64_t, std::pair<hipGraph_t, hipGraphExec_t>> g_sweep;   // key: nrhs*4 + fwd*2 + bwd

    // algorithmic flops per factorization, by kernel class
    double flops_update = 0, flops_potrf = 0, flops_trsm = 0, flops_trail = 0;

    // rank-k update/downdate (mxc_updown): column -> supernode, supernodal
    // parents and inverse-block bases, built on the first call; its device
    // buffers are kept for later calls
    std::vector<int64_t> col_super, sparent_h, winv_base_h;
    void *d_mw = nullptr;         // T[n * UD_RANK]: the columns of C of one pass
    void *d_mpar = nullptr;       // T[UD_RANK * NB * UD_PAR]: pivot data of the blocks of a launch
    void *d_malpha = nullptr;     // T[UD_RANK]
    int *d_mflag = nullptr;       // [0] first column whose new pivot is not positive, [1] bounded pivots
    void *d_mdesc = nullptr;      // blocks, row ranges and C's entries of one pass
    size_t mdesc_cap = 0;
    bool modified = false;        // L changed by mxc_updown since the last factorization

    // timing: an event pair around each launch of a class whose bit is set
    unsigned timing_mask = 0;
    double t_ms[10] = {0};
    int64_t counts[5] = {0};
    std::vector<hipEvent_t> ev;            // pool, grown on demand
    size_t ev_used = 0;
    std::vector<std::pair<int, size_t>> pending;   // (class, index of the start event)
};

static inline mxc_factor *active(mxc_factor *F);
static inline const mxc_factor *active(const mxc_factor *F);

// Device memory of a factor: from its pool when it has one, else hipMalloc.
static int falloc(mxc_factor *F, void **p, size_t bytes) {
    if (F->pool) {
        *p = pool_take(F->pool, bytes);
        if (!*p) {
            snprintf(g_err, sizeof(g_err), "device pool exhausted: %zu bytes requested, %zu of %zu in use",
                     bytes, F->pool->used, F->pool->cap);
            return MXC_OUT_OF_MEMORY;
        }
        return MXC_OK;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_hip_err(e, "hipMalloc", __LINE__);
    }
    return MXC_OK;
}

// Release memory of a factor.  A pool region can be handed out again at once,
// so the factor's streams are drained first (hipFree synchronizes by itself).
static void ffree(mxc_factor *F, void *p) {
    if (!p) return;
    if (F->pool) {
        if (F->stream) (void)hipStreamSynchronize(F->stream);
        if (F->stream2) (void)hipStreamSynchronize(F->stream2);
        pool_give(F->pool, p);
    } else {
        (void)hipFree(p);
    }
}

template <typename X>
static int dalloc(mxc_factor *F, X **p, size_t n) {
    if (n == 0) n = 1;
    int rc = falloc(F, (void **)p, n * sizeof(X));
    if (rc) return rc;
    F->bytes += n * sizeof(X);
    return MXC_OK;
}

template <typename X>
static int dupload(mxc_factor *F, X **p, const std::vector<X> &h, hipStream_t s) {
    int rc = dalloc(F, p, h.size());
    if (rc) return rc;
    if (!h.empty()) HIPTRY(hipMemcpyAsync(*p, h.data(), h.size() * sizeof(X), hipMemcpyHostToDevice, s));
    return MXC_OK;
}

// ---------------------------------------------------------------------------
// plan construction (host)
// ---------------------------------------------------------------------------
static inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Outer panel width of the in-supernode blocked Cholesky (a multiple of NB).
static constexpr int64_t PANEL = 384;

// local row of global row g inside supernode s (rows Ls[psi..psi+nsrow) sorted,
// first nscol of them = k1..k2-1)
static int64_t local_row(const mxc_symbolic *S, int64_t s, int64_t g) {
    int64_t k1 = S->Super[s], k2 = S->Super[s + 1];
    if (g >= k1 && g < k2) return g - k1;
    int64_t lo = S->Lpi[s] + (k2 - k1), hi = S->Lpi[s + 1];
    const int64_t *it = std::lower_bound(S->Ls + lo, S->Ls + hi, g);
    if (it == S->Ls + hi || *it != g) return -1;
    return (int64_t)(it - S->Ls) - S->Lpi[s];
}

static void drop_factor_graph(mxc_factor *F) {
    if (F->x_factor) (void)hipGraphExecDestroy(F->x_factor);
    if (F->g_factor) (void)hipGraphDestroy(F->g_factor);
    F->x_factor = nullptr;
    F->g_factor = nullptr;
}

// Turn the host block lists into device descriptors, tiles and the launch
// list, one launch per tile class (from F->tune) per phase.
// FP64 trailing launches of at most this many tiles carry the next step's
// diagonal blocks as tails: the FP64 tail kernels run at half the occupancy of
// the plain ones (2 waves / SIMD, two resident workgroups per CU).  The FP32
// tail kernels keep most of theirs and take tails in every trailing launch
// (96^3 27-point FP32 factor: 72.3 ms unbounded, 73.0 ms at 512 tiles).
constexpr long long PD_MAX_TILES_FP64 = 512;

// Tiles of edge ts over block u (its lower part when lower_only).
static long long lower_tiles(const UpdDesc &u, int ts, bool lower_only) {
    const long long mt = ceil_div(u.M, ts), nt = ceil_div(u.N, ts);
    return lower_only ? mt * nt - nt * (nt - 1) / 2 : mt * nt;
}

// Columns [a, b) of trailing block u (relative to its first column): the
// lower trapezoid of those columns, rows from a down.
static UpdDesc column_slice(const UpdDesc &u, int a, int b) {
    UpdDesc v = u;
    v.x_off += a;
    v.y_off += a;
    v.c_off += a + (long long)a * u.ldc;
    v.M = u.M - a;
    v.N = b - a;
    v.ls_off = u.ls_off + a;
    return v;
}

static int flatten_upload(mxc_factor *F) {
    std::vector<UpdDesc> descs;
    std::vector<Tile> tiles;
    std::vector<DiagTask> diags;
    std::vector<SmallTask> smalls;
    F->factor_launches.clear();
    int64_t nupd = 0;
    // FP64: supernodes of at most SMALL_COLS columns and SMALL_NROW rows that
    // have descendant updates are factored whole by k_small_level, one launch
    // at the head of their level (their update pairs, diagonal tasks and TRSM
    // blocks leave the level's other launches), in a level with at most
    // small_max of them: a short level is latency-bound, and the fused launch
    // replaces three launches there and runs a supernode's pairs four at a
    // time; a wide level keeps the separate launches, whose independent tiles
    // keep more waves in flight.  Those without updates (the leaves) keep the
    // one-lane and one-wave POTRF / TRSM launches.  Not in deterministic mode
    // (the fused launch adds a supernode's pairs with LDS atomics).
    // task_of[s]: the level's task of supernode s, or -1; npair[s]: its pairs.
    const bool fuse = F->small_max > 0 && F->precision == MXC_FP64 && !F->deterministic;
    std::vector<int> task_of(fuse ? F->nsuper : 0, -1), npair(fuse ? F->nsuper : 0, 0);
    auto small_s = [&](int64_t s) {
        return F->Super[s + 1] - F->Super[s] <= SMALL_COLS && F->Lpi[s + 1] - F->Lpi[s] <= SMALL_NROW;
    };
    auto super_of = [&](long long base) {
        return (int64_t)(std::upper_bound(F->Super.begin(), F->Super.end(), (int64_t)base) - F->Super.begin()) - 1;
    };
    F->nsmall = 0;
    // follow[i] (trailing blocks): the next step's diagonal task of block i's
    // supernode, or null.  A block tiled at 64 or 128 gets it as its tail (pd);
    // fused[i] records that, so the next step's diagonal launch skips it.
    // force[i] > 0 sets block i's tile edge (merged second-stream slices take the
    // edge of their supernode's own trailing block, so they share its launch).
    // foldable[i] (trailing blocks, FP64): the next step's TRSM of block i's
    // supernode may be applied by block i's tiles (see UpdDesc::fold); folded[i]
    // records that it was (64-wide tiles with a tail), so that TRSM is skipped.
    auto emit = [&](int kind, const std::vector<UpdDesc> &blocks, bool lower_only,
                    const std::vector<const DiagTask *> *follow = nullptr, std::vector<char> *fused = nullptr,
                    const std::vector<int> *force = nullptr, const std::vector<char> *foldable = nullptr,
                    std::vector<char> *folded = nullptr) {
        // FP64 descendant updates of at most 16 rows and columns take 16-wide tiles when a
        // level holds very many of them (the bottom levels of a forest of small systems:
        // millions); a 3D grid's levels (at most ~30,000 at 64^3) keep 32-wide tiles, which
        // measured faster there.  FP32 measured slower with them.
        auto tiny16 = [&](const UpdDesc &u) {
            return F->precision == MXC_FP64 && kind == L_UPD && u.M <= 16 && u.N <= 16 &&
                   F->tune[kind_of(kind)][tile_bucket(u.M, u.N, u.K)] == 32;
        };
        long long n16 = 0;
        for (const UpdDesc &u : blocks) n16 += tiny16(u) ? 1 : 0;
        const bool use16 = n16 >= 200000;
        auto edge = [&](size_t bi) {
            const UpdDesc &u = blocks[bi];
            if (force && (*force)[bi] > 0) return (*force)[bi];
            return (use16 && tiny16(u)) ? 16 : F->tune[kind_of(kind)][tile_bucket(u.M, u.N, u.K)];
        };
        long long total = 0;
        for (int ts : {16, 32, 64, 128}) {
            long long b = (long long)tiles.size();
            // tails only in launches that fit on the device at once (a tail's
            // kernel runs at lower occupancy)
            bool allow = follow != nullptr && ts >= 64;
            if (allow) {
                long long cnt = 0;
                for (size_t bi = 0; bi < blocks.size(); bi++)
                    if (edge(bi) == ts) cnt += lower_tiles(blocks[bi], ts, lower_only);
                allow = F->precision != MXC_FP64 || cnt <= PD_MAX_TILES_FP64;
            }
            bool tails = false;
            std::vector<Tile> producers;      // tile (0, 0) of each folding block: dispatched first
            for (size_t bi = 0; bi < blocks.size(); bi++) {
                const UpdDesc &u = blocks[bi];
                if (edge(bi) != ts) continue;
                int di = (int)descs.size();
                descs.push_back(u);
                descs.back().pd = -1;
                descs.back().fold = 0;
                if (allow && bi < follow->size() && (*follow)[bi]) {
                    descs.back().pd = (int)diags.size();
                    diags.push_back(*(*follow)[bi]);
                    (*fused)[bi] = 1;
                    tails = true;
                    if (ts == NB && foldable && bi < foldable->size() && (*foldable)[bi]) {
                        descs.back().fold = 1;
                        (*folded)[bi] = 1;
                    }
                }
                const bool prod_first = descs.back().fold != 0;
                if (prod_first) producers.push_back(Tile{di, 0, 0});
                int64_t mt = ceil_div(u.M, ts), nt = ceil_div(u.N, ts);
                for (int64_t ti = 0; ti < mt; ti++)
                    for (int64_t tj = 0; tj < nt; tj++) {
                        if (lower_only && ti < tj) continue;   // tile wholly above the diagonal
                        if (prod_first && ti == 0 && tj == 0) continue;
                        tiles.push_back(Tile{di, (int)ti, (int)tj});
                    }
            }
            if (!producers.empty()) tiles.insert(tiles.begin() + b, producers.begin(), producers.end());
            long long c = (long long)tiles.size() - b;
            if (c > 0) {
                Launch L{tails ? (int)L_TRAILD : kind, b, c, ts};
                L.nfirst = (int)producers.size();
                F->factor_launches.push_back(L);
            }
            total += c;
        }
        return total;
    };
    for (size_t l = 0; l < F->blk_upd.size(); l++) {
        bool level_forked = false;
        const std::vector<UpdDesc> *upd = &F->blk_upd[l];
        std::vector<UpdDesc> rest;                 // update pairs of the other targets
        std::vector<int64_t> fused_small;          // the level's small supernodes (ascending)
        if (fuse && !F->blk_steps[l].empty()) {
            const StepBlocks &s0 = F->blk_steps[l][0];
            const long long t0 = (long long)smalls.size();
            std::vector<int64_t> tgt(F->blk_upd[l].size());
            for (size_t q = 0; q < tgt.size(); q++) {
                tgt[q] = super_of(F->blk_upd[l][q].base);
                npair[tgt[q]]++;
            }
            for (size_t q = 0; q < s0.diags.size(); q++) {
                const int64_t s = s0.diag_s[q];
                if (!small_s(s) || npair[s] == 0) continue;
                task_of[s] = (int)(smalls.size() - t0);
                smalls.push_back(SmallTask{(long long)F->Lpx[s], s0.diags[q].winv,
                                           (int)(F->Lpi[s + 1] - F->Lpi[s]), (int)(F->Super[s + 1] - F->Super[s]),
                                           (int)F->Super[s], (int)F->Lpi[s], 0, 0});
                fused_small.push_back(s);
            }
            long long nt = (long long)smalls.size() - t0;
            if (nt > F->small_max) {            // a wide level: the separate launches
                for (int64_t s : fused_small) task_of[s] = -1;
                fused_small.clear();
                smalls.resize(t0);
                nt = 0;
            }
            if (nt > 0) {
                std::vector<std::vector<UpdDesc>> pairs(nt);
                for (size_t q = 0; q < tgt.size(); q++) {
                    const UpdDesc &u = F->blk_upd[l][q];
                    if (task_of[tgt[q]] >= 0) pairs[task_of[tgt[q]]].push_back(u);
                    else rest.push_back(u);
                }
                for (long long i = 0; i < nt; i++) {
                    SmallTask &st = smalls[t0 + i];
                    st.u0 = (int)descs.size();
                    st.nu = (int)pairs[i].size();
                    for (const UpdDesc &u : pairs[i]) {
                        descs.push_back(u);
                        descs.back().pd = -1;
                        descs.back().fold = 0;
                    }
                }
                for (int64_t s : fused_small) task_of[s] = -1;
                std::sort(fused_small.begin(), fused_small.end());
                F->factor_launches.push_back(Launch{L_SMALL, t0, nt, 0});
                F->nsmall += nt;
                upd = &rest;
            }
        }
        auto fused_here = [&](size_t t, int64_t s) {
            return t == 0 && std::binary_search(fused_small.begin(), fused_small.end(), s);
        };
        if (!F->deterministic) {
            nupd += emit(L_UPD, *upd, true);
        } else {
            // Rounds: round r holds the r-th pair (by descendant, ascending) of every
            // target supernode of the level, so no two tiles of a launch add into the
            // same entry of L and the sum into each entry has a fixed order: d
            // ascending, one launch after the other.  (Within one pair the relative
            // map sends distinct C entries to distinct entries of the target.)
            std::map<long long, int> rank;          // target (its L offset) -> pairs seen so far
            std::vector<std::vector<UpdDesc>> rounds;
            for (const UpdDesc &u : *upd) {
                const int r = rank[u.c_off]++;
                if ((int)rounds.size() <= r) rounds.resize(r + 1);
                rounds[r].push_back(u);
            }
            for (const auto &round : rounds) nupd += emit(L_UPD, round, true);
        }
        std::vector<int64_t> fused_s;              // supernodes whose diagonal block of this step is
                                                   // factored by the previous step's trailing launch
        std::vector<int64_t> folded_s;             // ... and whose TRSM of this step it applied
        const size_t nst = F->blk_steps[l].size();
        // second-stream updates merged into later trailing launches: per step, the
        // column slices and their supernodes
        std::vector<std::vector<UpdDesc>> merged(nst);
        std::vector<std::vector<int64_t>> merged_s(nst);
        auto edge_of = [&](const UpdDesc &u) { return F->tune[kind_of(L_TRAIL)][tile_bucket(u.M, u.N, u.K)]; };
        for (size_t t = 0; t < nst; t++) {
            const StepBlocks &sb = F->blk_steps[l][t];
            // blocks of at most TINY_DIAG columns: one lane each; at most 16: one
            // wave each; wider blocks: a workgroup each
            for (int cls = 2; cls >= 0; cls--) {
                long long db = (long long)diags.size();
                for (size_t q = 0; q < sb.diags.size(); q++) {
                    const DiagTask &d = sb.diags[q];
                    if (std::binary_search(fused_s.begin(), fused_s.end(), sb.diag_s[q])) continue;
                    if (fused_here(t, sb.diag_s[q])) continue;
                    const int c = d.nb <= TINY_DIAG ? 2 : d.nb <= 16 ? 1 : 0;
                    if (c == cls) diags.push_back(d);
                }
                long long dc = (long long)diags.size() - db;
                if (dc > 0)
                    F->factor_launches.push_back(
                        Launch{cls == 2 ? L_DIAG_TINY : cls == 1 ? L_DIAG_SMALL : L_DIAG, db, dc, 0});
            }
            {
                // TRSM of supernodes with at most TINY_DIAG columns and TINY_TRSM_ROWS rows below: one lane each
                std::vector<UpdDesc> big;
                long long b = (long long)tiles.size();
                for (size_t q = 0; q < sb.trsm.size(); q++) {
                    const UpdDesc &u = sb.trsm[q];
                    if (std::binary_search(folded_s.begin(), folded_s.end(), sb.trsm_s[q])) continue;
                    if (fused_here(t, sb.trsm_s[q])) continue;
                    if (u.K <= TINY_DIAG && u.M <= TINY_TRSM_ROWS) {
                        descs.push_back(u);
                        descs.back().pd = -1;
                        tiles.push_back(Tile{(int)descs.size() - 1, 0, 0});
                    } else {
                        big.push_back(u);
                    }
                }
                const long long c = (long long)tiles.size() - b;
                if (c > 0) F->factor_launches.push_back(Launch{L_TRSM_TINY, b, c, 0});
                emit(L_TRSM, big, false);
            }
            // At the end of panel P: wait for the second-stream update of panel
            // P-1 (it also wrote the columns of panel P+1), update panel P+1's
            // columns here, then start panel P's own second-stream update (columns
            // right of panel P+1; it needs only the finished panel P).  Panel P+1's
            // steps then overlap the second-stream update.
            // (a join only after a fork: the second stream must already be part of
            // the work, also inside a graph capture)
            if (sb.panel_end && level_forked) F->factor_launches.push_back(Launch{L_JOIN, 0, 0, 0});
            bool fork_t2 = false;
            if (!sb.trail2.empty()) {
                // FP64: fold the update right of the next panel into the next panel's
                // three trailing launches, one column slice each, where it runs beside
                // their diagonal tails (under graph replay the second stream's branch
                // ran after the main one).  Each slice takes its supernode's own
                // trailing edge; every merged launch must stay within the tail limit.
                bool merge = F->merge_t2 && F->precision == MXC_FP64 && t + 3 < nst;
                std::vector<std::vector<UpdDesc>> add(3);
                std::vector<std::vector<int64_t>> add_s(3);
                for (size_t q = 0; merge && q < sb.trail2.size(); q++) {
                    const UpdDesc &u = sb.trail2[q];
                    const double M = u.M, tot = (double)u.N * M - (double)u.N * u.N / 2.0;
                    int cut[4] = {0, 0, 0, u.N};
                    for (int k = 1; k <= 2; k++) {      // area-balanced cuts on 64-column boundaries
                        int a = cut[k - 1];
                        while (a < u.N && ((double)a * M - (double)a * a / 2.0) < tot * k / 3.0) a += NB;
                        cut[k] = std::min(a, u.N);
                    }
                    for (int k = 0; k < 3; k++) {
                        if (cut[k + 1] <= cut[k]) continue;
                        add[k].push_back(column_slice(u, cut[k], cut[k + 1]));
                        add_s[k].push_back(sb.trail2_s[q]);
                    }
                }
                for (int k = 0; merge && k < 3; k++) {
                    const StepBlocks &tg = F->blk_steps[l][t + 1 + k];
                    long long cnt[3] = {0, 0, 0};
                    auto slot = [](int ts) { return ts == 32 ? 0 : ts == 64 ? 1 : 2; };
                    for (const UpdDesc &u : tg.trail) cnt[slot(edge_of(u))] += lower_tiles(u, edge_of(u), true);
                    for (size_t q = 0; q < add[k].size(); q++) {
                        int e = 0;
                        for (size_t r = 0; r < tg.trail.size(); r++)
                            if (tg.trail_s[r] == add_s[k][q]) e = edge_of(tg.trail[r]);
                        if (e == 0) { merge = false; break; }
                        cnt[slot(e)] += lower_tiles(add[k][q], e, true);
                    }
                    for (int c = 0; merge && c < 3; c++) merge = cnt[c] <= PD_MAX_TILES_FP64;
                }
                if (merge) {
                    for (int k = 0; k < 3; k++) {
                        merged[t + 1 + k].insert(merged[t + 1 + k].end(), add[k].begin(), add[k].end());
                        merged_s[t + 1 + k].insert(merged_s[t + 1 + k].end(), add_s[k].begin(), add_s[k].end());
                    }
                } else {
                    fork_t2 = true;
                }
            }
            // the next step's diagonal blocks, one per trailing block, and (FP64)
            // whether its TRSM can be folded in: a full 64-column diagonal block
            // with rows below it, all inside this trailing block
            std::vector<const DiagTask *> follow(sb.trail.size(), nullptr);
            std::vector<char> fused(sb.trail.size(), 0), foldable(sb.trail.size(), 0),
                folded(sb.trail.size(), 0);
            if (F->diag_tails && t + 1 < F->blk_steps[l].size()) {
                // the lists ascend by supernode
                const StepBlocks &nx = F->blk_steps[l][t + 1];
                size_t r = 0, rt = 0;
                for (size_t q = 0; q < sb.trail.size(); q++) {
                    while (r < nx.diags.size() && nx.diag_s[r] < sb.trail_s[q]) r++;
                    if (r < nx.diags.size() && nx.diag_s[r] == sb.trail_s[q]) follow[q] = &nx.diags[r];
                    while (rt < nx.trsm.size() && nx.trsm_s[rt] < sb.trail_s[q]) rt++;
                    const bool has_trsm = rt < nx.trsm.size() && nx.trsm_s[rt] == sb.trail_s[q];
                    foldable[q] = F->fold_trsm && F->precision == MXC_FP64 && follow[q] &&
                                  follow[q]->nb == NB && has_trsm && nx.trsm[rt].K == NB &&
                                  sb.trail[q].N >= NB && sb.trail[q].M == nx.trsm[rt].M + NB;
                }
            }
            if (merged[t].empty()) {
                emit(L_TRAIL, sb.trail, true, &follow, &fused, nullptr, &foldable, &folded);
            } else {
                // this step's trailing blocks, then the merged slices at their
                // supernodes' edges
                std::vector<UpdDesc> blocks(sb.trail);
                std::vector<int> force(sb.trail.size(), 0);
                for (size_t q = 0; q < merged[t].size(); q++) {
                    int e = 0;
                    for (size_t r = 0; r < sb.trail.size(); r++)
                        if (sb.trail_s[r] == merged_s[t][q]) e = edge_of(sb.trail[r]);
                    blocks.push_back(merged[t][q]);
                    force.push_back(e);
                }
                emit(L_TRAIL, blocks, true, &follow, &fused, &force, &foldable, &folded);
            }
            if (fork_t2) {
                // the second-stream update starts when this step's own trailing
                // update has finished, together with the next diagonal block's
                // factorization: that single workgroup then finds a CU before the
                // second stream's tiles fill them all (started beside the trailing
                // update, it waited for the whole second-stream launch to drain)
                F->factor_launches.push_back(Launch{L_FORK, 0, 0, 0});
                emit(L_TRAIL2, sb.trail2, true);
                level_forked = true;
            }
            fused_s.clear();
            folded_s.clear();
            for (size_t q = 0; q < sb.trail.size(); q++) {
                if (fused[q]) fused_s.push_back(sb.trail_s[q]);
                if (folded[q]) folded_s.push_back(sb.trail_s[q]);
            }
        }
        // the next level reads this level's columns: join the second stream
        if (level_forked) F->factor_launches.push_back(Launch{L_JOIN, 0, 0, 0});
    }
    F->nupd_tiles = nupd;
    if (descs.size() >= (size_t)INT32_MAX || tiles.size() >= (size_t)INT32_MAX) return MXC_TOO_LARGE;
    std::lock_guard<std::recursive_mutex> lk0(g_rt_mutex);
    drop_factor_graph(F);
    ffree(F, F->d_desc);
    ffree(F, F->d_tiles);
    ffree(F, F->d_diag);
    ffree(F, F->d_small);
    F->d_desc = nullptr;
    F->d_tiles = nullptr;
    F->d_diag = nullptr;
    F->d_small = nullptr;
    F->bytes -= F->plan_bytes;
    size_t before = F->bytes;
    int rc;
    if ((rc = dupload(F, &F->d_desc, descs, F->stream))) return rc;
    if ((rc = dupload(F, &F->d_tiles, tiles, F->stream))) return rc;
    if ((rc = dupload(F, &F->d_diag, diags, F->stream))) return rc;
    if (!smalls.empty() && (rc = dupload(F, &F->d_small, smalls, F->stream))) return rc;
    // publication flags of the diagonal tasks (folded TRSMs) + the wait-abort flag
    ffree(F, F->d_ready);
    F->d_ready = nullptr;
    F->nready = 0;
    if ((rc = dalloc(F, &F->d_ready, diags.size() + 1))) return rc;
    F->nready = (int64_t)diags.size() + 1;
    if (F->precision != MXC_FP64) {
        // FP32 factors: room in the FP64 redo queue for every diagonal task
        ffree(F, F->d_redo);
        F->d_redo = nullptr;
        if ((rc = dalloc(F, &F->d_redo, diags.size()))) return rc;
    }
    F->plan_bytes = F->bytes - before;
    HIPTRY(hipStreamSynchronize(F->stream));
    return MXC_OK;
}

template <typename T>
static int build_plan(mxc_factor *F, const mxc_symbolic *S, const int64_t *Ap, const int64_t *Ai) {
    const int64_t n = S->n, ns = S->nsuper;
    F->n = n;
    F->nsuper = ns;
    F->xsize = S->xsize;
    F->fl = S->fl;
    F->lnz = S->lnz;
    F->Super.assign(S->Super, S->Super + ns + 1);
    F->Lpi.assign(S->Lpi, S->Lpi + ns + 1);
    F->Lpx.assign(S->Lpx, S->Lpx + ns + 1);
    F->Perm.assign(S->Perm, S->Perm + n);
    if (n >= (int64_t)INT32_MAX || S->ssize >= (int64_t)INT32_MAX) return MXC_TOO_LARGE;

    std::vector<int64_t> SuperMap(n);
    for (int64_t s = 0; s < ns; s++)
        for (int64_t k = S->Super[s]; k < S->Super[s + 1]; k++) SuperMap[k] = s;
    std::vector<int64_t> Pinv(n);
    for (int64_t k = 0; k < n; k++) Pinv[S->Perm[k]] = k;

    // ---- value positions of tril(A) inside L (permutation folded in)
    const int64_t nnz = Ap[n];
    F->nnz_a = nnz;
    std::vector<long long> pos(nnz > 0 ? nnz : 1, 0);
    for (int64_t j = 0; j < n; j++) {
        for (int64_t p = Ap[j]; p < Ap[j + 1]; p++) {
            int64_t i = Ai[p];
            if (i < j) {   // strictly upper entries are ignored, as CHOLMOD does for stype < 0
                pos[p] = F->xsize;   // dump slot (the first inverse-block entry, rewritten later)
                continue;
            }
            int64_t a = Pinv[i], b = Pinv[j];
            int64_t row = std::max(a, b), col = std::min(a, b);
            int64_t s = SuperMap[col];
            int64_t lr = local_row(S, s, row);
            if (lr < 0) return MXC_INVALID;   // pattern of A is not inside the analyzed L
            int64_t nsrow = S->Lpi[s + 1] - S->Lpi[s];
            pos[p] = S->Lpx[s] + lr + (col - S->Super[s]) * nsrow;
        }
    }
    std::vector<long long> dpos(n > 0 ? n : 1, 0);
    for (int64_t k = 0; k < n; k++) {
        int64_t s = SuperMap[k];
        int64_t nsrow = S->Lpi[s + 1] - S->Lpi[s];
        int64_t c = k - S->Super[s];
        dpos[k] = S->Lpx[s] + c + c * nsrow;
    }

    // ---- full CSR of A for the FP64 residual
    std::vector<long long> rowptr(n + 1, 0);
    for (int64_t j = 0; j < n; j++)
        for (int64_t p = Ap[j]; p < Ap[j + 1]; p++) {
            int64_t i = Ai[p];
            if (i < j) continue;
            rowptr[i + 1]++;
            if (i != j) rowptr[j + 1]++;
        }
    for (int64_t i = 0; i < n; i++) rowptr[i + 1] += rowptr[i];
    F->nnz_full = rowptr[n];
    std::vector<int> colv(F->nnz_full > 0 ? F->nnz_full : 1);
    std::vector<long long> src(F->nnz_full > 0 ? F->nnz_full : 1);
    {
        std::vector<long long> w(rowptr.begin(), rowptr.end() - 1);
        for (int64_t j = 0; j < n; j++)
            for (int64_t p = Ap[j]; p < Ap[j + 1]; p++) {
                int64_t i = Ai[p];
                if (i < j) continue;
                long long q = w[i]++;
                colv[q] = (int)j;
                src[q] = p;
                if (i != j) {
                    q = w[j]++;
                    colv[q] = (int)i;
                    src[q] = p;
                }
            }
    }

    // ---- supernodal etree and levels (height from the leaves)
    std::vector<int64_t> sparent(ns, -1);
    std::vector<int> level(ns, 0);
    for (int64_t s = 0; s < ns; s++) {
        int64_t nscol = S->Super[s + 1] - S->Super[s], nsrow = S->Lpi[s + 1] - S->Lpi[s];
        if (nsrow > nscol) sparent[s] = SuperMap[S->Ls[S->Lpi[s] + nscol]];
    }
    int nlev = ns > 0 ? 1 : 0;
    for (int64_t s = 0; s < ns; s++) {
        if (sparent[s] >= 0) level[sparent[s]] = std::max(level[sparent[s]], level[s] + 1);
        nlev = std::max(nlev, level[s] + 1);
    }
    F->nlevels = nlev;
    F->level_of = level;
    std::vector<std::vector<int64_t>> by_level(nlev);
    for (int64_t s = 0; s < ns; s++) by_level[level[s]].push_back(s);

    // ---- inverse-block slots: block (s,t) of width nb -> nb*nb entries
    std::vector<int64_t> winv_base(ns + 1, 0);
    int64_t woff = F->xsize;
    for (int64_t s = 0; s < ns; s++) {
        int64_t nscol = S->Super[s + 1] - S->Super[s];
        winv_base[s] = woff;
        for (int64_t c0 = 0; c0 < nscol; c0 += NB) {
            int64_t nb = std::min<int64_t>(NB, nscol - c0);
            woff += nb * nb;
        }
    }
    F->wsize = woff;

    // ---- (d, s) update pairs, grouped by the level of s
    std::vector<std::vector<UpdDesc>> pairs_by_level(nlev);
    int64_t npairs = 0;
    for (int64_t d = 0; d < ns; d++) {
        int64_t ndcol = S->Super[d + 1] - S->Super[d];
        int64_t pdi = S->Lpi[d], pdend = S->Lpi[d + 1], ndrow = pdend - pdi;
        int64_t q = pdi + ndcol;
        while (q < pdend) {
            int64_t s = SuperMap[S->Ls[q]];
            int64_t k2 = S->Super[s + 1];
            int64_t q2 = q;
            while (q2 < pdend && S->Ls[q2] < k2) q2++;
            UpdDesc u;
            memset(&u, 0, sizeof(u));
            int64_t p = q - pdi;   // Lpos: first row of d inside s
            u.x_off = S->Lpx[d] + p;
            u.y_off = u.x_off;
            u.ldx = u.ldy = (int)ndrow;
            u.K = (int)ndcol;
            u.M = (int)(pdend - q);       // ndrow2
            u.N = (int)(q2 - q);          // ndrow1
            u.c_off = S->Lpx[s];
            u.ldc = (int)(S->Lpi[s + 1] - S->Lpi[s]);
            u.ls_off = (int)q;
            u.base = (int)S->Super[s];
            u.t_psi = (int)S->Lpi[s];
            u.t_nscol = (int)(k2 - S->Super[s]);
            u.t_nsrow = u.ldc;
            pairs_by_level[level[s]].push_back(u);
            npairs++;
            {
                double n1 = (double)u.N, n3 = (double)(u.M - u.N), kk = (double)u.K;
                F->flops_update += kk * n1 * (n1 + 1) + 2.0 * n3 * n1 * kk;   // dsyrk + dgemm
            }
            q = q2;
        }
    }
    F->npairs = npairs;

    // ---- per level: the update blocks, then per 64-column step the diagonal
    // blocks, TRSM panels and trailing blocks (kept on the host so that the
    // tile assignment can be redone by the autotuner)
    F->blk_upd = std::move(pairs_by_level);
    F->blk_steps.assign(nlev, {});
    for (int l = 0; l < nlev; l++) {
        int64_t maxsteps = 0;
        for (int64_t s : by_level[l])
            maxsteps = std::max(maxsteps, ceil_div(S->Super[s + 1] - S->Super[s], NB));
        F->blk_steps[l].resize(maxsteps);
        for (int64_t t = 0; t < maxsteps; t++) {
            StepBlocks &sb = F->blk_steps[l][t];
            for (int64_t s : by_level[l]) {
                int64_t nscol = S->Super[s + 1] - S->Super[s];
                int64_t nsrow = S->Lpi[s + 1] - S->Lpi[s];
                int64_t c0 = t * NB;
                if (c0 >= nscol) continue;
                int64_t nb = std::min<int64_t>(NB, nscol - c0);
                int64_t psx = S->Lpx[s];
                int64_t wv = winv_base[s];
                for (int64_t cc = 0; cc < c0; cc += NB) {
                    int64_t b2 = std::min<int64_t>(NB, nscol - cc);
                    wv += b2 * b2;
                }
                DiagTask dt;
                dt.off = psx + c0 + c0 * nsrow;
                dt.winv = wv;
                dt.ld = (int)nsrow;
                dt.nb = (int)nb;
                dt.col0 = (int)(S->Super[s] + c0);
                dt.pad = 0;
                sb.diags.push_back(dt);
                sb.diag_s.push_back(s);
                F->flops_potrf += (double)nb * nb * nb / 3.0;
                int64_t r0 = c0 + nb;
                if (nsrow > r0) {
                    UpdDesc u;
                    memset(&u, 0, sizeof(u));
                    u.x_off = wv;
                    u.ldx = (int)nb;
                    u.y_off = psx + r0 + c0 * nsrow;
                    u.ldy = (int)nsrow;
                    u.K = (int)nb;
                    u.M = (int)(nsrow - r0);
                    u.N = (int)nb;
                    u.c_off = u.y_off;
                    u.ldc = (int)nsrow;
                    u.ls_off = (int)(S->Lpi[s] + r0);   // global rows (split-FP16 copy)
                    sb.trsm.push_back(u);
                    sb.trsm_s.push_back(s);
                    F->flops_trsm += (double)u.M * nb * nb;
                }
                // two-level blocking: a 64-column step updates the rest of its
                // PANEL-column panel only; the last step of a panel then updates
                // everything right of the panel with K = PANEL
                const int64_t pstart = (c0 / PANEL) * PANEL;
                const int64_t pend = std::min<int64_t>(nscol, pstart + PANEL);
                if (pend > r0) {
                    UpdDesc u;
                    memset(&u, 0, sizeof(u));
                    u.x_off = psx + r0 + c0 * nsrow;
                    u.y_off = u.x_off;
                    u.ldx = u.ldy = (int)nsrow;
                    u.K = (int)nb;
                    u.M = (int)(nsrow - r0);
                    u.N = (int)(pend - r0);
                    u.c_off = psx + r0 + r0 * nsrow;
                    u.ls_off = (int)(S->Lpi[s] + r0);   // global rows (for per-row scales)
                    u.ldc = (int)nsrow;
                    sb.trail.push_back(u);
                    sb.trail_s.push_back(s);
                    double n1 = (double)u.N, n3 = (double)(u.M - u.N), kk = (double)u.K;
                    F->flops_trail += kk * n1 * (n1 + 1) + 2.0 * n3 * n1 * kk;
                }
                if (r0 == pend && pend < nscol) {
                    // update by the finished panel of everything right of it, split
                    // at the end of the next panel: the next panel's columns (on the
                    // critical path) and the rest, which overlaps the next panel's
                    // steps on the second stream
                    const int64_t split = std::min<int64_t>(nscol, pend + PANEL);
                    for (int part = 0; part < 2; part++) {
                        const int64_t c_lo = part == 0 ? pend : split;
                        const int64_t c_hi = part == 0 ? split : nscol;
                        if (c_lo >= c_hi) continue;
                        UpdDesc u;
                        memset(&u, 0, sizeof(u));
                        u.x_off = psx + c_lo + pstart * nsrow;
                        u.y_off = u.x_off;
                        u.ldx = u.ldy = (int)nsrow;
                        u.K = (int)(pend - pstart);
                        u.M = (int)(nsrow - c_lo);
                        u.N = (int)(c_hi - c_lo);
                        u.c_off = psx + c_lo + c_lo * nsrow;
                        u.ls_off = (int)(S->Lpi[s] + c_lo);
                        u.ldc = (int)nsrow;
                        (part == 0 ? sb.trail : sb.trail2).push_back(u);
                        (part == 0 ? sb.trail_s : sb.trail2_s).push_back(s);
                        double n1 = (double)u.N, n3 = (double)(u.M - u.N), kk = (double)u.K;
                        F->flops_trail += kk * n1 * (n1 + 1) + 2.0 * n3 * n1 * kk;
                    }
                    sb.panel_end = true;
                }
            }
        }
    }

    // ---- solve tasks.  Per level, launch t carries the diag workgroups of
    // block t and the off workgroups of block t-1 of every supernode of the
    // level (t = 0 .. max steps; forward ascending, backward descending).
    std::vector<SweepTask> st;
    F->fwd_launches.clear();
    F->bwd_launches.clear();
    const int ROWS = 256;
    auto make_task = [&](int64_t s, int64_t blk) {
        SweepTask x{};
        int64_t nscol = S->Super[s + 1] - S->Super[s];
        x.lx = S->Lpx[s];
        int64_t wv = winv_base[s];
        for (int64_t cc = 0; cc < blk * NB; cc += NB) {
            int64_t b2 = std::min<int64_t>(NB, nscol - cc);
            wv += b2 * b2;
        }
        x.winv = wv;
        x.nsrow = (int)(S->Lpi[s + 1] - S->Lpi[s]);
        x.nscol = (int)nscol;
        x.k1 = (int)S->Super[s];
        x.psi = (int)S->Lpi[s];
        x.c0 = (int)(blk * NB);
        x.nb = (int)std::min<int64_t>(NB, nscol - blk * NB);
        return x;
    };
    auto nsteps = [&](int64_t s) { return ceil_div(S->Super[s + 1] - S->Super[s], NB); };
    // off workgroups of block t-1 cover the rows below block t
    auto push_off = [&](int64_t s, int64_t t, int *slot) {
        int64_t nscol = S->Super[s + 1] - S->Super[s];
        int64_t nsrow = S->Lpi[s + 1] - S->Lpi[s];
        for (int64_t r = std::min<int64_t>(nscol, (t + 1) * NB); r < nsrow; r += ROWS) {
            SweepTask x = make_task(s, t - 1);
            x.r0 = (int)r;
            x.r1 = (int)std::min<int64_t>(nsrow, r + ROWS);
            if (slot) x.p0 = (*slot)++;
            st.push_back(x);
        }
    };
    // small supernodes: one wave solves the whole supernode (one launch per level)
    auto is_small = [&](int64_t s) {
        int64_t nscol = S->Super[s + 1] - S->Super[s];
        int64_t nsrow = S->Lpi[s + 1] - S->Lpi[s];
        return nscol <= NB && nsrow - nscol <= SMALL_ROWS;
    };
    // tiny supernodes (one lane each) first, then the other small ones (one wave each)
    auto is_tiny = [&](int64_t s) {
        return S->Super[s + 1] - S->Super[s] <= TINY_COLS && S->Lpi[s + 1] - S->Lpi[s] <= TINY_ROWS;
    };
    // quarter-wave supernodes (sixteen lanes each), between the two
    auto size_class = [&](int64_t s) {
        if (is_tiny(s)) return 0;
        const int64_t nscol = S->Super[s + 1] - S->Super[s], nsrow = S->Lpi[s + 1] - S->Lpi[s];
        return (nscol <= QUAD_COLS && nsrow - nscol <= QUAD_ROWS) ? 1 : 2;
    };
    auto push_small = [&](int l, std::vector<Launch> &out, int kind) {
        const bool f = kind == L_FSMALL;
        const int kinds[3] = {f ? L_FTINY : L_BTINY, f ? L_FQUAD : L_BQUAD, kind};
        for (int c = 0; c < 3; c++) {
            long long b = (long long)st.size();
            for (int64_t s : by_level[l])
                if (is_small(s) && size_class(s) == c) st.push_back(make_task(s, 0));
            if ((long long)st.size() > b) out.push_back(Launch{kinds[c], b, (long long)st.size() - b, 0});
        }
    };
    for (int l = 0; l < nlev; l++) {
        push_small(l, F->fwd_launches, L_FSMALL);
        int64_t maxsteps = 0;
        for (int64_t s : by_level[l])
            if (!is_small(s)) maxsteps = std::max(maxsteps, nsteps(s));
        for (int64_t t = 0; t <= maxsteps; t++) {
            long long b = (long long)st.size();
            for (int64_t s : by_level[l])
                if (!is_small(s) && t < nsteps(s)) {
                    SweepTask x = make_task(s, t);
                    x.r0 = t >= 1 ? NB : 0;
                    st.push_back(x);
                }
            int ndiag = (int)((long long)st.size() - b);
            for (int64_t s : by_level[l])
                if (!is_small(s) && t >= 1 && t <= nsteps(s)) push_off(s, t, nullptr);
            if ((long long)st.size() > b)
                F->fwd_launches.push_back(Launch{L_FSTEP, b, (long long)st.size() - b, ndiag});
        }
    }
    int nslot = 0;
    std::vector<int> pend_p0(ns, 0), pend_np(ns, 0);
    for (int l = nlev - 1; l >= 0; l--) {
        push_small(l, F->bwd_launches, L_BSMALL);
        int64_t maxsteps = 0;
        for (int64_t s : by_level[l])
            if (!is_small(s)) maxsteps = std::max(maxsteps, nsteps(s));
        for (int64_t t = maxsteps; t >= 0; t--) {
            long long b = (long long)st.size();
            for (int64_t s : by_level[l])
                if (!is_small(s) && t < nsteps(s)) {
                    SweepTask x = make_task(s, t);
                    int64_t nscol = S->Super[s + 1] - S->Super[s];
                    x.r0 = (t + 1 < nsteps(s)) ? (int)std::min<int64_t>(NB, nscol - (t + 1) * NB) : 0;
                    x.p0 = pend_p0[s];
                    x.np = pend_np[s];
                    st.push_back(x);
                }
            int ndiag = (int)((long long)st.size() - b);
            for (int64_t s : by_level[l])
                if (!is_small(s) && t >= 1 && t <= nsteps(s)) {
                    int first = nslot;
                    push_off(s, t, &nslot);
                    pend_p0[s] = first;
                    pend_np[s] = nslot - first;
                }
            if ((long long)st.size() > b)
                F->bwd_launches.push_back(Launch{L_BSTEP, b, (long long)st.size() - b, ndiag});
        }
    }

    // ---- chained sweeps: per level, the small supernodes' launch, then one
    // launch whose workgroups are the blocks of the other supernodes in an
    // order where every block follows the blocks it depends on (forward: by
    // block index, then the rows below each supernode; backward: by distance
    // from the supernode's last block).  p0 = the supernode's counter.
    F->fwd_chain.clear();
    F->bwd_chain.clear();
    for (int l = 0; l < nlev; l++) {
        push_small(l, F->fwd_chain, L_FSMALL);
        int64_t maxsteps = 0;
        for (int64_t s : by_level[l])
            if (!is_small(s)) maxsteps = std::max(maxsteps, nsteps(s));
        long long b = (long long)st.size();
        for (int64_t t = 0; t < maxsteps; t++)
            for (int64_t s : by_level[l])
                if (!is_small(s) && t < nsteps(s)) {
                    SweepTask x = make_task(s, t);
                    x.p0 = (int)s;
                    x.np = (int)nsteps(s);
                    st.push_back(x);
                }
        for (int64_t s : by_level[l]) {
            if (is_small(s)) continue;
            int64_t nscol = S->Super[s + 1] - S->Super[s];
            int64_t nsrow = S->Lpi[s + 1] - S->Lpi[s];
            for (int64_t r = nscol; r < nsrow; r += NB) {
                SweepTask x = make_task(s, 0);
                x.r0 = (int)r;
                x.r1 = (int)std::min<int64_t>(nsrow, r + NB);
                x.p0 = (int)s;
                x.np = (int)nsteps(s);
                x.pad[0] = 1;
                st.push_back(x);
            }
        }
        if ((long long)st.size() > b) F->fwd_chain.push_back(Launch{L_FCHAIN, b, (long long)st.size() - b, 0});
    }
    // backward: the below workgroups of a level first (rows below each block,
    // CHAIN_ROWS at a time, one slot each; slots numbered per level; the last
    // block's first, as the chain needs them), then the blocks.  Counters:
    // [ns, 2 ns) blocks finished per supernode; [3 ns + bc_off[s] + t) below
    // workgroups done for block t of s.
    const int CHAIN_ROWS = 256;
    std::vector<int64_t> bc_off(ns, 0);
    int64_t nbc = 0;
    for (int64_t s = 0; s < ns; s++)
        if (!is_small(s)) {
            bc_off[s] = nbc;
            nbc += nsteps(s);
        }
    F->nbcnt = nbc;
    for (int l = nlev - 1; l >= 0; l--) {
        push_small(l, F->bwd_chain, L_BSMALL);
        int64_t maxsteps = 0;
        for (int64_t s : by_level[l])
            if (!is_small(s)) maxsteps = std::max(maxsteps, nsteps(s));
        long long b = (long long)st.size();
        int slot = 0;
        std::vector<int> first_slot(by_level[l].size(), 0), nchunk(by_level[l].size(), 0);
        for (size_t q = 0; q < by_level[l].size(); q++) {
            const int64_t s = by_level[l][q];
            if (is_small(s)) continue;
            int64_t nscol = S->Super[s + 1] - S->Super[s];
            int64_t nsrow = S->Lpi[s + 1] - S->Lpi[s];
            nchunk[q] = (int)ceil_div(nsrow - nscol, CHAIN_ROWS);
            first_slot[q] = slot;
            slot += (int)nsteps(s) * nchunk[q];     // slot of (t, chunk): first + t * nchunk + chunk
        }
        // the below workgroups, last block first, then the diag workgroups (below
        // workgroups placed between the blocks keep slots waiting: slower)
        for (int64_t k = 0; k < maxsteps; k++)
            for (size_t q = 0; q < by_level[l].size(); q++) {
                const int64_t s = by_level[l][q];
                if (is_small(s) || k >= nsteps(s)) continue;
                const int64_t t = nsteps(s) - 1 - k;
                int64_t nscol = S->Super[s + 1] - S->Super[s];
                int64_t nsrow = S->Lpi[s + 1] - S->Lpi[s];
                int c = 0;
                for (int64_t r = nscol; r < nsrow; r += CHAIN_ROWS, c++) {
                    SweepTask x = make_task(s, t);
                    x.r0 = (int)r;
                    x.r1 = (int)std::min<int64_t>(nsrow, r + CHAIN_ROWS);
                    x.p0 = (int)(3 * ns + bc_off[s] + t);
                    x.np = first_slot[q] + (int)t * nchunk[q] + c;
                    x.pad[0] = 1;
                    st.push_back(x);
                }
            }
        for (int64_t k = 0; k < maxsteps; k++)
            for (size_t q = 0; q < by_level[l].size(); q++) {
                const int64_t s = by_level[l][q];
                if (is_small(s) || k >= nsteps(s)) continue;
                const int64_t t = nsteps(s) - 1 - k;
                SweepTask x = make_task(s, t);
                x.p0 = (int)(ns + s);
                x.np = (int)nsteps(s);
                x.q0 = first_slot[q] + (int)t * nchunk[q];
                x.nq = nchunk[q];
                x.bc = (int)(3 * ns + bc_off[s] + t);
                x.bt = nchunk[q];
                st.push_back(x);
            }
        if ((long long)st.size() > b) F->bwd_chain.push_back(Launch{L_BCHAIN, b, (long long)st.size() - b, 0});
    }
    F->nslots = nslot;

    // ---- device upload (serialized with the other threads' runtime calls)
    std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
    std::vector<int> Ls32(S->ssize);
    for (int64_t k = 0; k < S->ssize; k++) Ls32[k] = (int)S->Ls[k];
    std::vector<long long> perm64(S->Perm, S->Perm + n);
    hipStream_t sm = F->stream;
    int rc;
    if ((rc = dalloc(F, (T **)&F->W, (size_t)F->wsize))) return rc;
    if (std::is_same<T, float>::value)
        if ((rc = dalloc(F, &F->d_Lh, (size_t)F->xsize))) return rc;
    if ((rc = dupload(F, &F->d_Ls, Ls32, sm))) return rc;
    if ((rc = dupload(F, &F->d_pos, pos, sm))) return rc;
    if ((rc = dupload(F, &F->d_dpos, dpos, sm))) return rc;
    if ((rc = dupload(F, &F->d_perm, perm64, sm))) return rc;
    if ((rc = flatten_upload(F))) return rc;
    if ((rc = dupload(F, &F->d_stask, st, sm))) return rc;
    if ((rc = dupload(F, &F->d_rowptr, rowptr, sm))) return rc;
    if ((rc = dupload(F, &F->d_col, colv, sm))) return rc;
    if ((rc = dupload(F, &F->d_src, src, sm))) return rc;
    if ((rc = dalloc(F, &F->d_vals, (size_t)std::max<int64_t>(F->nnz_full, 1)))) return rc;
    if ((rc = dalloc(F, &F->d_Ax, (size_t)std::max<int64_t>(nnz, 1)))) return rc;
    if ((rc = dalloc(F, &F->d_fail, 4))) return rc;
    {
        std::vector<int> pinv32(n > 0 ? n : 1, 0);
        for (int64_t k = 0; k < n; k++) pinv32[k] = (int)Pinv[k];
        if ((rc = dupload(F, &F->d_pinv, pinv32, sm))) return rc;
    }
    if ((rc = dalloc(F, &F->d_adiag, (size_t)std::max<int64_t>(n, 1)))) return rc;
    F->Ls_host.assign(S->Ls, S->Ls + S->ssize);
    F->Ap_host.assign(Ap, Ap + n + 1);
    F->Ai_host.assign(Ai, Ai + nnz);
    if ((rc = dalloc(F, &F->d_norm, 64))) return rc;
    if ((rc = dalloc(F, &F->d_beta, 1))) return rc;
    if ((rc = dalloc(F, &F->d_cnt, (size_t)(3 * ns + F->nbcnt)))) return rc;
    if ((rc = dalloc(F, &F->d_abort, 1))) return rc;
    HIPTRY(hipMemsetAsync(F->d_abort, 0, sizeof(int), sm));
    HIPTRY(hipStreamSynchronize(sm));
    return MXC_OK;
}

static void drop_sweep_graphs(mxc_factor *F) {
    for (auto &kv : F->g_sweep) {
        if (kv.second.second) (void)hipGraphExecDestroy(kv.second.second);
        if (kv.second.first) (void)hipGraphDestroy(kv.second.first);
    }
    F->g_sweep.clear();
}

static int ensure_work(mxc_factor *F, int64_t nrhs) {
    if (F->work_nrhs >= nrhs) return MXC_OK;
    std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
    drop_sweep_graphs(F);
    size_t esz = F->precision == MXC_FP64 ? sizeof(double) : sizeof(float);
    void **bufs[] = {&F->d_work, &F->d_part, (void **)&F->d_b, (void **)&F->d_x, (void **)&F->d_r};
    for (void **b : bufs) {
        ffree(F, *b);
        *b = nullptr;
    }
    F->work_nrhs = 0;
    size_t len = (size_t)std::max<int64_t>(F->n, 1) * (size_t)nrhs;
    const size_t sizes[] = {len * esz, (size_t)std::max<int64_t>(F->nslots, 1) * NB * (size_t)nrhs * esz,
                            len * sizeof(double), len * sizeof(double), len * sizeof(double)};
    for (int k = 0; k < 5; k++) {
        int rc = falloc(F, bufs[k], sizes[k]);
        if (rc) return rc;
    }
    F->work_nrhs = nrhs;
    return MXC_OK;
}

extern "C" void mxc_factor_free(mxc_factor *F) {
    if (!F) return;
    std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
    drop_sweep_graphs(F);
    drop_factor_graph(F);
    if (F->fallback) mxc_factor_free(F->fallback);
    void *ptrs[] = {F->W, F->d_Lh, F->d_Ls, F->d_pos, F->d_dpos, F->d_perm, F->d_desc, F->d_tiles,
                    F->d_diag, F->d_stask, F->d_fail, F->d_adiag, F->d_rowptr, F->d_src, F->d_col,
                    F->d_vals, F->d_Ax, F->d_norm, F->d_beta, F->d_work, F->d_part, F->d_b, F->d_x, F->d_r,
                    F->d_cnt, F->d_abort, F->d_seg, F->d_segout, F->d_pinv, F->d_redo, F->d_ready,
                    F->d_mw, F->d_mpar, F->d_malpha, F->d_mflag, F->d_mdesc, F->d_small};
    for (void *p : ptrs) ffree(F, p);

    for (hipEvent_t e : F->ev) (void)hipEventDestroy(e);
    F->ev.clear();
    for (hipEvent_t e : {F->ev_f0, F->ev_f1, F->ev_s0, F->ev_s1})
        if (e) (void)hipEventDestroy(e);
    if (F->ev_fork) (void)hipEventDestroy(F->ev_fork);
    if (F->ev_join) (void)hipEventDestroy(F->ev_join);
    if (F->stream2) (void)hipStreamDestroy(F->stream2);
    if (F->stream) (void)hipStreamDestroy(F->stream);
    delete F;
}

extern "C" int mxc_factor_create_in(const mxc_symbolic *S, const int64_t *Ap, const int64_t *Ai,
                                    int precision, mxc_pool *pool, mxc_factor **out) {
    if (!S || !out || (S->n > 0 && (!Ap || !Ai)) || !S->Super || !S->Lpi || !S->Lpx || !S->Ls ||
        !S->Perm)
        return MXC_INVALID;
    if (precision != MXC_FP64 && precision != MXC_FP32) return MXC_INVALID;
    if (pool && pool->device != g_device) {
        snprintf(g_err, sizeof(g_err), "pool is on device %d, current device is %d", pool->device, g_device);
        return MXC_INVALID;
    }
    *out = nullptr;
    mxc_factor *F = new (std::nothrow) mxc_factor();
    if (!F) return MXC_OUT_OF_MEMORY;
    F->pool = pool;
    F->precision = precision;
    F->rho = precision == MXC_FP32 ? 1e-6 : 0.0;
    for (int k = 0; k < NKIND; k++)
        for (int b = 0; b < NBUCKET; b++) {
            int mb = b / 15, nbk = (b / 3) % 5;
            F->tune[k][b] = (mb <= 2 && nbk <= 2) ? 32 : 64;
        }
    hipError_t e = hipSetDevice(g_device);
    {
        std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&F->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&F->stream2, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&F->ev_fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&F->ev_join, hipEventDisableTiming);
        for (hipEvent_t *ev : {&F->ev_f0, &F->ev_f1, &F->ev_s0, &F->ev_s1})
            if (e == hipSuccess) e = hipEventCreate(ev);
    }
    if (e != hipSuccess) {
        int rc = set_hip_err(e, "hipStreamCreate", __LINE__);
        mxc_factor_free(F);
        return rc;
    }
    int rc = precision == MXC_FP64 ? build_plan<double>(F, S, Ap, Ai) : build_plan<float>(F, S, Ap, Ai);
    if (rc != MXC_OK) {
        mxc_factor_free(F);
        return rc;
    }
    *out = F;
    return MXC_OK;
}

extern "C" int mxc_factor_create(const mxc_symbolic *S, const int64_t *Ap, const int64_t *Ai,
                                 int precision, mxc_factor **out) {
    return mxc_factor_create_in(S, Ap, Ai, precision, nullptr, out);
}

extern "C" int mxc_pool_create(size_t bytes, mxc_pool **out) {
    if (!out || bytes == 0) return MXC_INVALID;
    *out = nullptr;
    HIPTRY(hipSetDevice(g_device));
    mxc_pool *P = new (std::nothrow) mxc_pool();
    if (!P) return MXC_OUT_OF_MEMORY;
    P->device = g_device;
    P->cap = bytes / POOL_ALIGN * POOL_ALIGN;
    hipError_t e;
    {
        std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
        e = hipMalloc((void **)&P->base, P->cap);
    }
    if (e != hipSuccess) {
        delete P;
        return set_hip_err(e, "hipMalloc (pool)", __LINE__);
    }
    P->holes[0] = P->cap;
    *out = P;
    return MXC_OK;
}

extern "C" int mxc_pool_info(const mxc_pool *P, int64_t *info) {
    if (!P || !info) return MXC_INVALID;
    mxc_pool *Q = const_cast<mxc_pool *>(P);
    std::lock_guard<std::mutex> lk(Q->mu);
    info[0] = (int64_t)P->cap;
    info[1] = (int64_t)P->used;
    info[2] = (int64_t)P->peak;
    info[3] = (int64_t)P->live.size();
    return MXC_OK;
}

extern "C" int mxc_pool_free(mxc_pool *P) {
    if (!P) return MXC_OK;
    {
        std::lock_guard<std::mutex> lk(P->mu);
        if (!P->live.empty()) {
            snprintf(g_err, sizeof(g_err), "pool still holds %zu allocations (free their factors first)",
                     P->live.size());
            return MXC_INVALID;
        }
    }
    {
        std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
        (void)hipSetDevice(P->device);
        (void)hipFree(P->base);
    }
    delete P;
    return MXC_OK;
}

extern "C" int mxc_memory_info(int64_t *free_bytes, int64_t *total_bytes) {
    size_t f = 0, t = 0;
    HIPTRY(hipSetDevice(g_device));
    HIPTRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    return MXC_OK;
}

extern "C" int mxc_factor_stats(const mxc_factor *F, int64_t *counts, double *ms) {
    if (!F) return MXC_INVALID;
    F = active(F);
    if (counts) {
        for (int k = 0; k < 8; k++) counts[k] = 0;
        for (const Launch &L : F->factor_launches) {
            switch (L.kind) {
                case L_UPD: counts[0]++; break;
                case L_DIAG: case L_DIAG_SMALL: case L_DIAG_TINY: case L_SMALL: counts[1]++; break;
                case L_TRSM: case L_TRSM_TINY: counts[2]++; break;
                case L_TRAIL: case L_TRAIL2: case L_TRAILD: counts[3]++; break;
                default: break;
            }
        }
        counts[4] = counts[0] + counts[1] + counts[2] + counts[3];
        counts[5] = (int64_t)(F->chain ? F->fwd_chain.size() + F->bwd_chain.size()
                                       : F->fwd_launches.size() + F->bwd_launches.size());
        counts[6] = F->nfactor;
        counts[7] = F->nsolve;
    }
    if (ms) {
        ms[0] = F->ms_factor;
        ms[1] = F->ms_solve;
        ms[2] = F->ms_factor_sum;
        ms[3] = F->ms_solve_sum;
    }
    return MXC_OK;
}

extern "C" int mxc_factor_info(const mxc_factor *F, int64_t *info, double *dinfo) {
    if (!F) return MXC_INVALID;
    if (info) {
        info[0] = F->n;
        info[1] = F->nsuper;
        info[2] = F->xsize;
        info[3] = F->nlevels;
        info[4] = F->npairs;
        info[5] = F->nupd_tiles;
        info[6] = (int64_t)F->bytes;
        info[7] = F->precision;
        info[8] = F->nredo;
        info[9] = F->escalated ? 1 : 0;
    }
    if (dinfo) {
        dinfo[0] = F->fl;
        dinfo[1] = F->lnz;
        dinfo[2] = F->flops_update;
        dinfo[3] = F->flops_potrf;
        dinfo[4] = F->flops_trsm;
        dinfo[5] = F->flops_trail;
    }
    return MXC_OK;
}

// ---------------------------------------------------------------------------
// timing helpers: an event pair around each timed launch (classes selected by
// the mask given to mxc_set_timing); events come from a per-factor pool
// ---------------------------------------------------------------------------
static bool next_event(mxc_factor *F, size_t *idx) {
    if (F->ev_used == F->ev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return false;
        F->ev.push_back(e);
    }
    *idx = F->ev_used++;
    return hipEventRecord(F->ev[*idx], F->stream) == hipSuccess;
}

static void tick(mxc_factor *F, int cls, size_t *slot) {
    *slot = (size_t)-1;
    if (!(F->timing_mask & (1u << cls))) return;
    size_t i;
    if (next_event(F, &i)) *slot = i;
}

static void tock(mxc_factor *F, int cls, size_t start) {
    if (start == (size_t)-1) return;
    size_t i;
    if (next_event(F, &i) && i == start + 1) F->pending.push_back({cls, start});
}

static void collect_timing(mxc_factor *F) {
    if (F->pending.empty() && F->ev_used == 0) return;
    (void)hipStreamSynchronize(F->stream);
    for (auto &pr : F->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, F->ev[pr.second], F->ev[pr.second + 1]) == hipSuccess)
            F->t_ms[pr.first] += ms;
    }
    F->pending.clear();
    F->ev_used = 0;
}

extern "C" int mxc_set_timing(mxc_factor *F, int mask) {
    if (!F) return MXC_INVALID;
    collect_timing(F);
    F->timing_mask = (mask == 1) ? 0x3ffu : (unsigned)mask;
    for (double &t : F->t_ms) t = 0;
    for (int64_t &c : F->counts) c = 0;
    return MXC_OK;
}

extern "C" int mxc_get_timing(const mxc_factor *F, double *t, int64_t *counts) {
    if (!F) return MXC_INVALID;
    collect_timing(const_cast<mxc_factor *>(F));
    if (t) for (int i = 0; i < 10; i++) t[i] = F->t_ms[i];
    if (counts) for (int i = 0; i < 5; i++) counts[i] = F->counts[i];
    return MXC_OK;
}

// ---------------------------------------------------------------------------
// factorization
// ---------------------------------------------------------------------------
static inline unsigned grid_for(long long n, int bs, long long cap = 65536) {
    long long g = (n + bs - 1) / bs;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// Low-precision factors with the pivot-ratio safeguard on refactor the
// diagonal blocks whose pivots break down in FP64 (redo_block64).
template <typename T>
static inline bool redo_safe(const mxc_factor *F) {
    return std::is_same<T, float>::value && F->rho > 0 && F->d_redo != nullptr;
}

static inline Redo redo_args(const mxc_factor *F) {
    return Redo{F->d_fail + 1, F->d_redo, (const long long *)F->d_rowptr, (const int *)F->d_col,
                (const double *)F->d_vals, (const long long *)F->d_perm, (const int *)F->d_pinv,
                (const double *)F->d_beta};
}

// CUs of the current device (queried once)
static int device_cus() {
    static int cus = 0;
    if (cus == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, g_device) != hipSuccess || v <= 0) v = 256;
        cus = v;
    }
    return cus;
}

template <typename T, int MODE, bool PD = false>
static void launch_tile(mxc_factor *F, const Launch &L, T *W, hipStream_t st = nullptr) {
    if (!st) st = F->stream;
    if constexpr (std::is_same<T, double>::value && MODE == MODE_TRAIL && !PD) {
        // FP64 second-stream update: a grid of t2_per_cu workgroups per CU that
        // walk all its tiles, so that every CU keeps room for a diagonal
        // factorization of the main stream (k_potrf_diag: 144 registers per
        // lane and 66.5 KB of LDS beside three 96-register, 17 KB tiles)
        const long long cap = ((long long)F->t2_per_cu * device_cus()) / 8 * 8;
        if (L.kind == L_TRAIL2 && L.tile == 64 && cap > 0 && L.count > cap) {
            hipLaunchKernelGGL((k_tile_grid<double, MODE_TRAIL, 64>), dim3((unsigned)cap), dim3(256), 0, st,
                               F->d_desc, F->d_tiles + L.begin, W, F->d_Ls, F->d_adiag, (int)L.count);
            return;
        }
    }
    const bool h16 = std::is_same<T, float>::value && F->split16 && F->d_Lh;
    const double *adiag = F->d_adiag;
    const PdArgs pda{F->d_diag, F->d_fail, redo_safe<T>(F) ? F->rho : 0.0, redo_args(F), F->d_ready,
                     F->d_ready ? F->d_ready + (F->nready - 1) : nullptr, L.nfirst};
    const dim3 g((unsigned)L.count);
    if constexpr (std::is_same<T, float>::value && MODE != MODE_STORE) {
        // FP32 factor: split-FP16 products on the FP16 MFMA, operands from the
        // pairs the TRSM wrote (32-wide tiles are latency-bound: FP32 MFMA)
        if (h16 && L.tile > 32) {
            if (L.tile == 64)
                hipLaunchKernelGGL((k_tile_h<MODE, 64, PD>), g, dim3(256), 0, st, F->d_desc, F->d_tiles + L.begin,
                                   W, (const unsigned *)F->d_Lh, F->d_Ls, adiag, pda);
            else
                hipLaunchKernelGGL((k_tile_h<MODE, 128, PD>), g, dim3(256), 0, st, F->d_desc, F->d_tiles + L.begin,
                                   W, (const unsigned *)F->d_Lh, F->d_Ls, adiag, pda);
            return;
        }
    }
    unsigned *Lh = (MODE == MODE_STORE && h16) ? F->d_Lh : nullptr;
    if constexpr (!PD) {
        if (L.tile == 16) {
            hipLaunchKernelGGL((k_tile<T, MODE, 16>), g, dim3(64), 0, st, F->d_desc, F->d_tiles + L.begin, W,
                               F->d_Ls, Lh, adiag, pda);
            return;
        }
        if (L.tile == 32) {
            hipLaunchKernelGGL((k_tile<T, MODE, 32>), g, dim3(64), 0, st, F->d_desc, F->d_tiles + L.begin, W,
                               F->d_Ls, Lh, adiag, pda);
            return;
        }
    }
    if (L.tile == 64)
        hipLaunchKernelGGL((k_tile<T, MODE, 64, PD>), g, dim3(256), 0, st, F->d_desc, F->d_tiles + L.begin, W,
                           F->d_Ls, Lh, adiag, pda);
    else
        hipLaunchKernelGGL((k_tile<T, MODE, 128, PD>), g, dim3(256), 0, st, F->d_desc, F->d_tiles + L.begin, W,
                           F->d_Ls, Lh, adiag, pda);
}

// Enqueue one whole factorization (values already in d_Ax, beta in d_beta).
// Kernel launches only, so the sequence can be captured into a graph.
template <typename T>
static int enqueue_factor(mxc_factor *F) {
    hipStream_t sm = F->stream;
    T *W = (T *)F->W;
    size_t t0 = 0, tt = 0;
    tick(F, 8, &t0);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, sm, F->d_fail, INT_MAX);
    for (int k = 1; k < 4; k++) hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, sm, F->d_fail + k, 0);
    const bool safe = redo_safe<T>(F);
    const Redo rq = redo_args(F);
    // after a one-wave or one-lane POTRF launch: refactor in FP64 the blocks it
    // queued (the workgroup POTRF and the diagonal tails redo their own)
    auto redo = [&]() {
        if (safe) hipLaunchKernelGGL(k_potrf_redo64<T>, dim3(REDO_WG), dim3(256), 0, sm, rq, W, F->d_fail);
    };
    // assembly: L = 0, scatter A, + beta on the diagonal; also refresh the CSR values
    tick(F, 0, &tt);
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(F->xsize, 256)), dim3(256), 0, sm, W, (long long)F->xsize, (T)0);
    if (F->d_ready)
        hipLaunchKernelGGL(k_fill<int>, dim3(grid_for(F->nready, 256)), dim3(256), 0, sm, F->d_ready,
                           (long long)F->nready, 0);
    hipLaunchKernelGGL(k_scatter_a<T>, dim3(grid_for(F->nnz_a, 256)), dim3(256), 0, sm, F->d_Ax,
                       F->d_pos, (long long)F->nnz_a, W);
    hipLaunchKernelGGL(k_add_diag<T>, dim3(grid_for(F->n, 256, 1LL << 30)), dim3(256), 0, sm,
                       F->d_dpos, (long long)F->n, (const double *)F->d_beta, W, F->d_adiag);
    hipLaunchKernelGGL(k_gather_vals, dim3(grid_for(F->nnz_full, 256)), dim3(256), 0, sm,
                       F->d_Ax, F->d_src, (long long)F->nnz_full, F->d_vals);
    tock(F, 0, tt);
    // two streams (look-ahead) except when launches are timed one by one
    const bool two = F->lookahead && F->stream2 != nullptr && F->timing_mask == 0;
    for (const Launch &L : F->factor_launches) {
        if (L.kind == L_FORK || L.kind == L_JOIN) {
            if (!two) continue;
            if (L.kind == L_FORK) {
                HIPTRY(hipEventRecord(F->ev_fork, sm));
                HIPTRY(hipStreamWaitEvent(F->stream2, F->ev_fork, 0));
            } else {
                HIPTRY(hipEventRecord(F->ev_join, F->stream2));
                HIPTRY(hipStreamWaitEvent(sm, F->ev_join, 0));
            }
            continue;
        }
        const int cls = L.kind == L_UPD ? 1
                      : (L.kind == L_DIAG || L.kind == L_DIAG_SMALL || L.kind == L_DIAG_TINY ||
                         L.kind == L_SMALL) ? 2
                      : (L.kind == L_TRSM || L.kind == L_TRSM_TINY) ? 3 : 4;
        tick(F, cls, &tt);
        switch (L.kind) {
            case L_UPD:
                launch_tile<T, MODE_SCATTER>(F, L, W);
                F->counts[0]++;
                break;
            case L_DIAG:
                if (safe)
                    hipLaunchKernelGGL((k_potrf_diag<T, true>), dim3((unsigned)L.count), dim3(256), 0,
                                       sm, F->d_diag + L.begin, W, F->d_fail,
                                       (const double *)F->d_adiag, F->rho, rq);
                else
                    hipLaunchKernelGGL((k_potrf_diag<T, false>), dim3((unsigned)L.count), dim3(256), 0,
                                       sm, F->d_diag + L.begin, W, F->d_fail,
                                       (const double *)F->d_adiag, 0.0, rq);
                F->counts[1]++;
                break;
            case L_DIAG_SMALL: {
                const unsigned g = (unsigned)((L.count + 3) / 4);
                if (safe)
                    hipLaunchKernelGGL((k_potrf_small<T, true>), dim3(g), dim3(256), 0, sm,
                                       F->d_diag + L.begin, (int)L.count, W, F->d_fail,
                                       (const double *)F->d_adiag, F->rho, rq);
                else
                    hipLaunchKernelGGL((k_potrf_small<T, false>), dim3(g), dim3(256), 0, sm,
                                       F->d_diag + L.begin, (int)L.count, W, F->d_fail,
                                       (const double *)F->d_adiag, 0.0, rq);
                redo();
                F->counts[1]++;
                break;
            }
            case L_DIAG_TINY: {
                const unsigned g = (unsigned)((L.count + 255) / 256);
                if (safe)
                    hipLaunchKernelGGL((k_potrf_tiny<T, true>), dim3(g), dim3(256), 0, sm, F->d_diag + L.begin,
                                       (int)L.count, W, F->d_fail, (const double *)F->d_adiag, F->rho, rq);
                else
                    hipLaunchKernelGGL((k_potrf_tiny<T, false>), dim3(g), dim3(256), 0, sm, F->d_diag + L.begin,
                                       (int)L.count, W, F->d_fail, (const double *)F->d_adiag, 0.0, rq);
                redo();
                F->counts[1]++;
                break;
            }
            case L_SMALL:
                if constexpr (std::is_same<T, double>::value) {
                    const long long g = std::min<long long>(L.count, (long long)device_cus() * 8);
                    hipLaunchKernelGGL(k_small_level, dim3((unsigned)g), dim3(256), 0, sm, F->d_small + L.begin,
                                       (int)L.count, F->d_desc, W, F->d_Ls, F->d_fail);
                }
                F->counts[1]++;
                break;
            case L_TRSM_TINY: {
                const bool h16 = std::is_same<T, float>::value && F->split16 && F->d_Lh;
                hipLaunchKernelGGL(k_trsm_tiny<T>, dim3((unsigned)((L.count + 255) / 256)), dim3(256), 0, sm,
                                   F->d_desc, F->d_tiles + L.begin, (int)L.count, W, F->d_Ls,
                                   h16 ? F->d_Lh : nullptr, (const double *)F->d_adiag);
                F->counts[2]++;
                break;
            }
            case L_TRSM:
                launch_tile<T, MODE_STORE>(F, L, W);
                F->counts[2]++;
                break;
            case L_TRAIL:
                launch_tile<T, MODE_TRAIL>(F, L, W);
                F->counts[3]++;
                break;
            case L_TRAILD:
                launch_tile<T, MODE_TRAIL, true>(F, L, W);
                F->counts[3]++;
                break;
            case L_TRAIL2:
                launch_tile<T, MODE_TRAIL>(F, L, W, two ? F->stream2 : sm);
                F->counts[3]++;
                break;
        }
        tock(F, cls, tt);
    }
    HIPTRY(hipGetLastError());
    tock(F, 8, t0);
    return MXC_OK;
}

// Run the factorization: replay the captured graph (capturing it on first
// use), or enqueue directly when per-launch timing is on or graphs are off.
template <typename T>
static int run_factor(mxc_factor *F) {
    hipStream_t sm = F->stream;
    if (!F->use_graphs || F->timing_mask != 0) return enqueue_factor<T>(F);
    if (!F->x_factor) {
        std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
        HIPTRY(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_factor<T>(F);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(sm, &g);
        if (rc != MXC_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_hip_err(e, "hipStreamEndCapture", __LINE__);
        e = hipGraphInstantiate(&F->x_factor, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            F->x_factor = nullptr;
            return set_hip_err(e, "hipGraphInstantiate", __LINE__);
        }
        F->g_factor = g;
    }
    std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
    HIPTRY(hipGraphLaunch(F->x_factor, sm));
    return MXC_OK;
}

// Not positive definite: CHOLMOD's result is the factor of the leading
// (minor) columns, with the failing supernode refactored for its first
// info-1 columns and everything after it zero
// (t_cholmod_super_numeric_worker.c:1084-1146).  Rare path: redo the
// factorization with supernodes past the failing one excluded, on the host
// copy of the assembled values.
template <typename T>
static int finish_not_posdef(mxc_factor *F, int fail_col, double beta);
static int fp64_fallback(mxc_factor *F);

// The factor that serves solves: the FP64 one after an escalation.
static inline mxc_factor *active(mxc_factor *F) {
    return (F && F->escalated && F->fallback) ? F->fallback : F;
}
static inline const mxc_factor *active(const mxc_factor *F) {
    return (F && F->escalated && F->fallback) ? F->fallback : F;
}

extern "C" int mxc_factorize(mxc_factor *F, const double *Ax, int location, double beta,
                             int64_t *minor) {
    if (!F || (!Ax && F->nnz_a > 0)) return MXC_INVALID;
    HIPTRY(hipSetDevice(g_device));
    hipStream_t sm = F->stream;
    F->escalated = false;
    F->modified = false;
    HIPTRY(hipEventRecord(F->ev_f0, sm));
    if (F->nnz_a > 0)
        HIPTRY(hipMemcpyAsync(F->d_Ax, Ax, F->nnz_a * sizeof(double),
                              location == MXC_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, sm));
    HIPTRY(hipMemcpyAsync(F->d_beta, &beta, sizeof(double), hipMemcpyHostToDevice, sm));
    int rc = F->precision == MXC_FP64 ? run_factor<double>(F) : run_factor<float>(F);
    if (rc) {
        (void)hipStreamSynchronize(sm);       // the queued copies read host memory of the caller
        return rc;
    }
    HIPTRY(hipEventRecord(F->ev_f1, sm));
    int flags[3] = {0, 0, 0}, wait_abort = 0;
    HIPTRY(hipMemcpyAsync(flags, F->d_fail, sizeof(int) * 3, hipMemcpyDeviceToHost, sm));
    if (F->d_ready)
        HIPTRY(hipMemcpyAsync(&wait_abort, F->d_ready + (F->nready - 1), sizeof(int), hipMemcpyDeviceToHost, sm));
    HIPTRY(hipStreamSynchronize(sm));
    if (wait_abort) {
        snprintf(g_err, sizeof(g_err), "factorization: a folded TRSM gave up waiting for its diagonal block");
        return MXC_GPU_PROBLEM;
    }
    {
        float t = 0;
        if (hipEventElapsedTime(&t, F->ev_f0, F->ev_f1) == hipSuccess) {
            F->ms_factor = t;
            F->ms_factor_sum += t;
        }
        F->nfactor++;
    }
    int fail = flags[0];
    F->nredo = flags[2];
    F->values_loaded = true;
    if (fail == 0x7fffffff) {
        F->minor = F->n;
        if (minor) *minor = F->n;
        return MXC_OK;
    }
    if (F->precision != MXC_FP64 && F->rho > 0) {
        // A pivot <= 0 in the low-precision factor: FP64 decides.  The FP64
        // factor of the same values either succeeds (and serves the solves)
        // or reports CHOLMOD's not-positive-definite result (its minor and L).
        rc = fp64_fallback(F);
        if (rc != MXC_OK && rc != MXC_NOT_POSDEF) return rc;
        F->escalated = true;
        F->minor = F->fallback->minor;
        if (minor) *minor = F->minor;
        return rc;
    }
    F->minor = fail;
    rc = F->precision == MXC_FP64 ? finish_not_posdef<double>(F, fail, beta)
                                  : finish_not_posdef<float>(F, fail, beta);
    if (minor) *minor = fail;
    return rc == MXC_OK ? MXC_NOT_POSDEF : rc;
}

// Host completion of the not-positive-definite case.  Supernodes before the
// failing supernode s* are final on the device (none depends on s*).  s* is
// re-assembled from A and its descendants' updates on the host (a copy of
// the factored descendants is pulled back), its first info-1 columns are
// factored and solved for, and everything from column k1+info-1 on is zeroed.
template <typename T>
static int finish_not_posdef(mxc_factor *F, int fail_col, double beta) {
    hipStream_t sm = F->stream;
    const int64_t ns = F->nsuper;
    int64_t s = std::upper_bound(F->Super.begin(), F->Super.end(), (int64_t)fail_col) - F->Super.begin() - 1;
    const int64_t k1 = F->Super[s], nscol = F->Super[s + 1] - k1;
    const int64_t psi = F->Lpi[s], nsrow = F->Lpi[s + 1] - psi, psx = F->Lpx[s];
    const int64_t info = fail_col - k1 + 1;
    // host copy of everything up to and including s
    std::vector<T> Lx((size_t)F->Lpx[s + 1]);
    HIPTRY(hipMemcpyAsync(Lx.data(), F->W, Lx.size() * sizeof(T), hipMemcpyDeviceToHost, sm));
    std::vector<int> Ls((size_t)F->Lpi[ns]);
    HIPTRY(hipMemcpyAsync(Ls.data(), F->d_Ls, Ls.size() * sizeof(int), hipMemcpyDeviceToHost, sm));
    std::vector<double> Ax((size_t)std::max<int64_t>(F->nnz_a, 1));
    std::vector<long long> pos((size_t)std::max<int64_t>(F->nnz_a, 1));
    HIPTRY(hipMemcpyAsync(Ax.data(), F->d_Ax, F->nnz_a * sizeof(double), hipMemcpyDeviceToHost, sm));
    HIPTRY(hipMemcpyAsync(pos.data(), F->d_pos, F->nnz_a * sizeof(long long), hipMemcpyDeviceToHost, sm));
    HIPTRY(hipStreamSynchronize(sm));
    // re-assemble s: zero, scatter A entries that land in s, beta
    std::vector<double> Sx((size_t)(nsrow * nscol), 0.0);
    for (int64_t p = 0; p < F->nnz_a; p++) {
        long long q = pos[p];
        if (q >= psx && q < psx + nsrow * nscol) Sx[q - psx] = Ax[p];
    }
    for (int64_t c = 0; c < nscol; c++) Sx[c + c * nsrow] += beta;
    // local row map of s
    std::vector<int64_t> map_g;   // global row -> local, via binary search on Ls
    auto loc = [&](int64_t g) -> int64_t {
        if (g >= k1 && g < k1 + nscol) return g - k1;
        auto b = Ls.begin() + psi + nscol, e = Ls.begin() + psi + nsrow;
        auto it = std::lower_bound(b, e, (int)g);
        return (it != e && *it == (int)g) ? (int64_t)(it - Ls.begin()) - psi : -1;
    };
    // descendant updates into s
    for (int64_t d = 0; d < s; d++) {
        int64_t ndcol = F->Super[d + 1] - F->Super[d];
        int64_t pdi = F->Lpi[d], pdend = F->Lpi[d + 1], ndrow = pdend - pdi, pdx = F->Lpx[d];
        int64_t q = pdi + ndcol;
        while (q < pdend && Ls[q] < k1) q++;
        if (q >= pdend || Ls[q] >= k1 + nscol) continue;
        int64_t q2 = q;
        while (q2 < pdend && Ls[q2] < k1 + nscol) q2++;
        int64_t ndrow1 = q2 - q, ndrow2 = pdend - q;
        std::vector<int64_t> rel(ndrow2);
        for (int64_t i = 0; i < ndrow2; i++) rel[i] = loc(Ls[q + i]);
        for (int64_t j = 0; j < ndrow1; j++)
            for (int64_t i = j; i < ndrow2; i++) {
                double c = 0;
                for (int64_t k = 0; k < ndcol; k++)
                    c += (double)Lx[pdx + (q - pdi) + i + k * ndrow] * (double)Lx[pdx + (q - pdi) + j + k * ndrow];
                Sx[rel[i] + rel[j] * nsrow] -= c;
            }
    }
    // factor the first info-1 columns, solve the rows below them, zero the rest
    const int64_t nc = info - 1;
    for (int64_t j = 0; j < nc; j++) {
        double dj = Sx[j + j * nsrow];
        for (int64_t k = 0; k < j; k++) dj -= Sx[j + k * nsrow] * Sx[j + k * nsrow];
        dj = std::sqrt(dj);
        Sx[j + j * nsrow] = dj;
        for (int64_t i = j + 1; i < nsrow; i++) {
            double v = Sx[i + j * nsrow];
            for (int64_t k = 0; k < j; k++) v -= Sx[i + k * nsrow] * Sx[j + k * nsrow];
            Sx[i + j * nsrow] = v / dj;
        }
    }
    for (int64_t j = nc; j < nscol; j++)
        for (int64_t i = 0; i < nsrow; i++) Sx[i + j * nsrow] = 0;
    for (int64_t j = 0; j < nc; j++)       // strict upper part of the diagonal block stays 0
        for (int64_t i = 0; i < j; i++) Sx[i + j * nsrow] = 0;
    std::vector<T> St(Sx.size());
    for (size_t i = 0; i < Sx.size(); i++) St[i] = (T)Sx[i];
    T *W = (T *)F->W;
    HIPTRY(hipMemcpyAsync(W + psx, St.data(), St.size() * sizeof(T), hipMemcpyHostToDevice, sm));
    int64_t tail = F->xsize - F->Lpx[s + 1];
    if (tail > 0)
        hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(tail, 256)), dim3(256), 0, sm, W + F->Lpx[s + 1],
                           (long long)tail, (T)0);
    HIPTRY(hipStreamSynchronize(sm));
    return MXC_OK;
}

extern "C" int mxc_factor_get_Lx(const mxc_factor *F, double *Lx) {
    if (!F || !Lx) return MXC_INVALID;
    F = active(F);
    HIPTRY(hipSetDevice(g_device));
    if (F->precision == MXC_FP64) {
        HIPTRY(hipMemcpyAsync(Lx, F->W, F->xsize * sizeof(double), hipMemcpyDeviceToHost, F->stream));
        HIPTRY(hipStreamSynchronize(F->stream));
    } else {
        std::vector<float> tmp(F->xsize);
        HIPTRY(hipMemcpyAsync(tmp.data(), F->W, F->xsize * sizeof(float), hipMemcpyDeviceToHost, F->stream));
        HIPTRY(hipStreamSynchronize(F->stream));
        for (int64_t i = 0; i < F->xsize; i++) Lx[i] = tmp[i];
    }
    return MXC_OK;
}

// ---------------------------------------------------------------------------
// solves
// ---------------------------------------------------------------------------
// A level's chained sweep launch over nr <= CHAIN_RHS columns (one pass).
template <typename T, bool FWD>
static void launch_chain(mxc_factor *F, const Launch &L, T *Xw, int nr, T *part) {
    hipStream_t sm = F->stream;
    const T *W = (const T *)F->W;
    const SweepTask *tk = F->d_stask + L.begin;
    const dim3 g((unsigned)L.count), b(256);
    const long long n = (long long)F->n;
    if (FWD) {
        if (nr == 1)
            hipLaunchKernelGGL((k_chain_fwd<T, 1>), g, b, 0, sm, tk, W, F->d_Ls, Xw, n, nr, F->d_cnt, F->d_abort);
        else if (nr == 2)
            hipLaunchKernelGGL((k_chain_fwd<T, 2>), g, b, 0, sm, tk, W, F->d_Ls, Xw, n, nr, F->d_cnt, F->d_abort);
        else
            hipLaunchKernelGGL((k_chain_fwd<T, 4>), g, b, 0, sm, tk, W, F->d_Ls, Xw, n, nr, F->d_cnt, F->d_abort);
    } else {
        if (nr == 1)
            hipLaunchKernelGGL((k_chain_bwd<T, 1>), g, b, 0, sm, tk, W, F->d_Ls, Xw, part, n, nr,
                               F->d_cnt, F->d_abort);
        else if (nr == 2)
            hipLaunchKernelGGL((k_chain_bwd<T, 2>), g, b, 0, sm, tk, W, F->d_Ls, Xw, part, n, nr,
                               F->d_cnt, F->d_abort);
        else
            hipLaunchKernelGGL((k_chain_bwd<T, 4>), g, b, 0, sm, tk, W, F->d_Ls, Xw, part, n, nr,
                               F->d_cnt, F->d_abort);
    }
}

// A level's tiny supernodes, one lane each, or its quarter-wave ones, sixteen
// lanes each.
template <typename T>
static void launch_tiny(mxc_factor *F, const Launch &L, T *Xw, int nr) {
    const dim3 g((unsigned)((L.count + 255) / 256)), b(256);
    const T *W = (const T *)F->W;
    if (L.kind == L_FQUAD || L.kind == L_BQUAD) {
        const dim3 gq((unsigned)((L.count + 15) / 16));
        if (L.kind == L_FQUAD)
            hipLaunchKernelGGL(k_sweep_quad_fwd<T>, gq, b, 0, F->stream, F->d_stask + L.begin, (int)L.count, W,
                               F->d_Ls, Xw, (long long)F->n, nr);
        else
            hipLaunchKernelGGL(k_sweep_quad_bwd<T>, gq, b, 0, F->stream, F->d_stask + L.begin, (int)L.count, W,
                               F->d_Ls, Xw, (long long)F->n, nr);
        return;
    }
    if (L.kind == L_FTINY)
        hipLaunchKernelGGL(k_sweep_tiny_fwd<T>, g, b, 0, F->stream, F->d_stask + L.begin, (int)L.count, W, F->d_Ls,
                           Xw, (long long)F->n, nr);
    else
        hipLaunchKernelGGL(k_sweep_tiny_bwd<T>, g, b, 0, F->stream, F->d_stask + L.begin, (int)L.count, W, F->d_Ls,
                           Xw, (long long)F->n, nr);
}

// After the caller's synchronization of the factor's stream: a chained sweep
// whose wait gave up is a device error (the flag is cleared for the next call).
static int chain_status(mxc_factor *F) {
    if (!F->chain) return MXC_OK;
    int ab = 0;
    HIPTRY(hipMemcpyAsync(&ab, F->d_abort, sizeof(int), hipMemcpyDeviceToHost, F->stream));
    HIPTRY(hipStreamSynchronize(F->stream));
    if (!ab) return MXC_OK;
    HIPTRY(hipMemsetAsync(F->d_abort, 0, sizeof(int), F->stream));
    HIPTRY(hipStreamSynchronize(F->stream));
    snprintf(g_err, sizeof(g_err), "chained sweep: a wait for a block did not end");
    return MXC_GPU_PROBLEM;
}

// part: the backward sweep's partial-sum slots (nslots x 64 x nrhs of T);
// nullptr = the factor's own buffer.
template <typename T>
static int enqueue_sweep(mxc_factor *F, T *Xw, int64_t nrhs, bool fwd, bool bwd, T *part = nullptr) {
    hipStream_t sm = F->stream;
    if (!part) part = (T *)F->d_part;
    const T *W = (const T *)F->W;
    size_t tt = 0;
    int n = (int)F->n;
    int nr = (int)nrhs;
    const int ns = (int)F->nsuper;
    if (F->chain) {
        // chained sweeps: one pass per launch, so the columns go CHAIN_RHS at a
        // time (a block is published once all of its columns are final)
        for (int dir = 0; dir < 2; dir++) {
            if (!(dir == 0 ? fwd : bwd)) continue;
            tick(F, dir == 0 ? 5 : 6, &tt);
            for (int z0 = 0; z0 < nr; z0 += CHAIN_RHS) {
                const int zn = std::min(CHAIN_RHS, nr - z0);
                T *Xz = Xw + (size_t)z0 * (size_t)n;
                const int ncnt = 3 * ns + (int)F->nbcnt;
                if (ns > 0)
                    hipLaunchKernelGGL(k_chain_reset, dim3(grid_for(ncnt, 256, 1LL << 30)), dim3(256),
                                       0, sm, F->d_cnt, ncnt);
                for (const Launch &L : dir == 0 ? F->fwd_chain : F->bwd_chain) {
                    if (L.kind == L_FTINY || L.kind == L_BTINY || L.kind == L_FQUAD || L.kind == L_BQUAD)
                        launch_tiny<T>(F, L, Xz, zn);
                    else if (L.kind == L_FSMALL)
                        hipLaunchKernelGGL(k_sweep_small_fwd<T>, dim3((unsigned)((L.count + 3) / 4)), dim3(256),
                                           0, sm, F->d_stask + L.begin, (int)L.count, W, F->d_Ls, Xz,
                                           (long long)n, zn);
                    else if (L.kind == L_BSMALL)
                        hipLaunchKernelGGL(k_sweep_small_bwd<T>, dim3((unsigned)((L.count + 3) / 4)), dim3(256),
                                           0, sm, F->d_stask + L.begin, (int)L.count, W, F->d_Ls, Xz,
                                           (long long)n, zn);
                    else if (dir == 0)
                        launch_chain<T, true>(F, L, Xz, zn, part);
                    else
                        launch_chain<T, false>(F, L, Xz, zn, part);
                    F->counts[4]++;
                }
            }
            tock(F, dir == 0 ? 5 : 6, tt);
        }
        HIPTRY(hipGetLastError());
        return MXC_OK;
    }
    if (fwd) {
        tick(F, 5, &tt);
        for (const Launch &L : F->fwd_launches) {
            if (L.kind == L_FTINY || L.kind == L_FQUAD)
                launch_tiny<T>(F, L, Xw, nr);
            else if (L.kind == L_FSMALL)
                hipLaunchKernelGGL(k_sweep_small_fwd<T>, dim3((unsigned)((L.count + 3) / 4)), dim3(256), 0,
                                   sm, F->d_stask + L.begin, (int)L.count, W, F->d_Ls, Xw, (long long)n, nr);
            else
                hipLaunchKernelGGL(k_sweep_fwd<T>, dim3((unsigned)L.count), dim3(256), 0, sm,
                                   F->d_stask + L.begin, L.tile, W, F->d_Ls, Xw, (long long)n, nr);
            F->counts[4]++;
        }
        tock(F, 5, tt);
    }
    if (bwd) {
        tick(F, 6, &tt);
        for (const Launch &L : F->bwd_launches) {
            if (L.kind == L_BTINY || L.kind == L_BQUAD)
                launch_tiny<T>(F, L, Xw, nr);
            else if (L.kind == L_BSMALL)
                hipLaunchKernelGGL(k_sweep_small_bwd<T>, dim3((unsigned)((L.count + 3) / 4)), dim3(256), 0,
                                   sm, F->d_stask + L.begin, (int)L.count, W, F->d_Ls, Xw, (long long)n, nr);
            else
                hipLaunchKernelGGL(k_sweep_bwd<T>, dim3((unsigned)L.count), dim3(256), 0, sm,
                                   F->d_stask + L.begin, L.tile, W, F->d_Ls, Xw, part,
                                   (long long)n, nr);
            F->counts[4]++;
        }
        tock(F, 6, tt);
    }
    HIPTRY(hipGetLastError());
    return MXC_OK;
}

// Sweeps over the factor's own work vector are replayed from a graph per
// (nrhs, direction); sweeps over caller memory are enqueued directly.
template <typename T>
static int run_sweep(mxc_factor *F, T *Xw, int64_t nrhs, bool fwd, bool bwd) {
    if (!F->use_graphs || F->timing_mask != 0 || (void *)Xw != F->d_work)
        return enqueue_sweep<T>(F, Xw, nrhs, fwd, bwd);
    hipStream_t sm = F->stream;
    int64_t key = nrhs * 4 + (fwd ? 2 : 0) + (bwd ? 1 : 0);
    auto it = F->g_sweep.find(key);
    if (it == F->g_sweep.end()) {
        std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
        HIPTRY(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_sweep<T>(F, Xw, nrhs, fwd, bwd);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(sm, &g);
        if (rc != MXC_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_hip_err(e, "hipStreamEndCapture", __LINE__);
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            return set_hip_err(e, "hipGraphInstantiate", __LINE__);
        }
        it = F->g_sweep.emplace(key, std::make_pair(g, x)).first;
    }
    std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
    HIPTRY(hipGraphLaunch(it->second.second, sm));
    return MXC_OK;
}

// device-resident full solve: dX = A \ dB (both FP64, device, n x nrhs)
template <typename T>
static int solve_device(mxc_factor *F, const double *dB, int64_t nrhs, double *dX) {
    hipStream_t sm = F->stream;
    T *Xw = (T *)F->d_work;
    unsigned g = grid_for(F->n, 256, 1LL << 30);
    size_t tt = 0;
    tick(F, 9, &tt);
    hipLaunchKernelGGL(k_perm_gather<T>, dim3(g), dim3(256), 0, sm, dB, F->d_perm, (long long)F->n,
                       (int)nrhs, Xw);
    int rc = run_sweep<T>(F, Xw, nrhs, true, true);
    if (rc) return rc;
    hipLaunchKernelGGL(k_perm_scatter<T>, dim3(g), dim3(256), 0, sm, Xw, F->d_perm, (long long)F->n,
                       (int)nrhs, dX);
    tock(F, 9, tt);
    HIPTRY(hipGetLastError());
    return MXC_OK;
}

// Device time of a solve call: events around its work on the factor's stream,
// read after the call's final synchronization.
static void solve_t0(mxc_factor *F) { (void)hipEventRecord(F->ev_s0, F->stream); }
static void solve_t1(mxc_factor *F) { (void)hipEventRecord(F->ev_s1, F->stream); }
static void solve_done(mxc_factor *F) {
    float t = 0;
    if (hipEventElapsedTime(&t, F->ev_s0, F->ev_s1) == hipSuccess) {
        F->ms_solve = t;
        F->ms_solve_sum += t;
    }
    F->nsolve++;
}

static int stage_in(mxc_factor *F, const double *B, int64_t nrhs, int location, const double **dB) {
    if (location == MXC_DEVICE) {
        *dB = B;
        return MXC_OK;
    }
    HIPTRY(hipMemcpyAsync(F->d_b, B, F->n * nrhs * sizeof(double), hipMemcpyHostToDevice, F->stream));
    *dB = F->d_b;
    return MXC_OK;
}

extern "C" int mxc_solve(mxc_factor *F, const double *B, int64_t nrhs, double *X, int location) {
    if (!F || !B || !X || nrhs < 0) return MXC_INVALID;
    F = active(F);
    if (!F->values_loaded) return MXC_INVALID;
    if (nrhs == 0 || F->n == 0) return MXC_OK;
    HIPTRY(hipSetDevice(g_device));
    int rc = ensure_work(F, nrhs);
    if (rc) return rc;
    const double *dB;
    solve_t0(F);
    if ((rc = stage_in(F, B, nrhs, location, &dB))) return rc;
    double *dX = location == MXC_DEVICE ? X : F->d_x;
    rc = F->precision == MXC_FP64 ? solve_device<double>(F, dB, nrhs, dX)
                                  : solve_device<float>(F, dB, nrhs, dX);
    if (rc) {
        (void)hipStreamSynchronize(F->stream);
        return rc;
    }
    if (location != MXC_DEVICE)
        HIPTRY(hipMemcpyAsync(X, dX, F->n * nrhs * sizeof(double), hipMemcpyDeviceToHost, F->stream));
    solve_t1(F);
    HIPTRY(hipStreamSynchronize(F->stream));
    solve_done(F);
    return chain_status(F);
}

static int super_sweep(mxc_factor *F, double *X, int64_t nrhs, int location, bool fwd, bool bwd) {
    if (!F || !X || nrhs < 0) return MXC_INVALID;
    F = active(F);
    if (!F->values_loaded) return MXC_INVALID;
    if (nrhs == 0 || F->n == 0) return MXC_OK;
    HIPTRY(hipSetDevice(g_device));
    int rc = ensure_work(F, nrhs);
    if (rc) return rc;
    hipStream_t sm = F->stream;
    size_t len = (size_t)F->n * nrhs;
    solve_t0(F);
    if (location != MXC_DEVICE)
        HIPTRY(hipMemcpyAsync(F->d_x, X, len * sizeof(double), hipMemcpyHostToDevice, sm));
    double *dX = location == MXC_DEVICE ? X : F->d_x;
    if (F->precision == MXC_FP64) {
        rc = run_sweep<double>(F, dX, nrhs, fwd, bwd);
    } else {
        // through the FP32 work vector, converted on the device
        float *Xw = (float *)F->d_work;
        const dim3 g(grid_for((int64_t)len, 256, 1LL << 16));
        hipLaunchKernelGGL((k_cast<double, float>), g, dim3(256), 0, sm, (const double *)dX, (long long)len, Xw);
        rc = run_sweep<float>(F, Xw, nrhs, fwd, bwd);
        if (rc) return rc;
        hipLaunchKernelGGL((k_cast<float, double>), g, dim3(256), 0, sm, (const float *)Xw, (long long)len, dX);
        HIPTRY(hipGetLastError());
    }
    if (rc) {
        (void)hipStreamSynchronize(sm);
        return rc;
    }
    if (location != MXC_DEVICE)
        HIPTRY(hipMemcpyAsync(X, dX, len * sizeof(double), hipMemcpyDeviceToHost, sm));
    solve_t1(F);
    HIPTRY(hipStreamSynchronize(sm));
    solve_done(F);
    return chain_status(F);
}

extern "C" int mxc_super_lsolve(mxc_factor *F, double *X, int64_t nrhs, int location) {
    return super_sweep(F, X, nrhs, location, true, false);
}
extern "C" int mxc_super_ltsolve(mxc_factor *F, double *X, int64_t nrhs, int location) {
    return super_sweep(F, X, nrhs, location, false, true);
}

// Sparse right-hand side (cholmod_l_solve2 with Bset, cholmod_solve.c:290-640).
// Yset = the columns of L reached from C = the (permuted) Bset, as
// cholmod_l_lsolve_pattern orders them (mxc_reach, symbolic.c).  Yset holds
// every parent of its members, so the solution over
// Yset needs no other column: the dense device solve restricted to Yset gives
// those values.  X receives them at Xset = P(Yset) (bset_iperm,
// t_cholmod_solve_worker.c:252-283); its other entries are left as they were.
extern "C" int mxc_solve_bset(mxc_factor *F, int sys, const double *B, const int64_t *Bset, int64_t blen,
                              double *X, int64_t *Xset, int64_t *xlen) {
    if (!F || !B || !X || !Xset || !xlen || blen < 0 || (blen > 0 && !Bset)) return MXC_INVALID;
    if (sys < MXC_SYS_A || sys > MXC_SYS_D) return MXC_INVALID;
    F = active(F);
    if (!F->values_loaded) return MXC_INVALID;
    const int64_t n = F->n;
    for (int64_t p = 0; p < blen; p++)
        if (Bset[p] < 0 || Bset[p] >= n) return MXC_INVALID;
    const bool perm = sys == MXC_SYS_A;
    if (F->IPerm.size() != (size_t)n) {
        F->IPerm.assign(n, 0);
        for (int64_t k = 0; k < n; k++) F->IPerm[F->Perm[k]] = k;
    }
    std::vector<int64_t> C(blen), Y;
    for (int64_t p = 0; p < blen; p++) C[p] = perm ? F->IPerm[Bset[p]] : Bset[p];
    if (sys == MXC_SYS_D) {
        Y = C;                     // D = I for an LL' factor
    } else {
        Y.resize(n);
        int64_t ylen = 0;
        int rc = mxc_reach(n, F->nsuper, F->Super.data(), F->Lpi.data(), F->Ls_host.data(), C.data(), blen,
                           Y.data(), &ylen);
        if (rc) return rc;
        Y.resize(ylen);
    }
    std::vector<double> y(n, 0.0);
    for (int64_t p = 0; p < blen; p++) y[C[p]] = B[Bset[p]];
    if (sys != MXC_SYS_D) {
        const bool fwd = sys == MXC_SYS_A || sys == MXC_SYS_LDLt || sys == MXC_SYS_LD || sys == MXC_SYS_L;
        const bool bwd = sys == MXC_SYS_A || sys == MXC_SYS_LDLt || sys == MXC_SYS_DLt || sys == MXC_SYS_Lt;
        int rc = super_sweep(F, y.data(), 1, MXC_HOST, fwd, bwd);
        if (rc) return rc;
    }
    for (size_t p = 0; p < Y.size(); p++) {
        const int64_t iold = perm ? F->Perm[Y[p]] : Y[p];
        X[iold] = y[Y[p]];
        Xset[p] = iold;
    }
    *xlen = (int64_t)Y.size();
    return MXC_OK;
}

// ---------------------------------------------------------------------------
// residual and iterative refinement
// ---------------------------------------------------------------------------
// lanes per row from the average row length: 4 (2D 5-point), 8 (3D 7-point), 16 (27-point)
static void launch_residual(mxc_factor *F, const double *dB, const double *dX, double *dR, int64_t nrhs) {
    const double avg = F->n > 0 ? (double)F->nnz_full / (double)F->n : 0.0;
    const int lpr = avg <= 6.0 ? 4 : avg <= 20.0 ? 8 : 16;
    const dim3 g(grid_for(F->n * lpr, 256, 1LL << 30)), b(256);
    if (lpr == 4)
        hipLaunchKernelGGL(k_residual<4>, g, b, 0, F->stream, F->d_rowptr, F->d_col, F->d_vals, dB, dX, dR,
                           (long long)F->n, (int)nrhs);
    else if (lpr == 8)
        hipLaunchKernelGGL(k_residual<8>, g, b, 0, F->stream, F->d_rowptr, F->d_col, F->d_vals, dB, dX, dR,
                           (long long)F->n, (int)nrhs);
    else
        hipLaunchKernelGGL(k_residual<16>, g, b, 0, F->stream, F->d_rowptr, F->d_col, F->d_vals, dB, dX, dR,
                           (long long)F->n, (int)nrhs);
}

extern "C" int mxc_residual(mxc_factor *F, const double *B, const double *X, int64_t nrhs, double *R,
                            int location) {
    if (!F || !B || !X || !R || nrhs < 0) return MXC_INVALID;
    F = active(F);
    if (!F->values_loaded) return MXC_INVALID;
    if (nrhs == 0 || F->n == 0) return MXC_OK;
    HIPTRY(hipSetDevice(g_device));
    int rc = ensure_work(F, nrhs);
    if (rc) return rc;
    hipStream_t sm = F->stream;
    size_t len = (size_t)F->n * nrhs;
    const double *dB = B, *dX = X;
    double *dR = R;
    if (location != MXC_DEVICE) {
        HIPTRY(hipMemcpyAsync(F->d_b, B, len * sizeof(double), hipMemcpyHostToDevice, sm));
        HIPTRY(hipMemcpyAsync(F->d_x, X, len * sizeof(double), hipMemcpyHostToDevice, sm));
        dB = F->d_b;
        dX = F->d_x;
        dR = F->d_r;
    }
    size_t tt = 0;
    tick(F, 7, &tt);
    launch_residual(F, dB, dX, dR, nrhs);
    tock(F, 7, tt);
    HIPTRY(hipGetLastError());
    if (location != MXC_DEVICE) HIPTRY(hipMemcpyAsync(R, dR, len * sizeof(double), hipMemcpyDeviceToHost, sm));
    HIPTRY(hipStreamSynchronize(sm));
    return MXC_OK;
}

// 2-norms of the columns of V, per segment of rows when segments are given
// (out[g*nrhs + r]), else of whole columns (out[r]).
static int colnorms(mxc_factor *F, const double *dV, int64_t nrhs, int64_t nseg, const long long *d_seg,
                    double *d_out, std::vector<double> &out) {
    hipStream_t sm = F->stream;
    if (nseg > 0) {
        hipLaunchKernelGGL(k_seg_sumsq, dim3((unsigned)nseg), dim3(256), 0, sm, dV, (long long)F->n,
                           (int)nrhs, d_seg, d_out);
        out.assign((size_t)(nseg * nrhs), 0);
    } else {
        HIPTRY(hipMemsetAsync(d_out, 0, sizeof(double) * nrhs, sm));
        hipLaunchKernelGGL(k_sumsq, dim3(grid_for(F->n, 256, 1024)), dim3(256), 0, sm, dV, (long long)F->n,
                           (int)nrhs, d_out);
        out.assign((size_t)nrhs, 0);
    }
    HIPTRY(hipMemcpyAsync(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, sm));
    HIPTRY(hipStreamSynchronize(sm));
    for (auto &v : out) v = std::sqrt(v);
    return MXC_OK;
}

template <typename T>
static int refine_device(mxc_factor *F, const double *dB, int64_t nrhs, double *dX, double tol,
                         int maxiter, int64_t nseg, const int64_t *segoff, double *out) {
    hipStream_t sm = F->stream;
    size_t len = (size_t)F->n * nrhs;
    std::vector<double> bn, rn;
    long long *d_seg = nullptr;
    double *d_out = F->d_norm;
    if (nseg > 0) {
        // segment offsets and per-segment norms: buffers of the factor, grown on demand
        if (F->seg_cap < nseg + 1 || F->segout_cap < nseg * nrhs) {
            std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
            ffree(F, F->d_seg);
            ffree(F, F->d_segout);
            F->d_seg = nullptr;
            F->d_segout = nullptr;
            F->seg_cap = F->segout_cap = 0;
            int rc = falloc(F, (void **)&F->d_seg, sizeof(long long) * (nseg + 1));
            if (!rc) rc = falloc(F, (void **)&F->d_segout, sizeof(double) * nseg * std::max<int64_t>(nrhs, 4));
            if (rc) return rc;
            F->seg_cap = nseg + 1;
            F->segout_cap = nseg * std::max<int64_t>(nrhs, 4);
        }
        d_seg = F->d_seg;
        d_out = F->d_segout;
        HIPTRY(hipMemcpyAsync(d_seg, segoff, sizeof(long long) * (nseg + 1), hipMemcpyHostToDevice, sm));
    }
    int rc = colnorms(F, dB, nrhs, nseg, d_seg, d_out, bn);
    if (rc) return rc;
    // x0 = A \ b
    T *Xw = (T *)F->d_work;
    unsigned g = grid_for(F->n, 256, 1LL << 30);
    hipLaunchKernelGGL(k_perm_gather<T>, dim3(g), dim3(256), 0, sm, dB, F->d_perm, (long long)F->n, (int)nrhs, Xw);
    if ((rc = run_sweep<T>(F, Xw, nrhs, true, true))) return rc;
    hipLaunchKernelGGL(k_perm_scatter<T>, dim3(g), dim3(256), 0, sm, Xw, F->d_perm, (long long)F->n, (int)nrhs, dX);
    int it = 0;
    double rel = 0, rel0 = -1;
    for (;;) {
        launch_residual(F, dB, dX, F->d_r, nrhs);
        if ((rc = colnorms(F, F->d_r, nrhs, nseg, d_seg, d_out, rn))) return rc;
        if ((rc = chain_status(F))) return rc;   // colnorms synchronized the stream
        rel = 0;
        for (size_t q = 0; q < rn.size(); q++) {
            double v = bn[q] > 0 ? rn[q] / bn[q] : rn[q];
            rel = std::max(rel, v);
        }
        if (rel0 < 0) rel0 = rel;
        if (rel <= tol || it >= maxiter || !std::isfinite(rel)) break;
        // d = A \ r ; x += d
        hipLaunchKernelGGL(k_perm_gather<T>, dim3(g), dim3(256), 0, sm, F->d_r, F->d_perm, (long long)F->n, (int)nrhs, Xw);
        if ((rc = run_sweep<T>(F, Xw, nrhs, true, true))) return rc;
            hipLaunchKernelGGL(k_perm_scatter<T>, dim3(g), dim3(256), 0, sm, Xw, F->d_perm, (long long)F->n, (int)nrhs, F->d_r);
        hipLaunchKernelGGL(k_axpy_update<double>, dim3(grid_for((long long)len, 256)), dim3(256), 0, sm, dX,
                           (const double *)F->d_r, (long long)len);
        it++;
    }
    HIPTRY(hipGetLastError());
    if (out) {
        out[0] = it;
        out[1] = rel;
        out[2] = rel0;
    }
    return MXC_OK;
}

// Build (once) and factor an FP64 twin of a low-precision factor, from the
// kept symbolic arrays and the values of the last factorization.
static int fp64_fallback(mxc_factor *F) {
    if (!F->fallback) {
        mxc_symbolic S;
        memset(&S, 0, sizeof(S));
        S.n = F->n;
        S.nsuper = F->nsuper;
        S.ssize = (int64_t)F->Ls_host.size();
        S.xsize = F->xsize;
        S.fl = F->fl;
        S.lnz = F->lnz;
        S.Perm = F->Perm.data();
        S.Super = F->Super.data();
        S.Lpi = F->Lpi.data();
        S.Lpx = F->Lpx.data();
        S.Ls = F->Ls_host.data();
        int rc = mxc_factor_create_in(&S, F->Ap_host.data(), F->Ai_host.data(), MXC_FP64, F->pool,
                                      &F->fallback);
        if (rc) return rc;
    }
    double beta = 0;
    HIPTRY(hipMemcpy(&beta, F->d_beta, sizeof(double), hipMemcpyDeviceToHost));
    int64_t minor = 0;
    return mxc_factorize(F->fallback, F->d_Ax, MXC_DEVICE, beta, &minor);   // NOT_POSDEF propagates
}

static int solve_refine_impl(mxc_factor *F, const double *B, int64_t nrhs, double *X, int location,
                             double tol, int maxiter, int64_t nseg, const int64_t *segoff, double *out) {
    if (!F || !B || !X || nrhs < 0 || nrhs > 64) return MXC_INVALID;
    if (nseg < 0 || (nseg > 0 && !segoff)) return MXC_INVALID;
    if (nseg > 0) {
        if (segoff[0] != 0 || segoff[nseg] != F->n) return MXC_INVALID;
        for (int64_t g = 0; g < nseg; g++)
            if (segoff[g + 1] < segoff[g]) return MXC_INVALID;
    }
    if (!F->values_loaded) return MXC_INVALID;
    if (F->modified) {
        snprintf(g_err, sizeof(g_err), "refinement: the factor was changed by mxc_updown and no longer matches "
                                       "the stored A; factorize A +- CC' to refine against it");
        return MXC_INVALID;
    }
    if (nrhs == 0 || F->n == 0) return MXC_OK;
    HIPTRY(hipSetDevice(g_device));
    const bool escalated = F->escalated && F->fallback;
    F = active(F);
    int rc = ensure_work(F, nrhs);
    if (rc) return rc;
    const double *dB;
    solve_t0(F);
    if ((rc = stage_in(F, B, nrhs, location, &dB))) return rc;
    double *dX = location == MXC_DEVICE ? X : F->d_x;
    double o[3] = {0, 0, 0};
    rc = F->precision == MXC_FP64 ? refine_device<double>(F, dB, nrhs, dX, tol, maxiter, nseg, segoff, o)
                                  : refine_device<float>(F, dB, nrhs, dX, tol, maxiter, nseg, segoff, o);
    if (rc) return rc;
    int used_fallback = escalated ? 1 : 0;
    if (F->precision != MXC_FP64 && !(o[1] <= tol)) {
        // the low-precision factor did not reach tol: refactor in FP64 and refine with it
        if ((rc = fp64_fallback(F))) return rc;
        mxc_factor *G = F->fallback;
        if ((rc = ensure_work(G, nrhs))) return rc;
        HIPTRY(hipStreamSynchronize(F->stream));
        double o2[3] = {0, 0, 0};
        rc = refine_device<double>(G, dB, nrhs, dX, tol, maxiter, nseg, segoff, o2);
        if (rc) return rc;
        HIPTRY(hipStreamSynchronize(G->stream));
        o[0] += o2[0];
        o[1] = o2[1];
        used_fallback = 1;
    }
    if (out) {
        out[0] = o[0];
        out[1] = o[1];
        out[2] = o[2];
        out[3] = used_fallback;
    }
    if (location != MXC_DEVICE)
        HIPTRY(hipMemcpyAsync(X, dX, F->n * nrhs * sizeof(double), hipMemcpyDeviceToHost, F->stream));
    solve_t1(F);
    HIPTRY(hipStreamSynchronize(F->stream));
    solve_done(F);
    return MXC_OK;
}

// ---------------------------------------------------------------------------
// autotuner: per (M, N) bucket of update blocks, time each tile class on (a
// sample of) that bucket's blocks and keep the fastest
// ---------------------------------------------------------------------------
template <typename T>
static int time_tiles(mxc_factor *F, int kind, const std::vector<UpdDesc> &blocks, int ts, int reps,
                      float *ms) {
    std::vector<UpdDesc> descs(blocks);
    std::vector<Tile> tiles;
    for (size_t di = 0; di < descs.size(); di++) {
        const UpdDesc &u = descs[di];
        int64_t mt = ceil_div(u.M, ts), nt = ceil_div(u.N, ts);
        for (int64_t ti = 0; ti < mt; ti++)
            for (int64_t tj = 0; tj < nt; tj++)
                if (kind == 2 || ti >= tj) tiles.push_back(Tile{(int)di, (int)ti, (int)tj});
    }
    UpdDesc *d_desc = nullptr;
    Tile *d_tiles = nullptr;
    HIPTRY(hipMalloc((void **)&d_desc, descs.size() * sizeof(UpdDesc)));
    hipError_t e = hipMalloc((void **)&d_tiles, tiles.size() * sizeof(Tile));
    if (e != hipSuccess) {
        (void)hipFree(d_desc);
        return set_hip_err(e, "hipMalloc", __LINE__);
    }
    hipStream_t sm = F->stream;
    (void)hipMemcpyAsync(d_desc, descs.data(), descs.size() * sizeof(UpdDesc), hipMemcpyHostToDevice, sm);
    (void)hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, sm);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    T *W = (T *)F->W;
    Launch L{kind == 0 ? L_UPD : kind == 1 ? L_TRAIL : L_TRSM, 0, (long long)tiles.size(), ts};
    UpdDesc *save_desc = F->d_desc;
    Tile *save_tiles = F->d_tiles;
    F->d_desc = d_desc;
    F->d_tiles = d_tiles;
    auto launch = [&]() {
        if (kind == 0) launch_tile<T, MODE_SCATTER>(F, L, W);
        else if (kind == 1) launch_tile<T, MODE_TRAIL>(F, L, W);
        else launch_tile<T, MODE_STORE>(F, L, W);
    };
    launch();                        // warm-up
    (void)hipEventRecord(e0, sm);
    for (int r = 0; r < reps; r++) launch();
    (void)hipEventRecord(e1, sm);
    e = hipEventSynchronize(e1);
    F->d_desc = save_desc;
    F->d_tiles = save_tiles;
    float t = 0;
    if (e == hipSuccess) (void)hipEventElapsedTime(&t, e0, e1);
    *ms = t / (float)std::max(reps, 1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(d_desc);
    (void)hipFree(d_tiles);
    if (e != hipSuccess) return set_hip_err(e, "autotune launch", __LINE__);
    HIPTRY(hipGetLastError());
    return MXC_OK;
}

template <typename T>
static int autotune_impl(mxc_factor *F, int reps, int *table) {
    const size_t cap = 4096;         // blocks timed per bucket (evenly strided sample)
    for (int kind = 0; kind < NKIND; kind++) {
        std::vector<std::vector<UpdDesc>> byb(NBUCKET);
        for (size_t l = 0; l < F->blk_upd.size(); l++) {
            if (kind == 0) {
                for (const UpdDesc &u : F->blk_upd[l]) byb[tile_bucket(u.M, u.N, u.K)].push_back(u);
            } else {
                for (const StepBlocks &sb : F->blk_steps[l])
                    for (const UpdDesc &u : (kind == 1 ? sb.trail : sb.trsm))
                        byb[tile_bucket(u.M, u.N, u.K)].push_back(u);
            }
        }
        for (int b = 0; b < NBUCKET; b++) {
            if (byb[b].empty()) continue;
            std::vector<UpdDesc> sample;
            size_t stride = std::max<size_t>(1, byb[b].size() / cap);
            for (size_t q = 0; q < byb[b].size() && sample.size() < cap; q += stride)
                sample.push_back(byb[b][q]);
            float best = -1;
            int best_ts = F->tune[kind][b];
            for (int ts : {32, 64, 128}) {
                float ms = 0;
                int rc = time_tiles<T>(F, kind, sample, ts, reps, &ms);
                if (rc) return rc;
                if (best < 0 || ms < best) {
                    best = ms;
                    best_ts = ts;
                }
            }
            F->tune[kind][b] = best_ts;
        }
    }
    F->values_loaded = false;        // the timing launches scribbled on L
    int rc = flatten_upload(F);
    if (rc) return rc;
    if (table)
        for (int k = 0; k < NKIND; k++)
            for (int b = 0; b < NBUCKET; b++) table[k * NBUCKET + b] = F->tune[k][b];
    return MXC_OK;
}

extern "C" int mxc_autotune(mxc_factor *F, int reps, int *table) {
    if (!F || reps < 1) return MXC_INVALID;
    HIPTRY(hipSetDevice(g_device));
    std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);   // times launches on scratch descriptors
    return F->precision == MXC_FP64 ? autotune_impl<double>(F, reps, table)
                                    : autotune_impl<float>(F, reps, table);
}

extern "C" int mxc_set_tiles(mxc_factor *F, const int *table) {
    if (!F || !table) return MXC_INVALID;
    for (int q = 0; q < NKIND * NBUCKET; q++)
        if (table[q] != 32 && table[q] != 64 && table[q] != 128) return MXC_INVALID;
    for (int k = 0; k < NKIND; k++)
        for (int b = 0; b < NBUCKET; b++) F->tune[k][b] = table[k * NBUCKET + b];
    return flatten_upload(F);
}

extern "C" int mxc_get_tiles(const mxc_factor *F, int *table) {
    if (!F || !table) return MXC_INVALID;
    for (int k = 0; k < NKIND; k++)
        for (int b = 0; b < NBUCKET; b++) table[k * NBUCKET + b] = F->tune[k][b];
    return MXC_OK;
}

extern "C" int mxc_set_option(mxc_factor *F, int option, int64_t value) {
    if (!F) return MXC_INVALID;
    std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);     // graphs are destroyed below
    switch (option) {
        case MXC_OPT_GRAPHS:
            F->use_graphs = value != 0;
            return MXC_OK;
        case MXC_OPT_PIVOT_RHO_E9:
            if (value < 0) return MXC_INVALID;
            F->rho = (double)value * 1e-9;
            drop_factor_graph(F);
            return MXC_OK;
        case MXC_OPT_LOOKAHEAD:
            F->lookahead = value != 0;
            drop_factor_graph(F);
            return MXC_OK;
        case MXC_OPT_SPLIT16:
            F->split16 = value != 0;
            drop_factor_graph(F);
            return MXC_OK;
        case MXC_OPT_SOLVE_CHAIN:
            F->chain = value != 0;
            drop_sweep_graphs(F);
            return MXC_OK;
        case MXC_OPT_DETERMINISTIC:
            if (F->deterministic == (value != 0)) return MXC_OK;
            F->deterministic = value != 0;
            return flatten_upload(F);           // new launch list (drops the factor graph)
        case MXC_OPT_T2_GRID:
            if (value < 0 || value > 64) return MXC_INVALID;
            F->t2_per_cu = (int)value;
            drop_factor_graph(F);
            return MXC_OK;
        case MXC_OPT_FOLD_TRSM:
            if (F->fold_trsm == (value != 0)) return MXC_OK;
            F->fold_trsm = value != 0;
            return flatten_upload(F);
        case MXC_OPT_SMALL_FUSED:
            if (value < 0) return MXC_INVALID;
            if (F->small_max == value) return MXC_OK;
            F->small_max = value;
            return flatten_upload(F);
        default:
            return MXC_INVALID;
    }
}

extern "C" int mxc_solve_refine(mxc_factor *F, const double *B, int64_t nrhs, double *X, int location,
                                double tol, int maxiter, double *out) {
    return solve_refine_impl(F, B, nrhs, X, location, tol, maxiter, 0, nullptr, out);
}

extern "C" int mxc_solve_refine_segments(mxc_factor *F, const double *B, int64_t nrhs, double *X,
                                         int location, double tol, int maxiter, int64_t nseg,
                                         const int64_t *segoff, double *out) {
    return solve_refine_impl(F, B, nrhs, X, location, tol, maxiter, nseg, segoff, out);
}

// (min diag(L) / max diag(L))^2, 0 if any diagonal entry is NaN or the
// factorization stopped early, 1 for n = 0 (CHOLMOD/Cholesky/cholmod_rcond.c:51-112,
// t_cholmod_rcond_worker.c:53-143).  The diagonal's min / max / NaN flag are
// reduced per block by k_diag_minmax into d_out (3 * RCOND_BLOCKS doubles) and
// finished on the host.
static int rcond_partials(mxc_factor *F, double *d_out, double *rcond) {
    if (F->precision == MXC_FP64)
        hipLaunchKernelGGL(k_diag_minmax<double>, dim3(RCOND_BLOCKS), dim3(256), 0, F->stream,
                           (const double *)F->W, F->d_dpos, (long long)F->n, d_out);
    else
        hipLaunchKernelGGL(k_diag_minmax<float>, dim3(RCOND_BLOCKS), dim3(256), 0, F->stream,
                           (const float *)F->W, F->d_dpos, (long long)F->n, d_out);
    double h[3 * RCOND_BLOCKS];
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(h, d_out, sizeof(h), hipMemcpyDeviceToHost, F->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(F->stream);
    if (e != hipSuccess) return set_hip_err(e, "rcond", __LINE__);
    double lo = INFINITY, hi = -INFINITY;
    for (int b = 0; b < RCOND_BLOCKS; b++) {
        if (h[3 * b + 2] != 0) { *rcond = 0; return MXC_OK; }
        lo = std::min(lo, h[3 * b]);
        hi = std::max(hi, h[3 * b + 1]);
    }
    double r = lo / hi;
    *rcond = r * r;
    return MXC_OK;
}

// The cases decided without the device: 1 for n = 0, 0 after a factorization
// that stopped early.  Returns 1 when *rcond is set.
static int rcond_trivial(const mxc_factor *F, double *rcond) {
    if (F->n == 0) { *rcond = 1; return 1; }
    if (F->minor < F->n) { *rcond = 0; return 1; }
    return 0;
}

extern "C" int mxc_rcond(mxc_factor *F, double *rcond) {
    if (!F || !rcond) return MXC_INVALID;
    F = active(F);
    if (!F->values_loaded) return MXC_INVALID;
    if (rcond_trivial(F, rcond)) return MXC_OK;
    HIPTRY(hipSetDevice(g_device));
    double *d_out = nullptr;
    {
        std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
        HIPTRY(hipMalloc((void **)&d_out, sizeof(double) * 3 * RCOND_BLOCKS));
    }
    int rc = rcond_partials(F, d_out, rcond);
    {
        std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
        (void)hipFree(d_out);
    }
    return rc;
}

// mxc_rcond without any allocation: the partials go to this GPU's `scratch`
// array (through its device address), so the call makes no hipMalloc / hipFree
// and no host allocation.  Calls on one GPU share the array and are serialized
// by g_scratch_mutex; the call is synchronous, so its contents are dead on
// return (the next call overwrites all 3 * RCOND_BLOCKS entries before reading).
extern "C" int mxc_rcond_noalloc(mxc_factor *F, double *rcond) {
    if (!F || !rcond) return MXC_INVALID;
    F = active(F);
    if (!F->values_loaded) return MXC_INVALID;
    if (rcond_trivial(F, rcond)) return MXC_OK;
    static_assert(3 * RCOND_BLOCKS * sizeof(double) == sizeof(scratch), "scratch holds one partial per block");
    if (g_device < 0 || g_device >= MAX_DEVICES || !g_scratch[g_device]) {
        snprintf(g_err, sizeof(g_err),
                 "mxc_rcond_noalloc: no scratch address for device %d (call mxc_device_init first; "
                 "devices 0..%d are supported)", g_device, MAX_DEVICES - 1);
        return MXC_INVALID;
    }
    HIPTRY(hipSetDevice(g_device));
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    return rcond_partials(F, g_scratch[g_device], rcond);
}


// ---------------------------------------------------------------------------
// rank-k update/downdate of the resident factor (cholmod_updown,
// CHOLMOD/Modify/cholmod_updown.c:74; paths and modfl as
// Modify/t_cholmod_updown_worker.c:166-240 and :784)
// ---------------------------------------------------------------------------
static void updown_structure(mxc_factor *F) {
    if (!F->col_super.empty() || F->n == 0) return;
    const int64_t ns = F->nsuper;
    F->col_super.assign(F->n, 0);
    F->sparent_h.assign(ns, -1);
    F->winv_base_h.assign(ns, 0);
    int64_t woff = F->xsize;                 // the same slots as the factorization's plan
    for (int64_t s = 0; s < ns; s++) {
        const int64_t k1 = F->Super[s], nscol = F->Super[s + 1] - k1;
        for (int64_t j = k1; j < k1 + nscol; j++) F->col_super[j] = s;
        F->winv_base_h[s] = woff;
        for (int64_t c0 = 0; c0 < nscol; c0 += NB) {
            const int64_t nb = std::min<int64_t>(NB, nscol - c0);
            woff += nb * nb;
        }
    }
    for (int64_t s = 0; s < ns; s++) {
        const int64_t nscol = F->Super[s + 1] - F->Super[s], nsrow = F->Lpi[s + 1] - F->Lpi[s];
        if (nsrow > nscol) F->sparent_h[s] = F->col_super[F->Ls_host[F->Lpi[s] + nscol]];
    }
}

// sum over columns jl = st .. nscol-1 of supernode s of (6 + 4 |L(:, k1 + jl)|)
static double updown_fl(const mxc_factor *F, int64_t s, int64_t st) {
    const double nscol = (double)(F->Super[s + 1] - F->Super[s]), nsrow = (double)(F->Lpi[s + 1] - F->Lpi[s]);
    const double m = nscol - (double)st;
    return 6.0 * m + 4.0 * (m * nsrow - ((double)st + nscol - 1.0) * m / 2.0);
}

struct UdLaunch {
    int64_t b0, nb, r0, nr;
};

template <typename T>
static int updown_run(mxc_factor *F, int update, int64_t ncol, const int64_t *Cp, const int64_t *Ci,
                      const double *Cx, double dbound, double *modfl, int64_t *fail, int64_t *nbound) {
    const int64_t n = F->n, ns = F->nsuper;
    hipStream_t sm = F->stream;
    {
        std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
        int rc;
        if (!F->d_mw && (rc = falloc(F, &F->d_mw, sizeof(T) * (size_t)n * UD_RANK))) return rc;
        if (!F->d_mpar && (rc = falloc(F, &F->d_mpar, sizeof(T) * (size_t)UD_RANK * NB * UD_PAR))) return rc;
        if (!F->d_malpha && (rc = falloc(F, &F->d_malpha, sizeof(T) * UD_RANK))) return rc;
        if (!F->d_mflag && (rc = falloc(F, (void **)&F->d_mflag, sizeof(int) * 2))) return rc;
    }
    T *W = (T *)F->W, *Wc = (T *)F->d_mw, *par = (T *)F->d_mpar, *alpha = (T *)F->d_malpha;
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, sm, F->d_mflag, INT_MAX);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, sm, F->d_mflag + 1, 0);
    double fl = 0;
    std::vector<int64_t> task_of(ns, -1);
    std::map<long long, DiagTask> inv;          // modified blocks (by inverse slot), inverted at the end
    for (int64_t r0 = 0; r0 < ncol; r0 += UD_RANK) {
        const int k = (int)std::min<int64_t>(UD_RANK, ncol - r0);
        // the supernodes on the paths, with each rank's first column in them
        std::vector<int64_t> ts;
        std::vector<std::array<int, UD_RANK>> starts;
        std::vector<long long> sidx;
        std::vector<double> sval;
        for (int r = 0; r < k; r++) {
            const int64_t c = r0 + r;
            const bool numeric = Cp[c + 1] > Cp[c];
            const int64_t j = numeric ? Ci[Cp[c]] : n - 1;
            for (int64_t p = Cp[c]; p < Cp[c + 1]; p++) {
                if (p + 1 < Cp[c + 1] && Ci[p + 1] == Ci[p]) continue;   // a later entry of the row wins
                sidx.push_back((long long)Ci[p] * UD_RANK + r);
                sval.push_back(Cx[p]);
            }
            int64_t s = F->col_super[j], st = j - F->Super[s];
            while (s >= 0) {
                fl += updown_fl(F, s, st);
                if (numeric) {
                    if (task_of[s] < 0) {
                        task_of[s] = (int64_t)ts.size();
                        ts.push_back(s);
                        std::array<int, UD_RANK> a;
                        a.fill(INT_MAX);
                        starts.push_back(a);
                    }
                    starts[task_of[s]][r] = (int)st;
                }
                // the path leaves s at its first row below, which may lie inside the parent
                const int64_t nscol = F->Super[s + 1] - F->Super[s], nsrow = F->Lpi[s + 1] - F->Lpi[s];
                if (nsrow == nscol) break;
                const int64_t pj = F->Ls_host[F->Lpi[s] + nscol];
                s = F->sparent_h[s];
                st = pj - F->Super[s];
            }
        }
        for (int64_t s : ts) task_of[s] = -1;
        if (ts.empty()) continue;
        // blocks by level, then by 64-column step within the level
        std::vector<int64_t> order(ts.size());
        for (size_t q = 0; q < ts.size(); q++) order[q] = (int64_t)q;
        std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
            const int la = F->level_of[ts[a]], lb = F->level_of[ts[b]];
            return la != lb ? la < lb : ts[a] < ts[b];
        });
        std::vector<UdBlock> blocks;
        std::vector<UdRows> rows;
        std::vector<UdLaunch> launches;
        for (size_t q0 = 0; q0 < order.size();) {
            size_t q1 = q0;
            while (q1 < order.size() && F->level_of[ts[order[q1]]] == F->level_of[ts[order[q0]]]) q1++;
            int64_t steps = 0;
            std::vector<int64_t> fb(q1 - q0), nblk(q1 - q0);
            for (size_t q = q0; q < q1; q++) {
                const int64_t s = ts[order[q]], nscol = F->Super[s + 1] - F->Super[s];
                int first = INT_MAX;
                for (int r = 0; r < UD_RANK; r++) first = std::min(first, starts[order[q]][r]);
                fb[q - q0] = first / NB;
                nblk[q - q0] = ceil_div(nscol, NB);
                steps = std::max(steps, nblk[q - q0] - fb[q - q0]);
            }
            for (int64_t step = 0; step < steps; step++) {
                UdLaunch L{(int64_t)blocks.size(), 0, (int64_t)rows.size(), 0};
                int slot = 0;
                for (size_t q = q0; q < q1; q++) {
                    const int64_t t = fb[q - q0] + step;
                    if (t >= nblk[q - q0]) continue;
                    const int64_t s = ts[order[q]], k1 = F->Super[s], nscol = F->Super[s + 1] - k1;
                    const int64_t nsrow = F->Lpi[s + 1] - F->Lpi[s], c0 = t * NB;
                    UdBlock ub;
                    memset(&ub, 0, sizeof(ub));
                    ub.lx = F->Lpx[s];
                    ub.winv = F->winv_base_h[s] + c0 * NB;      // every earlier block is 64 wide
                    ub.nsrow = (int)nsrow;
                    ub.nscol = (int)nscol;
                    ub.k1 = (int)k1;
                    ub.psi = (int)F->Lpi[s];
                    ub.c0 = (int)c0;
                    ub.nb = (int)std::min<int64_t>(NB, nscol - c0);
                    for (int r = 0; r < UD_RANK; r++) ub.start[r] = starts[order[q]][r];
                    ub.slot = slot++;
                    blocks.push_back(ub);
                    DiagTask dt;
                    dt.off = ub.lx + c0 * nsrow + c0;
                    dt.winv = ub.winv;
                    dt.ld = ub.nsrow;
                    dt.nb = ub.nb;
                    dt.col0 = (int)(k1 + c0);
                    dt.pad = 0;
                    inv[dt.winv] = dt;
                    for (int64_t rr = c0 + ub.nb; rr < nsrow; rr += 64)
                        rows.push_back(UdRows{(int)(blocks.size() - 1), (int)rr});
                }
                L.nb = (int64_t)blocks.size() - L.b0;
                L.nr = (int64_t)rows.size() - L.r0;
                launches.push_back(L);
            }
            q0 = q1;
        }
        // one upload: blocks, row ranges, C's entries (positions, values)
        const size_t ob = 0, orr = ob + blocks.size() * sizeof(UdBlock);
        const size_t oi = (orr + rows.size() * sizeof(UdRows) + 255) / 256 * 256;
        const size_t ov = oi + sidx.size() * sizeof(long long);
        const size_t need = ov + sval.size() * sizeof(double) + 8;
        if (need > F->mdesc_cap) {
            std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
            HIPTRY(hipStreamSynchronize(sm));
            ffree(F, F->d_mdesc);
            F->d_mdesc = nullptr;
            F->mdesc_cap = 0;
            const size_t cap = std::max<size_t>(need, 1 << 16) * 3 / 2;
            int rc = falloc(F, &F->d_mdesc, cap);
            if (rc) return rc;
            F->mdesc_cap = cap;
        }
        char *dm = (char *)F->d_mdesc;
        HIPTRY(hipMemcpyAsync(dm + ob, blocks.data(), blocks.size() * sizeof(UdBlock), hipMemcpyHostToDevice, sm));
        if (!rows.empty())
            HIPTRY(hipMemcpyAsync(dm + orr, rows.data(), rows.size() * sizeof(UdRows), hipMemcpyHostToDevice, sm));
        if (!sidx.empty()) {
            HIPTRY(hipMemcpyAsync(dm + oi, sidx.data(), sidx.size() * sizeof(long long), hipMemcpyHostToDevice, sm));
            HIPTRY(hipMemcpyAsync(dm + ov, sval.data(), sval.size() * sizeof(double), hipMemcpyHostToDevice, sm));
        }
        const UdBlock *d_blocks = (const UdBlock *)(dm + ob);
        const UdRows *d_rows = (const UdRows *)(dm + orr);
        hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n * UD_RANK, 256)), dim3(256), 0, sm, Wc,
                           (long long)(n * UD_RANK), (T)0);
        hipLaunchKernelGGL(k_fill<T>, dim3(1), dim3(256), 0, sm, alpha, (long long)UD_RANK, (T)1);
        if (!sidx.empty())
            hipLaunchKernelGGL(k_updown_scatter<T>, dim3(grid_for((int64_t)sidx.size(), 256)), dim3(256), 0, sm,
                               (const long long *)(dm + oi), (const double *)(dm + ov), (long long)sidx.size(), Wc);
        for (const UdLaunch &L : launches) {
            if (L.nb > 0)
                hipLaunchKernelGGL(k_updown_block<T>, dim3((unsigned)L.nb), dim3(64), 0, sm, d_blocks + L.b0, W, Wc,
                                   alpha, par, update, dbound, F->d_mflag);
            if (L.nr > 0)
                hipLaunchKernelGGL(k_updown_rows<T>, dim3((unsigned)L.nr), dim3(64), 0, sm, d_blocks,
                                   d_rows + L.r0, W, (const int *)F->d_Ls, Wc, (const T *)par);
        }
        HIPTRY(hipGetLastError());
        HIPTRY(hipStreamSynchronize(sm));     // the host arrays above are read by the queued copies
    }
    if (!inv.empty()) {
        std::vector<DiagTask> tasks;
        tasks.reserve(inv.size());
        for (const auto &kv : inv) tasks.push_back(kv.second);
        const size_t need = tasks.size() * sizeof(DiagTask);
        if (need > F->mdesc_cap) {
            std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
            ffree(F, F->d_mdesc);
            F->d_mdesc = nullptr;
            F->mdesc_cap = 0;
            const size_t cap = std::max<size_t>(need, 1 << 16) * 3 / 2;
            int rc = falloc(F, &F->d_mdesc, cap);
            if (rc) return rc;
            F->mdesc_cap = cap;
        }
        HIPTRY(hipMemcpyAsync(F->d_mdesc, tasks.data(), need, hipMemcpyHostToDevice, sm));
        hipLaunchKernelGGL(k_updown_inv<T>, dim3((unsigned)tasks.size()), dim3(64), 0, sm,
                           (const DiagTask *)F->d_mdesc, W);
        HIPTRY(hipGetLastError());
        HIPTRY(hipStreamSynchronize(sm));
    }
    int flags[2] = {INT_MAX, 0};
    HIPTRY(hipMemcpyAsync(flags, F->d_mflag, sizeof(flags), hipMemcpyDeviceToHost, sm));
    HIPTRY(hipStreamSynchronize(sm));
    *modfl = fl;
    *fail = flags[0] == INT_MAX ? n : flags[0];
    *nbound = flags[1];
    return MXC_OK;
}

extern "C" int mxc_updown(mxc_factor *F, int update, int64_t ncol, const int64_t *Cp, const int64_t *Ci,
                          const double *Cx, double dbound, double *modfl, int64_t *info) {
    if (modfl) *modfl = 0;
    if (info) info[0] = info[1] = 0;
    if (!F || ncol < 0 || (ncol > 0 && !Cp)) return MXC_INVALID;
    const int64_t n = F->n;
    if (ncol > 0 && (Cp[0] != 0 || (Cp[ncol] > 0 && (!Ci || !Cx)))) return MXC_INVALID;
    if (!F->values_loaded) {
        snprintf(g_err, sizeof(g_err), "updown: the factor holds no numeric values");
        return MXC_INVALID;
    }
    mxc_factor *G = active(F);
    updown_structure(G);
    for (int64_t c = 0; c < ncol; c++) {
        if (Cp[c + 1] < Cp[c]) return MXC_INVALID;
        if (Cp[c + 1] == Cp[c]) continue;
        const int64_t k = Ci[Cp[c]];
        if (k < 0 || k >= n) return MXC_INVALID;
        const int64_t s = G->col_super[k], k2 = G->Super[s + 1];
        int64_t q = G->Lpi[s] + (k2 - G->Super[s]);
        const int64_t qend = G->Lpi[s + 1];
        for (int64_t p = Cp[c]; p < Cp[c + 1]; p++) {
            const int64_t i = Ci[p];
            if (i < 0 || i >= n || (p > Cp[c] && i < Ci[p - 1])) {
                snprintf(g_err, sizeof(g_err), "updown: column %lld of C is not sorted", (long long)c);
                return MXC_INVALID;
            }
            if (i < k2) continue;
            while (q < qend && G->Ls_host[q] < i) q++;
            if (q == qend || G->Ls_host[q] != i) {
                snprintf(g_err, sizeof(g_err),
                         "updown: C(%lld, %lld) lies outside the pattern of L(:, %lld); the modification "
                         "needs new fill (analyze A + CC' instead)", (long long)i, (long long)c, (long long)k);
                return MXC_INVALID;
            }
        }
    }
    if (ncol == 0 || n == 0) return MXC_OK;
    HIPTRY(hipSetDevice(g_device));
    if (F->fallback && !F->escalated) {          // an unused FP64 twin no longer matches
        std::lock_guard<std::recursive_mutex> lk(g_rt_mutex);
        mxc_factor_free(F->fallback);
        F->fallback = nullptr;
    }
    double fl = 0;
    int64_t fail = n, nbound = 0;
    int rc = G->precision == MXC_FP64
                 ? updown_run<double>(G, update, ncol, Cp, Ci, Cx, dbound, &fl, &fail, &nbound)
                 : updown_run<float>(G, update, ncol, Cp, Ci, Cx, dbound, &fl, &fail, &nbound);
    if (rc) return rc;
    F->modified = G->modified = true;
    if (modfl) *modfl = fl;
    if (info) {
        info[0] = fail;
        info[1] = nbound;
    }
    if (fail < n) {
        F->values_loaded = G->values_loaded = false;
        F->minor = G->minor = fail;
        snprintf(g_err, sizeof(g_err), "updown: the new pivot of column %lld is not positive; the factor is "
                                       "no longer valid", (long long)fail);
        return MXC_NOT_POSDEF;
    }
    return MXC_OK;
}

extern "C" int mxc_synchronize(mxc_factor *F) {
    if (!F) return MXC_INVALID;
    HIPTRY(hipStreamSynchronize(F->stream));
    return MXC_OK;
}
