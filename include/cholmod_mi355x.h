/* cholmod_mi355x.h -- C ABI of the MI355X supernodal Cholesky numeric path.
 *
 * A drop-in for the numeric half of CHOLMOD's supernodal LL' path: the
 * symbolic analysis stays on the CPU (bit-identical Super/Lpi/Lpx/Ls), while
 * the numeric factorization, the triangular solves and FP64 iterative
 * refinement run on one MI355X with the whole factor resident in HBM.
 *
 * Plain C types only: int64_t indices, double values, opaque handles.
 * Every function returns one of the status codes below, which carry the same
 * values as CHOLMOD's Common->status (CHOLMOD/Include/cholmod.h:833-840), so a
 * caller can store them in cholmod_common.status unchanged.
 *
 * Reference interfaces replaced (paths relative to the SuiteSparse root):
 *   mxc_analyze            cholmod_l_analyze_p   CHOLMOD/Include/cholmod.h:2734
 *                          (CHOLMOD/Cholesky/cholmod_analyze.c:384-909: GIVEN, NATURAL,
 *                          AMD, METIS, and the default GIVEN -> AMD -> METIS strategy)
 *   mxc_amd_order          cholmod_l_amd       CHOLMOD/Cholesky/cholmod_amd.c:40-194
 *   mxc_metis_order        cholmod_l_metis     CHOLMOD/Partition/cholmod_metis.c:557-846
 *                          (METIS_NodeND, SuiteSparse_metis/libmetis/ometis.c:43)
 *   mxc_symbolic_import    the supernodal fields of cholmod_factor, cholmod.h:1584-1594
 *   mxc_device_init        cholmod_l_gpu_start / cholmod_l_gpu_allocate  cholmod.h:4148-4180
 *                          (CHOLMOD/GPU/cholmod_gpu.c:234, :426)
 *   mxc_factor_create      the per-analysis device set-up done by gpu_init
 *                          (CHOLMOD/GPU/t_cholmod_gpu.c:84-160) and the relative maps of
 *                          kernelCreateMap / kernelCreateRelativeMap
 *                          (CHOLMOD/GPU/cholmod_gpu_kernels.cu:21-41)
 *   mxc_factorize          cholmod_l_factorize_p cholmod.h:2794 (permutation,
 *                          CHOLMOD/Cholesky/cholmod_factorize.c:220-250) +
 *                          cholmod_l_super_numeric cholmod.h:4026
 *                          (CHOLMOD/Supernodal/t_cholmod_super_numeric_worker.c:335-1279)
 *   mxc_factor_get_Lx      L->x on return from cholmod_l_factorize (cholmod.h:1584-1594)
 *   mxc_solve              cholmod_l_solve(CHOLMOD_A, ...) cholmod.h:2823, supernodal branch
 *                          CHOLMOD/Cholesky/cholmod_solve.c:656-724
 *   mxc_super_lsolve       cholmod_l_super_lsolve  cholmod.h:4047
 *   mxc_super_ltsolve      cholmod_l_super_ltsolve cholmod.h:4068
 *   mxc_solve_bset         cholmod_l_solve2 cholmod.h:2845 with a sparse right-hand side
 *                          (Bset / Xset, CHOLMOD/Cholesky/cholmod_solve.c:290-640)
 *   mxc_lsolve_pattern     cholmod_l_lsolve_pattern (CHOLMOD/Cholesky/cholmod_rowfac.c:383)
 *   mxc_residual           cholmod_l_sdmult(A, 0, -1, 1, X, B) used for r = b - A x
 *                          (CHOLMOD/MatrixOps/cholmod_sdmult.c:62)
 *   mxc_solve_refine       the iterative-refinement loop of CHOLMOD/Demo/cholmod_dl_demo.c:627-663,
 *                          extended to a low-precision factor (no single routine in the reference)
 */
#ifndef CHOLMOD_MI355X_H
#define CHOLMOD_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: same values as CHOLMOD_OK, CHOLMOD_NOT_INSTALLED, ... */
#define MXC_OK              0
#define MXC_NOT_INSTALLED (-1)
#define MXC_OUT_OF_MEMORY (-2)
#define MXC_TOO_LARGE     (-3)
#define MXC_INVALID       (-4)
#define MXC_GPU_PROBLEM   (-5)
#define MXC_NOT_POSDEF      1

/* orderings (values of CHOLMOD_NATURAL, CHOLMOD_GIVEN, CHOLMOD_AMD, CHOLMOD_METIS) */
#define MXC_ORDER_NATURAL 0
#define MXC_ORDER_GIVEN   1
#define MXC_ORDER_AMD     2
#define MXC_ORDER_METIS   3
/* CHOLMOD's default strategy (Common->nmethods = 0): GIVEN if a UserPerm is
 * passed, then AMD, then METIS unless AMD already predicts fewer than 500
 * flops per entry of L or fill below 5x nnz(A); the smallest nnz(L) wins */
#define MXC_ORDER_DEFAULT (-1)

/* working precision of the numeric factor */
#define MXC_FP64 0
#define MXC_FP32 1

/* where a value/vector pointer lives */
#define MXC_HOST   0
#define MXC_DEVICE 1

typedef struct mxc_symbolic mxc_symbolic;   /* CPU symbolic factor */
typedef struct mxc_factor mxc_factor;       /* device plan + numeric factor */
typedef struct mxc_pool mxc_pool;           /* device memory pool (cholmod_common.dev_mempool) */

/* ---- symbolic analysis (CPU) ------------------------------------------- */

/* Analyze tril(A) (n x n, CSC, Ap[n+1], Ai[Ap[n]], rows i >= j) under an
 * ordering (MXC_ORDER_*; UserPerm is used by GIVEN and, when not NULL, by
 * DEFAULT).  Produces Perm, ColCount, Super, Lpi, Lpx, Ls exactly as
 * cholmod_l_analyze_p does with default parameters and supernodal = SUPERNODAL
 * (with nmethods = 1 and that ordering, or nmethods = 0 for DEFAULT). */
int mxc_analyze(int64_t n, const int64_t *Ap, const int64_t *Ai, int ordering,
                const int64_t *UserPerm, mxc_symbolic **S);

/* info[0..6] = n, nsuper, ssize, xsize, maxcsize, maxesize, ordering;
 * dinfo[0..1] = fl (Common->fl), lnz (Common->lnz). */
int mxc_symbolic_info(const mxc_symbolic *S, int64_t *info, double *dinfo);

/* What the ordering strategy left in Common: selected = index of the chosen
 * method (0 GIVEN, 1 AMD, 2 METIS under DEFAULT; 0 for a single ordering),
 * method_fl[3] / method_lnz[3] = Common->method[0..2].fl / lnz (AMD's are its
 * own predictions; -1 where a method was not tried), anz = Common->anz
 * (-1 if no ordering computed it).  Any output pointer may be NULL. */
int mxc_symbolic_strategy(const mxc_symbolic *S, int64_t *selected, double *method_fl,
                          double *method_lnz, double *anz);

/* Copy out Perm[n], ColCount[n], Super[nsuper+1], Lpi[nsuper+1], Lpx[nsuper+1],
 * Ls[ssize] (any pointer may be NULL). */
int mxc_symbolic_export(const mxc_symbolic *S, int64_t *Perm, int64_t *ColCount,
                        int64_t *Super, int64_t *Lpi, int64_t *Lpx, int64_t *Ls);

/* Wrap an existing supernodal symbolic factor (e.g. from cholmod_l_analyze). */
int mxc_symbolic_import(int64_t n, int64_t nsuper, const int64_t *Perm, const int64_t *Super,
                        const int64_t *Lpi, const int64_t *Lpx, const int64_t *Ls,
                        mxc_symbolic **S);

void mxc_symbolic_free(mxc_symbolic *S);

/* Approximate minimum degree ordering of tril(A) as used by cholmod_l_analyze
 * with CHOLMOD_AMD: Perm[k] = original index of the k-th pivot. */
int mxc_amd_order(int64_t n, const int64_t *Ap, const int64_t *Ai, int64_t *Perm);

/* The same, plus the statistics cholmod_l_amd leaves in Common:
 * stats[0] = fl, stats[1] = lnz (AMD's predictions), stats[2] = anz. */
int mxc_amd_order_stats(int64_t n, const int64_t *Ap, const int64_t *Ai, int64_t *Perm,
                        double *stats);

/* Nested-dissection ordering of tril(A) exactly as cholmod_l_metis computes it
 * inside cholmod_l_analyze (METIS_NodeND with default options on A + A',
 * no postorder): Perm[k] = original index of the k-th pivot.  nswitch and
 * dswitch are Common->metis_nswitch / metis_dswitch (defaults 3000, 0.66):
 * a matrix with n > nswitch and density > dswitch gets the identity. */
int mxc_metis_order(int64_t n, const int64_t *Ap, const int64_t *Ai, double nswitch, double dswitch,
                    int64_t *Perm);

/* ---- device ------------------------------------------------------------- */

/* Select the HIP device used by subsequent factor objects. */
int mxc_device_init(int device);

/* info[0..3] = compute units, HBM bytes, gcn arch number (950), clock kHz */
int mxc_device_info(int64_t *info);

/* Free and total device memory of the current device (cholmod_l_gpu_memorysize,
 * CHOLMOD/GPU/cholmod_gpu.c:63-110). */
int mxc_memory_info(int64_t *free_bytes, int64_t *total_bytes);

/* Device memory pool: one allocation of `bytes` on the current device, from
 * which factors created with mxc_factor_create_in take all their device
 * memory (first fit, coalescing frees).  Replaces the Common-owned
 * dev_mempool of cholmod_l_gpu_allocate (CHOLMOD/GPU/cholmod_gpu.c:426-543) and
 * its release in cholmod_l_gpu_deallocate (:234-260).  A factor that does not
 * fit gets MXC_OUT_OF_MEMORY.  info[0..3] = capacity, bytes in use, peak bytes
 * in use, live allocations.  mxc_pool_free returns MXC_INVALID (and frees
 * nothing) while factors still hold memory from the pool. */
int mxc_pool_create(size_t bytes, mxc_pool **P);
int mxc_pool_info(const mxc_pool *P, int64_t *info);
int mxc_pool_free(mxc_pool *P);

/* ---- numeric factorization (device) ------------------------------------ */

/* Build the device plan for S and the pattern of tril(A) (the pattern the
 * values passed to mxc_factorize follow).  precision: MXC_FP64 or MXC_FP32. */
int mxc_factor_create(const mxc_symbolic *S, const int64_t *Ap, const int64_t *Ai,
                      int precision, mxc_factor **F);

/* As mxc_factor_create, with every device buffer of the factor (and of its
 * FP64 fallback, if one is ever built) taken from pool P (NULL: hipMalloc). */
int mxc_factor_create_in(const mxc_symbolic *S, const int64_t *Ap, const int64_t *Ai,
                         int precision, mxc_pool *P, mxc_factor **F);

void mxc_factor_free(mxc_factor *F);

/* info[0..9] = n, nsuper, xsize, levels, update pairs, update tiles,
 *              device bytes held, precision, diagonal blocks the last
 *              factorization refactored in FP64 (see MXC_OPT_PIVOT_RHO_E9), 1 if
 *              the last factorization escalated to FP64 (see mxc_factorize);
 * dinfo[0..5] = fl, lnz, and the algorithmic flops of one factorization by
 *               kernel class: descendant updates (dsyrk + dgemm counts), diagonal
 *               POTRF, panel TRSM, in-supernode trailing updates */
int mxc_factor_info(const mxc_factor *F, int64_t *info, double *dinfo);

/* Numeric factorization of tril(A) + beta*I with values Ax (pattern of
 * create, unpermuted).  Ax is a host or device pointer (location).  On
 * return *minor = n on success, else the first failing column, with
 * MXC_NOT_POSDEF returned (L then holds the factor of the leading minor
 * columns and zeros after, as CHOLMOD leaves it).  A low-precision factor
 * (MXC_FP32 with a pivot ratio rho > 0) that meets a pivot <= 0 escalates:
 * the same values are factored in FP64, whose result decides the status and
 * minor, and the FP64 factor serves every solve, L copy and rcond until the
 * next mxc_factorize (mxc_factor_info info[9] = 1). */
int mxc_factorize(mxc_factor *F, const double *Ax, int location, double beta, int64_t *minor);

/* Copy the factor to host memory as double, in CHOLMOD's L->x layout. */
int mxc_factor_get_Lx(const mxc_factor *F, double *Lx);

/* X = A \ B (sys = CHOLMOD_A): B, X are n x nrhs column-major (ld n), host or
 * device (both at location).  X may alias B. */
int mxc_solve(mxc_factor *F, const double *B, int64_t nrhs, double *X, int location);

/* In-place L \ X and L' \ X on vectors already in the permuted order
 * (cholmod_l_super_lsolve / ltsolve). */
int mxc_super_lsolve(mxc_factor *F, double *X, int64_t nrhs, int location);
int mxc_super_ltsolve(mxc_factor *F, double *X, int64_t nrhs, int location);

/* System codes of cholmod_l_solve / solve2 (cholmod.h:2805-2813). */
#define MXC_SYS_A 0
#define MXC_SYS_LDLt 1
#define MXC_SYS_LD 2
#define MXC_SYS_DLt 3
#define MXC_SYS_L 4
#define MXC_SYS_Lt 5
#define MXC_SYS_D 6

/* cholmod_l_solve2 (cholmod.h:2845) with Bset: one right-hand side b (dense,
 * n entries, host) of which only b[Bset[0..blen)] is read.  sys is
 * MXC_SYS_A .. MXC_SYS_D (an LL' factor: D = I; CHOLMOD ignores Bset for
 * P and P').  On return Xset[0..*xlen) is the pattern the reference returns
 * in Xset, in its order (the columns of L reached from Bset, permuted back
 * for MXC_SYS_A), and X (host, n entries) holds the solution at those rows;
 * its other entries are not written.  Xset must hold n entries.  L stays
 * supernodal (the reference converts it to simplicial; the pattern and
 * values returned are the same). */
int mxc_solve_bset(mxc_factor *F, int sys, const double *B, const int64_t *Bset, int64_t blen,
                   double *X, int64_t *Xset, int64_t *xlen);

/* The pattern alone (host only, no device): Xset[0..*xlen) as mxc_solve_bset
 * returns it for these sys and Bset -- cholmod_l_lsolve_pattern
 * (CHOLMOD/Cholesky/cholmod_rowfac.c:383-398) on the simplicial form of L,
 * permuted as cholmod_l_solve2 permutes it.  Xset must hold n entries. */
int mxc_lsolve_pattern(const mxc_symbolic *S, int sys, const int64_t *Bset, int64_t blen, int64_t *Xset,
                       int64_t *xlen);

/* Reciprocal condition estimate of the last factorization, as
 * cholmod_l_rcond (CHOLMOD/Cholesky/cholmod_rcond.c:51-112) computes it for an
 * LL' factor: (min diag(L) / max diag(L))^2; 0 if the factorization stopped
 * early (minor < n) or a diagonal entry is NaN; 1 when n = 0. */
int mxc_rcond(mxc_factor *F, double *rcond);

/* mxc_rcond for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops): no counterpart in the reference.  Same result, bit
 * for bit, but it calls no hipMalloc / hipFree: the per-block partials
 * (3 x 256 doubles, independent of n) live in a static device array, one per
 * GPU, set up by mxc_device_init (which must have been called for the current
 * device; else MXC_INVALID).  Calls on one GPU share that array and take turns
 * (a process-wide lock; the call is synchronous). */
int mxc_rcond_noalloc(mxc_factor *F, double *rcond);

/* Rank-k update or downdate of the resident factor in place:
 *   Lnew Lnew' = L L' + C C' (update != 0) or L L' - C C' (update == 0),
 * replacing cholmod_l_updown (CHOLMOD/Include/cholmod.h:3465,
 * CHOLMOD/Modify/cholmod_updown.c:74).  C is n x ncol, CSC (Cp[ncol+1],
 * Ci, Cx on the host, rows sorted within each column, a later entry of a
 * repeated row winning), its rows in the factor's order: L->Perm is not
 * applied, as in the reference.  Each entry of C(:, c) must lie in the
 * pattern of L(:, k), k the column's first row; otherwise the modification
 * needs new fill and the call returns MXC_INVALID with L unchanged (the
 * reference adds the fill to a simplicial factor).  Columns are taken 8 at a
 * time, as the reference's maxrank.  *modfl = the reference's
 * Common->modfl; info[0] = the first column whose new pivot is not positive
 * (n if none), info[1] = pivots raised to dbound (dbound > 0, as
 * cholmod_dbound; 0 = off).  L stays a supernodal LL' factor on the device
 * (the reference turns it into a simplicial LDL' one), and its diagonal-block
 * inverses are renewed for the solves.  A downdate that loses positive
 * definiteness returns MXC_NOT_POSDEF and leaves the factor without values
 * (the reference keeps an indefinite LDL').  After a modification the stored
 * A no longer matches: mxc_solve_refine refuses until the next
 * mxc_factorize, and an FP64 twin of a low-precision factor is dropped. */
int mxc_updown(mxc_factor *F, int update, int64_t ncol, const int64_t *Cp, const int64_t *Ci, const double *Cx,
               double dbound, double *modfl, int64_t *info);

/* R = B - A X in FP64 with the values of the last mxc_factorize. */
int mxc_residual(mxc_factor *F, const double *B, const double *X, int64_t nrhs, double *R,
                 int location);

/* Solve with FP64 iterative refinement: x_0 = A\b through the factor, then
 * r = b - A x (FP64), x += A\r until ||r||_2 / ||b||_2 <= tol for every
 * column or maxiter corrections.  If a low-precision factor does not reach tol
 * (e.g. its pivots broke down: see MXC_OPT_PIVOT_RHO_E9), an FP64 factor of
 * the same matrix is built, factored and used instead.  out[0] = corrections
 * done, out[1] = final max relative residual, out[2] = initial max relative
 * residual, out[3] = 1 if the FP64 fallback factor was used. */
int mxc_solve_refine(mxc_factor *F, const double *B, int64_t nrhs, double *X, int location,
                     double tol, int maxiter, double *out);

/* Options.  MXC_OPT_GRAPHS (default 1): capture the factorization and the
 * solve sweeps into hipGraphs on first use and replay them afterwards (the
 * launch sequence depends only on the pattern); 0 enqueues every launch. */
#define MXC_OPT_GRAPHS 1
/* MXC_OPT_PIVOT_RHO_E9: pivot-ratio threshold rho in units of 1e-9 for FP32
 * factors (default rho = 1e-6; FP64 factors ignore it).  With rho > 0, a
 * 64-column (or smaller) diagonal block in which a pivot falls to
 * rho * a_jj or below (a_jj the assembled diagonal of A) is refactored in
 * FP64 on the device before anything reads it: from A's FP64 values plus
 * the descendant updates accumulated in FP32, FP64 POTRF and inverse, the
 * result stored in FP32; the factorization then continues in FP32
 * (mxc_factor_info info[8] counts such blocks).  A pivot <= 0 even in FP64
 * escalates the factorization to FP64 (see mxc_factorize).  rho = 0 keeps
 * CHOLMOD's semantics in FP32: a pivot <= 0 stops with MXC_NOT_POSDEF. */
#define MXC_OPT_PIVOT_RHO_E9 2
/* MXC_OPT_LOOKAHEAD (default 1): run the update of each finished 384-column
 * panel of a large supernode right of the next panel on a second stream, so
 * that it overlaps the next panel's diagonal-block steps.  0: one stream. */
#define MXC_OPT_LOOKAHEAD 3
/* MXC_OPT_SPLIT16 (default 1, FP32 factors only): the descendant-update and
 * trailing tiles of 64 and 128 columns compute each product as
 * hi*hi + hi*lo + lo*hi of FP16 halves on the FP16 MFMA, FP32 accumulation,
 * after a per-row power-of-two scaling by the assembled diagonal
 * (|L_ij| <= sqrt(a_ii)); the TRSM stores each value's pair once.  32-column
 * tiles, and all tiles with 0, use the FP32 MFMA. */
#define MXC_OPT_SPLIT16 4
/* MXC_OPT_SOLVE_CHAIN (default 1): the solve sweeps take one launch per etree
 * level; inside it each 64-column block of a supernode waits, through a
 * per-supernode counter in device memory, for the blocks it depends on.  A
 * wait is bounded: a sweep whose wait gives up returns MXC_GPU_PROBLEM.
 * 0: one launch per 64-column step of a level. */
#define MXC_OPT_SOLVE_CHAIN 5
/* MXC_OPT_DETERMINISTIC (default 0): bitwise-repeatable factorization.  The
 * descendant updates of a level are split into rounds, round r holding the
 * r-th (d, s) pair (by descendant, ascending) of every target supernode s, one
 * launch after the other: no two tiles of a launch add into the same entry of
 * L, so every entry receives its updates in a fixed order, d ascending (the
 * order in which the reference's left-looking loop applies them,
 * t_cholmod_super_numeric_worker.c:896-930).  In-supernode steps already have
 * one writer per entry.  Costs one launch per round: as many rounds per level
 * as the most descendants of any of its supernodes.  The solve sweeps are not
 * affected (the forward sweep adds ancestor contributions atomically). */
#define MXC_OPT_DETERMINISTIC 6
/* MXC_OPT_FOLD_TRSM (default 0, FP64 factors): a trailing launch whose tile
 * (0, 0) factors the next 64-column diagonal block also applies that block's
 * TRSM to the rows below it (the tiles of its first column wait for the
 * block's inverse), instead of a separate TRSM launch per step.  Bitwise the
 * same factor; off by default because the waiting tiles hold their CUs
 * through the diagonal factorization, which costs more than the launch. */
#define MXC_OPT_FOLD_TRSM 7
/* MXC_OPT_T2_GRID (default 3, FP64 factors): the look-ahead's second-stream
 * trailing update runs on a grid of this many workgroups per CU, each walking
 * its share of the tiles (0: one workgroup per tile).  Every CU then keeps
 * room for the main stream's diagonal factorizations, which otherwise wait
 * for the whole second-stream launch to drain. */
#define MXC_OPT_T2_GRID 8
/* MXC_OPT_SMALL_FUSED (default 0, FP64 factors): in every etree level with at
 * most this many supernodes of at most 16 columns and 64 rows that have
 * descendant updates, those supernodes are factored whole by one launch, a
 * workgroup per supernode on a persistent grid: its descendant updates (four
 * pairs at a time), diagonal factorization, inverse and TRSM, held in LDS
 * throughout (0: never; not in deterministic mode).  Off by default: every
 * threshold measured slower than the level's update, POTRF and TRSM launches
 * (a 300 x 300 grid: 2.70 ms with 64 against 2.55 ms; 4,096 2D systems with
 * every level fused: 49.8 against 31.3 ms). */
#define MXC_OPT_SMALL_FUSED 9
int mxc_set_option(mxc_factor *F, int option, int64_t value);

/* As mxc_solve_refine for a block-diagonal batch of independent systems: rows
 * segoff[g] .. segoff[g+1]-1 are system g (segoff[0] = 0, segoff[nseg] = n), and
 * refinement continues until every system's relative residual is <= tol.
 * out[1], out[2] are the max over systems and columns. */
int mxc_solve_refine_segments(mxc_factor *F, const double *B, int64_t nrhs, double *X, int location,
                              double tol, int maxiter, int64_t nseg, const int64_t *segoff,
                              double *out);

/* Tile autotuner.  Blocks are bucketed by rows M (<=16, <=32, <=64, <=128,
 * >128), columns N (<=8, <=16, <=32, <=64, >64) and inner dimension K (<=32,
 * <=128, >128): bucket (5*mb + nb)*3 + kb, one table of 75 buckets for each
 * kind of block (0 descendant update, 1 in-supernode trailing update, 2 TRSM).
 * mxc_autotune times every tile edge (32, 64, 128) on a sample of each
 * bucket's blocks (reps launches each, HIP events), keeps the fastest, rebuilds
 * the plan and returns the table (3*75 entries, kind-major; may be NULL).  It
 * scribbles on the numeric factor: call it before mxc_factorize.
 * mxc_set_tiles installs a saved table; mxc_get_tiles reads the current one. */
int mxc_autotune(mxc_factor *F, int reps, int *table);
int mxc_set_tiles(mxc_factor *F, const int *table);
int mxc_get_tiles(const mxc_factor *F, int *table);

/* Wait for all work queued on the factor's stream. */
int mxc_synchronize(mxc_factor *F);

/* Device times (ms, from HIP events on the factor's stream) accumulated since
 * mxc_set_timing, per class: t[0] assembly, t[1] updates, t[2] diag potrf,
 * t[3] trsm, t[4] trailing, t[5] forward solve, t[6] backward solve,
 * t[7] residual, t[8] factor total, t[9] solve total.  mask: bit c times
 * class c (1 = all classes, 0 = off).  counts[0..4] = launches of update,
 * potrf, trsm, trailing, solve. */
int mxc_set_timing(mxc_factor *F, int mask);
int mxc_get_timing(const mxc_factor *F, double *t, int64_t *counts);

/* Per-factor statistics for cholmod_common's GPU fields (gpuNumKernelLaunches,
 * gpuKernelTime, cholmod_gpu_*_calls; CHOLMOD/Include/cholmod.h:921-986).
 * counts[0..7] = launches per factorization of the descendant-update,
 * diagonal-POTRF, TRSM and trailing-update classes, their sum, launches per
 * solve (both sweeps), factorizations and solves done so far;
 * ms[0..3] = device time (HIP events on the factor's stream) of the last
 * factorization and the last solve call, and their sums so far. */
int mxc_factor_stats(const mxc_factor *F, int64_t *counts, double *ms);

/* Last HIP error text seen by the library (empty when none). */
const char *mxc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
