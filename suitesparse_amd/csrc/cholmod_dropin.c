/* cholmod_dropin.c -- CHOLMOD's int64 analyze / factorize / solve entry points,
 * with CHOLMOD's own signatures, structs and status behaviour, running the
 * supernodal numeric factorization and the solves on one MI355X through the
 * C ABI of libcholmod_mi355x.so (include/cholmod_mi355x.h).
 *
 * Built against the host CHOLMOD's cholmod.h (the struct layouts and the
 * prototypes below are that header's; a mismatch is a compile error) and
 * linked in front of the host CHOLMOD library, which keeps everything else
 * (start/finish, allocation, read, sdmult, print, ...).  The entry points
 * replaced, paths relative to the SuiteSparse root:
 *
 *   cholmod_l_analyze       CHOLMOD/Include/cholmod.h:2714  (Cholesky/cholmod_analyze.c:384-909)
 *   cholmod_l_analyze_p     cholmod.h:2734
 *   cholmod_l_factorize     cholmod.h:2775  (Cholesky/cholmod_factorize.c:93-418)
 *   cholmod_l_factorize_p   cholmod.h:2794
 *   cholmod_l_solve         cholmod.h:2823  (Cholesky/cholmod_solve.c:130-724)
 *   cholmod_l_solve2        cholmod.h:2845
 *   cholmod_l_updown        cholmod.h:3465  (Modify/cholmod_updown.c:74; on the device
 *                                            factor when C stays in L's pattern)
 *   cholmod_l_free_factor   cholmod.h:1692  (releases the factor's device plan first)
 *   cholmod_l_finish        cholmod.h:1035  (releases the Common's device pool:
 *                                            GPU/cholmod_gpu.c:234-260, gpu_deallocate)
 *   cholmod_l_gpu_stats     cholmod.h:2307  (Check/cholmod_check.c:615-674)
 *
 * What runs where.  The device takes the supernodal LL' factorization of a
 * real double symmetric A (stype -1 or +1) and every solve with such a factor,
 * as the reference's GPU module does (GPU/t_cholmod_gpu.c).  As in the
 * reference, the host CHOLMOD keeps the rest: Common->useGPU = 0 (or -1 with
 * CHOLMOD_USE_GPU unset or 0, cholmod_super_symbolic.c:250-289), unsymmetric
 * A (A*A', stype 0), single precision or complex values, simplicial analyses
 * (Common->supernodal, supernodal_switch), non-default orderings or
 * postordering, and a factor that does not fit the Common's device pool.
 * Those calls go to the next definition of the same entry point (the host
 * library), found with dlsym(RTLD_NEXT).
 *
 * Analysis is this library's own: the given or natural ordering, AMD and
 * METIS (mxc_amd_order_stats, mxc_metis_order), and CHOLMOD's default
 * strategy over them (UserPerm if given, AMD, then METIS unless AMD's
 * prediction is already good: cholmod_analyze.c:554-782), bit-identical to
 * cholmod_l_analyze_p's.  NESDIS, non-default AMD controls and postorder =
 * FALSE go to the host library.
 *
 * Device state.  Each factor L that reaches the device owns an entry in a side
 * table keyed by the cholmod_factor pointer (with L->Perm / L->super as stamps
 * against reuse of the address): the symbolic import, the device plan built
 * from A's lower-triangular pattern, and the gather map from A's storage (any
 * stype, packed or not, sorted or not) to that pattern.  All device memory of
 * the plans comes from one pool per Common (Common->dev_mempool /
 * dev_mempool_size), sized as cholmod_l_gpu_allocate sizes it
 * (maxGpuMemBytes, maxGpuMemFraction, CHOLMOD_GPU_MEM_BYTES /
 * CHOLMOD_GPU_MEM_FRACTION; all free memory less 50 MB when neither is set).
 *
 * After cholmod_l_factorize, L->x holds the factor on the host, as the API
 * promises (cholmod.h:1584-1594): the device copy is written back into L->x
 * on every factorization.  CHOLMOD_MI355X_KEEP_L_ON_DEVICE=1 skips that copy
 * (L->x is then allocated but not filled; solves through this library still
 * use the device factor).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cholmod.h"
#include "../../include/cholmod_mi355x.h"

#define DROPIN_FILE "cholmod_dropin.c"

/* ------------------------------------------------------------------------ */
/* the host library's definitions                                           */
/* ------------------------------------------------------------------------ */

typedef cholmod_factor *(*analyze_p_fn)(cholmod_sparse *, int64_t *, int64_t *, size_t, cholmod_common *);
typedef int (*factorize_p_fn)(cholmod_sparse *, double *, int64_t *, size_t, cholmod_factor *,
                              cholmod_common *);
typedef cholmod_dense *(*solve_fn)(int, cholmod_factor *, cholmod_dense *, cholmod_common *);
typedef int (*solve2_fn)(int, cholmod_factor *, cholmod_dense *, cholmod_sparse *, cholmod_dense **,
                         cholmod_sparse **, cholmod_dense **, cholmod_dense **, cholmod_common *);
typedef int (*free_factor_fn)(cholmod_factor **, cholmod_common *);
typedef int (*updown_fn)(int, cholmod_sparse *, cholmod_factor *, cholmod_common *);
typedef int (*finish_fn)(cholmod_common *);
typedef int (*gpu_stats_fn)(cholmod_common *);

static void *next_sym(const char *name) {
    void *f = dlsym(RTLD_NEXT, name);
    if (!f) {
        fprintf(stderr, "cholmod_mi355x drop-in: the host CHOLMOD library does not define %s\n", name);
        abort();
    }
    return f;
}
#define NEXT(type, name) ((type)next_sym(name))

/* ------------------------------------------------------------------------ */
/* side table: device state per factor                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    cholmod_factor *L;            /* key */
    void *stamp_perm, *stamp_super;
    cholmod_common *cm;
    int64_t n;
    /* the input pattern the gather map was built from */
    int stype, packed;
    int64_t *in_p, *in_i, *in_nz;  /* copies of A->p (n+1), A->i (nnz), A->nz (n, unpacked A) */
    int64_t in_nnz;
    /* lower-triangular pattern of the plan (sorted CSC) and gather map */
    int64_t *Cp, *Ci, *src, cnz;
    double *vals;                 /* values gathered for one factorization */
    mxc_symbolic *S;
    mxc_factor *F;
    int fresh;                    /* the device factor holds L's current values */
} entry;

static entry *g_tab = NULL;
static size_t g_ntab = 0, g_cap = 0;
static pthread_mutex_t g_tab_mu = PTHREAD_MUTEX_INITIALIZER;

static void entry_release(entry *e) {
    if (e->F) mxc_factor_free(e->F);
    if (e->S) mxc_symbolic_free(e->S);
    free(e->in_p);
    free(e->in_i);
    free(e->in_nz);
    free(e->Cp);
    free(e->Ci);
    free(e->src);
    free(e->vals);
    memset(e, 0, sizeof(*e));
}

/* caller holds g_tab_mu */
static entry *tab_find(cholmod_factor *L) {
    for (size_t k = 0; k < g_ntab; k++)
        if (g_tab[k].L == L) {
            if (g_tab[k].stamp_perm == L->Perm && g_tab[k].stamp_super == L->super) return &g_tab[k];
            entry_release(&g_tab[k]);          /* a freed factor's address, reused */
            g_tab[k] = g_tab[--g_ntab];
            return NULL;
        }
    return NULL;
}

static void tab_drop(cholmod_factor *L) {
    pthread_mutex_lock(&g_tab_mu);
    for (size_t k = 0; k < g_ntab; k++)
        if (g_tab[k].L == L) {
            entry_release(&g_tab[k]);
            g_tab[k] = g_tab[--g_ntab];
            break;
        }
    pthread_mutex_unlock(&g_tab_mu);
}

static entry *tab_add(cholmod_factor *L, cholmod_common *cm) {
    if (g_ntab == g_cap) {
        size_t cap = g_cap ? 2 * g_cap : 16;
        entry *t = (entry *)realloc(g_tab, cap * sizeof(entry));
        if (!t) return NULL;
        g_tab = t;
        g_cap = cap;
    }
    entry *e = &g_tab[g_ntab++];
    memset(e, 0, sizeof(*e));
    e->L = L;
    e->stamp_perm = L->Perm;
    e->stamp_super = L->super;
    e->cm = cm;
    e->n = (int64_t)L->n;
    return e;
}

/* ------------------------------------------------------------------------ */
/* Common: GPU request, device pool, statistics                             */
/* ------------------------------------------------------------------------ */

static pthread_mutex_t g_dev_mu = PTHREAD_MUTEX_INITIALIZER;
static int g_dev_ready = 0;

/* Common->useGPU as cholmod_super_symbolic.c:250-289 resolves it: -1 asks the
 * CHOLMOD_USE_GPU environment variable (unset or 0: no GPU), which also
 * brings CHOLMOD_GPU_MEM_BYTES / CHOLMOD_GPU_MEM_FRACTION; 1 needs a device. */
static int use_gpu(cholmod_common *cm) {
    if (cm->useGPU == -1) {
        const char *env = getenv("CHOLMOD_USE_GPU");
        if (env && atoi(env) != 0) {
            cm->useGPU = 1;
            const char *b = getenv("CHOLMOD_GPU_MEM_BYTES");
            const char *f = getenv("CHOLMOD_GPU_MEM_FRACTION");
            if (b) cm->maxGpuMemBytes = (size_t)atol(b);
            if (f) {
                double x = atof(f);
                cm->maxGpuMemFraction = x < 0 ? 0 : x > 1 ? 1 : x;
            }
        } else {
            cm->useGPU = 0;
        }
    }
    if (cm->useGPU != 1) return 0;
    pthread_mutex_lock(&g_dev_mu);
    if (!g_dev_ready) g_dev_ready = mxc_device_init(0) == MXC_OK ? 1 : -1;
    pthread_mutex_unlock(&g_dev_mu);
    if (g_dev_ready != 1) cm->useGPU = 0;      /* no device: the probe fails, as gpu_probe does */
    return cm->useGPU == 1;
}

/* The Common's pool, created on first use with cholmod_l_gpu_allocate's size
 * rule (GPU/cholmod_gpu.c:451-489). */
static mxc_pool *common_pool(cholmod_common *cm) {
    if (cm->dev_mempool) return (mxc_pool *)cm->dev_mempool;
    int64_t fdm = 0, tdm = 0;
    if (mxc_memory_info(&fdm, &tdm) != MXC_OK) return NULL;
    size_t bytes = cm->maxGpuMemBytes;
    double frac = cm->maxGpuMemFraction < 0 ? 0 : cm->maxGpuMemFraction > 1 ? 1 : cm->maxGpuMemFraction;
    size_t req;
    if (bytes == 0 && frac == 0) {
        req = (size_t)fdm > (50u << 20) ? (size_t)fdm - (50u << 20) : (size_t)fdm / 2;
    } else if (bytes > 0 && frac > 0) {
        req = bytes;
        if ((double)req > (double)tdm * frac) req = (size_t)((double)tdm * frac);
    } else if (frac > 0) {
        req = (size_t)(frac * (double)tdm);
    } else {
        req = bytes > (size_t)fdm ? (size_t)fdm : bytes;
    }
    mxc_pool *P = NULL;
    if (req == 0 || mxc_pool_create(req, &P) != MXC_OK) return NULL;
    cm->dev_mempool = P;
    cm->dev_mempool_size = req;
    cm->gpuMemorySize = (size_t)tdm;
    return P;
}

static void add_factor_stats(cholmod_common *cm, const mxc_factor *F) {
    int64_t c[8];
    double ms[4];
    if (mxc_factor_stats(F, c, ms) != MXC_OK) return;
    cm->gpuKernelTime += ms[0] / 1e3;
    cm->gpuNumKernelLaunches += (int)c[4];
    /* descendant updates (fused syrk + gemm), trailing updates, POTRF, TRSM */
    cm->cholmod_gpu_syrk_calls += (size_t)c[0];
    cm->cholmod_gpu_gemm_calls += (size_t)c[3];
    cm->cholmod_gpu_potrf_calls += (size_t)c[1];
    cm->cholmod_gpu_trsm_calls += (size_t)c[2];
}

static void add_solve_stats(cholmod_common *cm, const mxc_factor *F) {
    int64_t c[8];
    double ms[4];
    if (mxc_factor_stats(F, c, ms) != MXC_OK) return;
    cm->gpuKernelTime += ms[1] / 1e3;
    cm->gpuNumKernelLaunches += (int)c[5];
}

static int report(cholmod_common *cm, int status, const char *msg) {
    cholmod_l_error(status, DROPIN_FILE, __LINE__, msg, cm);
    return status;
}

/* ------------------------------------------------------------------------ */
/* A's lower-triangular pattern and the gather map                          */
/* ------------------------------------------------------------------------ */

static int cmp_pair(const void *a, const void *b) {
    const int64_t *x = (const int64_t *)a, *y = (const int64_t *)b;
    return x[0] < y[0] ? -1 : x[0] > y[0] ? 1 : 0;
}

/* tril of the symmetric matrix A stores (stype -1: its lower part; +1: the
 * transpose of its upper part), as sorted CSC with, per entry, the index of
 * its value in A->x.  Entries in the other triangle are ignored, as CHOLMOD
 * ignores them.  Returns 0 on success. */
static int lower_pattern(const cholmod_sparse *A, int64_t **pCp, int64_t **pCi, int64_t **psrc,
                         int64_t *pcnz) {
    const int64_t n = (int64_t)A->ncol;
    const int64_t *Ap = (const int64_t *)A->p, *Ai = (const int64_t *)A->i;
    const int64_t *Anz = (const int64_t *)A->nz;
    const int upper = A->stype > 0;
    int64_t *Cp = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    if (!Cp) return -1;
    for (int64_t j = 0; j < n; j++) {
        const int64_t p0 = Ap[j], p1 = A->packed ? Ap[j + 1] : Ap[j] + Anz[j];
        for (int64_t p = p0; p < p1; p++) {
            const int64_t i = Ai[p];
            if (i < 0 || i >= n) { free(Cp); return -2; }
            if (upper ? i <= j : i >= j) Cp[(upper ? i : j) + 1]++;
        }
    }
    for (int64_t j = 0; j < n; j++) Cp[j + 1] += Cp[j];
    const int64_t cnz = Cp[n];
    int64_t *Ci = (int64_t *)malloc((size_t)(cnz > 0 ? cnz : 1) * sizeof(int64_t));
    int64_t *src = (int64_t *)malloc((size_t)(cnz > 0 ? cnz : 1) * sizeof(int64_t));
    int64_t *w = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    if (!Ci || !src || !w) { free(Cp); free(Ci); free(src); free(w); return -1; }
    memcpy(w, Cp, (size_t)n * sizeof(int64_t));
    for (int64_t j = 0; j < n; j++) {
        const int64_t p0 = Ap[j], p1 = A->packed ? Ap[j + 1] : Ap[j] + Anz[j];
        for (int64_t p = p0; p < p1; p++) {
            const int64_t i = Ai[p];
            if (upper ? i <= j : i >= j) {
                const int64_t col = upper ? i : j, row = upper ? j : i;
                const int64_t q = w[col]++;
                Ci[q] = row;
                src[q] = p;
            }
        }
    }
    free(w);
    /* rows ascending in every column (an unsorted A; the upper case comes out sorted) */
    int64_t *pair = NULL;
    size_t pair_cap = 0;
    for (int64_t j = 0; j < n; j++) {
        const int64_t a = Cp[j], b = Cp[j + 1];
        int sorted = 1;
        for (int64_t q = a + 1; q < b && sorted; q++) sorted = Ci[q - 1] < Ci[q];
        if (sorted) continue;
        if ((size_t)(b - a) > pair_cap) {
            pair_cap = (size_t)(b - a);
            free(pair);
            pair = (int64_t *)malloc(2 * pair_cap * sizeof(int64_t));
            if (!pair) { free(Cp); free(Ci); free(src); return -1; }
        }
        for (int64_t q = a; q < b; q++) {
            pair[2 * (q - a)] = Ci[q];
            pair[2 * (q - a) + 1] = src[q];
        }
        qsort(pair, (size_t)(b - a), 2 * sizeof(int64_t), cmp_pair);
        for (int64_t q = a; q < b; q++) {
            if (q > a && pair[2 * (q - a)] == pair[2 * (q - a - 1)]) {
                free(pair); free(Cp); free(Ci); free(src);
                return -3;                         /* duplicate entry */
            }
            Ci[q] = pair[2 * (q - a)];
            src[q] = pair[2 * (q - a) + 1];
        }
    }
    free(pair);
    *pCp = Cp;
    *pCi = Ci;
    *psrc = src;
    *pcnz = cnz;
    return 0;
}

static int64_t *dup64(const void *p, int64_t len) {
    int64_t *q = (int64_t *)malloc((size_t)(len > 0 ? len : 1) * sizeof(int64_t));
    if (q && len > 0) memcpy(q, p, (size_t)len * sizeof(int64_t));
    return q;
}

static int64_t stored_nnz(const cholmod_sparse *A) {
    const int64_t n = (int64_t)A->ncol;
    const int64_t *Ap = (const int64_t *)A->p;
    return A->packed ? Ap[n] : (n > 0 ? Ap[n - 1] + ((const int64_t *)A->nz)[n - 1] : 0);
}

/* Same stored pattern as the one e's gather map came from. */
static int same_input(const entry *e, const cholmod_sparse *A) {
    const int64_t n = (int64_t)A->ncol;
    if (e->stype != A->stype || e->packed != A->packed || !e->in_p) return 0;
    const int64_t len = A->packed ? ((const int64_t *)A->p)[n] : (int64_t)A->nzmax;
    if (len != e->in_nnz) return 0;
    if (memcmp(e->in_p, A->p, (size_t)(n + 1) * sizeof(int64_t))) return 0;
    if (!A->packed && memcmp(e->in_nz, A->nz, (size_t)n * sizeof(int64_t))) return 0;
    return memcmp(e->in_i, A->i, (size_t)len * sizeof(int64_t)) == 0;
}

/* The input-pattern part of e for A; the plan is rebuilt only when the lower
 * pattern itself changes.  Returns a CHOLMOD status. */
static int bind_input(entry *e, const cholmod_sparse *A) {
    if (same_input(e, A)) return CHOLMOD_OK;
    int64_t *Cp, *Ci, *src, cnz;
    int rc = lower_pattern(A, &Cp, &Ci, &src, &cnz);
    if (rc == -1) return CHOLMOD_OUT_OF_MEMORY;
    if (rc) return CHOLMOD_INVALID;
    const int64_t n = (int64_t)A->ncol;
    const int same_plan = e->Cp && e->cnz == cnz && !memcmp(e->Cp, Cp, (size_t)(n + 1) * sizeof(int64_t)) &&
                          !memcmp(e->Ci, Ci, (size_t)(cnz > 0 ? cnz : 0) * sizeof(int64_t));
    if (!same_plan && e->F) {
        mxc_factor_free(e->F);
        e->F = NULL;
        e->fresh = 0;
    }
    free(e->Cp); free(e->Ci); free(e->src); free(e->vals);
    free(e->in_p); free(e->in_i); free(e->in_nz);
    e->Cp = Cp;
    e->Ci = Ci;
    e->src = src;
    e->cnz = cnz;
    e->vals = (double *)malloc((size_t)(cnz > 0 ? cnz : 1) * sizeof(double));
    e->stype = A->stype;
    e->packed = A->packed;
    e->in_nnz = A->packed ? ((const int64_t *)A->p)[n] : (int64_t)A->nzmax;
    e->in_p = dup64(A->p, n + 1);
    e->in_i = dup64(A->i, e->in_nnz);
    e->in_nz = A->packed ? NULL : dup64(A->nz, n);
    if (!e->vals || !e->in_p || !e->in_i || (!A->packed && !e->in_nz)) return CHOLMOD_OUT_OF_MEMORY;
    return CHOLMOD_OK;
}

/* The inputs the device path takes; anything else goes to the host CHOLMOD. */
static int device_matrix(const cholmod_sparse *A) {
    return A && A->stype != 0 && A->nrow == A->ncol && A->itype == CHOLMOD_LONG &&
           (A->xtype == CHOLMOD_REAL || A->xtype == CHOLMOD_PATTERN) && A->dtype == CHOLMOD_DOUBLE;
}

/* ------------------------------------------------------------------------ */
/* analyze                                                                  */
/* ------------------------------------------------------------------------ */

cholmod_factor *cholmod_l_analyze_p(cholmod_sparse *A, int64_t *UserPerm, int64_t *fset, size_t fsize,
                                    cholmod_common *cm) {
    if (!cm) return NULL;
    if (!A) {
        report(cm, CHOLMOD_INVALID, "argument missing");
        return NULL;
    }
    const int nm = cm->nmethods < 0 ? 0 : cm->nmethods > CHOLMOD_MAXMETHODS ? CHOLMOD_MAXMETHODS : cm->nmethods;
    /* anything the restated orderings do not cover goes to the host library:
     * NESDIS, other methods, non-default AMD controls (cholmod_amd.c:160-164),
     * no postorder */
    int custom = !cm->postorder || (nm == 0 && cm->default_nesdis);
    for (int m = 0; m < (nm == 0 ? 3 : nm); m++) {
        const int o = nm == 0 ? (m == 1 ? CHOLMOD_AMD : CHOLMOD_GIVEN) : cm->method[m].ordering;
        if (o != CHOLMOD_NATURAL && o != CHOLMOD_GIVEN && o != CHOLMOD_AMD && o != CHOLMOD_COLAMD &&
            o != CHOLMOD_METIS)
            custom = 1;
        if ((o == CHOLMOD_AMD || o == CHOLMOD_COLAMD) &&
            (cm->method[m].prune_dense != 10.0 || !cm->method[m].aggressive))
            custom = 1;
    }
    /* the AMD backup after failed methods runs with the controls of the slot after the last */
    if (nm > 0 && nm <= CHOLMOD_MAXMETHODS &&
        (cm->method[nm].prune_dense != 10.0 || !cm->method[nm].aggressive))
        custom = 1;
    if (!use_gpu(cm) || !device_matrix(A) || custom)
        return NEXT(analyze_p_fn, "cholmod_l_analyze_p")(A, UserPerm, fset, fsize, cm);

    cm->status = CHOLMOD_OK;
    cm->selected = -1;
    cm->called_nd = 0;
    const int64_t n = (int64_t)A->nrow;
    int64_t *Cp, *Ci, *src, cnz;
    int rc = lower_pattern(A, &Cp, &Ci, &src, &cnz);
    free(src);
    if (rc) {
        report(cm, rc == -1 ? CHOLMOD_OUT_OF_MEMORY : CHOLMOD_INVALID,
               rc == -3 ? "duplicate entries in A" : "invalid A");
        return NULL;
    }
    cm->anz = -1;                        /* set by the AMD and METIS orderings */

    /* the methods to try (cholmod_analyze.c:444-490) */
    int methods[CHOLMOD_MAXMETHODS + 1], nmeth, amd_backup;
    const int default_strategy = nm == 0;
    if (default_strategy) {
        methods[0] = CHOLMOD_GIVEN;      /* skipped without a UserPerm */
        methods[1] = CHOLMOD_AMD;
        methods[2] = CHOLMOD_METIS;      /* default_nesdis goes to the host library */
        nmeth = 3;
        amd_backup = 0;
        for (int m = 0; m < 3; m++) cm->method[m].ordering = methods[m];
    } else {
        for (int m = 0; m < nm; m++) methods[m] = cm->method[m].ordering;
        nmeth = nm;
        amd_backup = nm > 1 || (nm == 1 && (methods[0] == CHOLMOD_METIS || methods[0] == CHOLMOD_NESDIS ||
                                            (methods[0] == CHOLMOD_GIVEN && UserPerm == NULL)));
    }

    /* Each method yields a permutation and (fl, lnz): AMD its own prediction
     * (cholmod_amd.c:177-180), the others their exact analysis.  The first
     * strictly smallest lnz wins; AMD's symbolic factor is only built if it
     * wins (cholmod_analyze.c:808-822). */
    mxc_symbolic *best = NULL;
    int64_t *best_perm = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    int64_t *cand = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    if (!best_perm || !cand) {
        free(best_perm); free(cand); free(Cp); free(Ci);
        report(cm, CHOLMOD_OUT_OF_MEMORY, "out of memory");
        return NULL;
    }
    int best_order = 0, status = CHOLMOD_OK;
    double best_lnz = -1;
    for (int m = 0; m <= nmeth; m++) {
        int ordering;
        if (m == nmeth) {
            if (cm->selected == -1 && amd_backup) ordering = CHOLMOD_AMD;
            else break;
        } else {
            ordering = methods[m];
        }
        cm->current = m;
        cm->fl = cm->lnz = -1;
        mxc_symbolic *S = NULL;
        int st = MXC_OK, amd = 0;
        double mfl = -1, mlnz = -1;
        if (ordering == CHOLMOD_NATURAL) {
            st = mxc_analyze(n, Cp, Ci, MXC_ORDER_NATURAL, NULL, &S);
        } else if (ordering == CHOLMOD_GIVEN) {
            if (!UserPerm) continue;
            st = mxc_analyze(n, Cp, Ci, MXC_ORDER_GIVEN, UserPerm, &S);
        } else if (ordering == CHOLMOD_AMD || ordering == CHOLMOD_COLAMD) {
            double stats[3];
            amd_backup = 0;
            amd = 1;
            st = mxc_amd_order_stats(n, Cp, Ci, cand, stats);
            if (st == MXC_OK) {
                mfl = stats[0];
                mlnz = stats[1];
                cm->anz = stats[2];
            }
        } else if (ordering == CHOLMOD_METIS) {
            cm->called_nd = 1;
            st = mxc_metis_order(n, Cp, Ci, cm->metis_nswitch, cm->metis_dswitch, cand);
            if (st == MXC_OK) {
                int64_t nz = 0;
                for (int64_t j = 0; j < n; j++)
                    for (int64_t p = Cp[j]; p < Cp[j + 1]; p++) nz += Ci[p] > j;
                cm->anz = (double)(nz + n);
                st = mxc_analyze(n, Cp, Ci, MXC_ORDER_GIVEN, cand, &S);
            }
        } else {
            status = status < CHOLMOD_NOT_INSTALLED ? status : CHOLMOD_NOT_INSTALLED;
            continue;                     /* NESDIS: not on this path */
        }
        if (st != MXC_OK) {
            status = status < st ? status : st;
            continue;
        }
        if (S) {
            int64_t info[7];
            double dinfo[2];
            mxc_symbolic_info(S, info, dinfo);
            mfl = dinfo[0];
            mlnz = dinfo[1];
        }
        cm->fl = mfl;
        cm->lnz = mlnz;
        if (m < CHOLMOD_MAXMETHODS) {
            cm->method[m].fl = mfl;
            cm->method[m].lnz = mlnz;
        }
        if (cm->selected == -1 || mlnz < best_lnz) {
            if (best) mxc_symbolic_free(best);
            best = S;                     /* NULL for AMD: built below from best_perm */
            S = NULL;
            if (amd) memcpy(best_perm, cand, (size_t)n * sizeof(int64_t));
            best_lnz = mlnz;
            best_order = ordering == CHOLMOD_COLAMD ? CHOLMOD_AMD : ordering;
            cm->selected = m;
        }
        if (S) mxc_symbolic_free(S);
        /* AMD already good: skip METIS (cholmod_analyze.c:767-781) */
        if (default_strategy && amd && (cm->fl < 500 * cm->lnz || cm->lnz < 5 * cm->anz)) break;
    }
    free(cand);
    if (cm->selected != -1 && !best) {
        int st = mxc_analyze(n, Cp, Ci, MXC_ORDER_GIVEN, best_perm, &best);
        if (st != MXC_OK) status = status < st ? status : st;
    }
    free(best_perm);
    if (!best) {
        free(Cp);
        free(Ci);
        report(cm, status >= CHOLMOD_OK ? CHOLMOD_INVALID : status, "all methods failed");
        return NULL;
    }

    int64_t info[7];
    double dinfo[2];
    mxc_symbolic_info(best, info, dinfo);
    const int64_t nsuper = info[1], ssize = info[2], xsize = info[3];
    cm->fl = dinfo[0];
    cm->lnz = dinfo[1];
    cholmod_factor *L = cholmod_l_alloc_factor((size_t)n, A->dtype, cm);
    if (!L) {
        mxc_symbolic_free(best);
        free(Cp);
        free(Ci);
        return NULL;
    }
    mxc_symbolic_export(best, (int64_t *)L->Perm, (int64_t *)L->ColCount, NULL, NULL, NULL, NULL);
    /* a natural ordering is postordered (cholmod_analyze.c:866-870) */
    L->ordering = best_order == CHOLMOD_NATURAL ? CHOLMOD_POSTORDERED : best_order;
    const int supernodal = cm->supernodal > CHOLMOD_AUTO ||
                           (cm->supernodal == CHOLMOD_AUTO && cm->lnz > 0 &&
                            cm->fl / cm->lnz >= cm->supernodal_switch);
    if (supernodal) {
        L->nsuper = (size_t)nsuper;
        L->ssize = (size_t)ssize;
        L->xsize = (size_t)xsize;
        cholmod_l_change_factor(CHOLMOD_PATTERN, 1, 1, 1, 1, L, cm);
        if (cm->status < CHOLMOD_OK) {
            cholmod_l_free_factor(&L, cm);
            mxc_symbolic_free(best);
            free(Cp);
            free(Ci);
            return NULL;
        }
        mxc_symbolic_export(best, NULL, NULL, (int64_t *)L->super, (int64_t *)L->pi, (int64_t *)L->px,
                            (int64_t *)L->s);
        L->maxcsize = (size_t)info[4];
        L->maxesize = (size_t)info[5];
        L->is_super = 1;
        L->is_ll = 1;
        L->useGPU = 1;
    }
    mxc_symbolic_free(best);
    free(Cp);
    free(Ci);
    cm->status = CHOLMOD_OK;
    return L;
}

cholmod_factor *cholmod_l_analyze(cholmod_sparse *A, cholmod_common *cm) {
    return cholmod_l_analyze_p(A, NULL, NULL, 0, cm);
}

/* ------------------------------------------------------------------------ */
/* factorize                                                                */
/* ------------------------------------------------------------------------ */

/* Build (or reuse) the device plan of L for A.  Returns a CHOLMOD status;
 * *e_out is the entry (table lock held by the caller). */
static int device_plan(cholmod_factor *L, cholmod_sparse *A, cholmod_common *cm, entry **e_out) {
    entry *e = tab_find(L);
    if (!e) {
        e = tab_add(L, cm);
        if (!e) return CHOLMOD_OUT_OF_MEMORY;
    }
    *e_out = e;
    int st = bind_input(e, A);
    if (st != CHOLMOD_OK) return st;
    if (!e->S) {
        st = mxc_symbolic_import((int64_t)L->n, (int64_t)L->nsuper, (const int64_t *)L->Perm,
                                 (const int64_t *)L->super, (const int64_t *)L->pi, (const int64_t *)L->px,
                                 (const int64_t *)L->s, &e->S);
        if (st != MXC_OK) return st;
    }
    if (!e->F) {
        mxc_pool *P = common_pool(cm);
        if (!P) return CHOLMOD_GPU_PROBLEM;
        st = mxc_factor_create_in(e->S, e->Cp, e->Ci, MXC_FP64, P, &e->F);
        if (st != MXC_OK) return st;
    }
    return CHOLMOD_OK;
}

int cholmod_l_factorize_p(cholmod_sparse *A, double *beta, int64_t *fset, size_t fsize, cholmod_factor *L,
                          cholmod_common *cm) {
    if (!cm) return 0;
    if (!A || !L) {
        report(cm, CHOLMOD_INVALID, "argument missing");
        return 0;
    }
    const int device = use_gpu(cm) && device_matrix(A) && A->xtype == CHOLMOD_REAL && L->is_super &&
                       L->n == A->nrow && L->itype == CHOLMOD_LONG && L->dtype == CHOLMOD_DOUBLE;
    if (device) {
        pthread_mutex_lock(&g_tab_mu);
        entry *e = NULL;
        int st = device_plan(L, A, cm, &e);
        if (st == CHOLMOD_OUT_OF_MEMORY && e && !e->F && cm->dev_mempool) {
            /* the factor does not fit the Common's pool: the host factorizes it,
               as the reference keeps the supernodes that do not fit on the CPU */
            entry_release(e);
            *e = g_tab[--g_ntab];
            pthread_mutex_unlock(&g_tab_mu);
            return NEXT(factorize_p_fn, "cholmod_l_factorize_p")(A, beta, fset, fsize, L, cm);
        }
        if (st != CHOLMOD_OK) {
            pthread_mutex_unlock(&g_tab_mu);
            report(cm, st, st == CHOLMOD_INVALID ? "A's pattern is not inside the analyzed L" : mxc_last_error());
            return 0;
        }
        const double *Ax = (const double *)A->x;
        for (int64_t k = 0; k < e->cnz; k++) e->vals[k] = Ax[e->src[k]];
        int64_t minor = 0;
        st = mxc_factorize(e->F, e->vals, MXC_HOST, beta ? beta[0] : 0.0, &minor);
        if (st < 0) {
            e->fresh = 0;
            pthread_mutex_unlock(&g_tab_mu);
            report(cm, st, mxc_last_error());
            return 0;
        }
        e->fresh = 1;
        add_factor_stats(cm, e->F);
        cm->gpuFlops += (int64_t)cm->fl;
        mxc_factor *F = e->F;
        pthread_mutex_unlock(&g_tab_mu);
        /* L->x on the host, in the reference's layout (API contract) */
        cholmod_l_change_factor(CHOLMOD_REAL, 1, 1, 1, 1, L, cm);
        if (cm->status < CHOLMOD_OK) return 0;
        const char *keep = getenv("CHOLMOD_MI355X_KEEP_L_ON_DEVICE");
        if (!(keep && atoi(keep) != 0)) {
            int g = mxc_factor_get_Lx(F, (double *)L->x);
            if (g != MXC_OK) {
                report(cm, g, mxc_last_error());
                return 0;
            }
        }
        L->minor = (size_t)minor;
        L->is_ll = 1;
        cm->status = CHOLMOD_OK;
        if (st == MXC_NOT_POSDEF) report(cm, CHOLMOD_NOT_POSDEF, "matrix not positive definite");
        return 1;
    }
    tab_drop(L);                       /* the host path changes L's values */
    return NEXT(factorize_p_fn, "cholmod_l_factorize_p")(A, beta, fset, fsize, L, cm);
}

int cholmod_l_factorize(cholmod_sparse *A, cholmod_factor *L, cholmod_common *cm) {
    double zero[2] = {0, 0};
    return cholmod_l_factorize_p(A, zero, NULL, 0, L, cm);
}

/* ------------------------------------------------------------------------ */
/* solve                                                                    */
/* ------------------------------------------------------------------------ */

/* The device factor of L when it holds L's current values (table lock held). */
static entry *fresh_entry(cholmod_factor *L) {
    entry *e = tab_find(L);
    return (e && e->fresh && e->F) ? e : NULL;
}

/* X = solve(sys, L, B) on the device: X, B n x nrhs real double (any ld).
 * Returns a CHOLMOD status. */
static int device_solve(entry *e, cholmod_factor *L, int sys, const cholmod_dense *B, cholmod_dense *X) {
    const int64_t n = e->n, nrhs = (int64_t)B->ncol;
    if (n == 0 || nrhs == 0) return CHOLMOD_OK;
    double *Bp = (double *)B->x, *Xp = (double *)X->x;
    double *tmp = NULL;
    if (B->d != (size_t)n) {                /* pack B to ld = n */
        tmp = (double *)malloc((size_t)(n * nrhs) * sizeof(double));
        if (!tmp) return CHOLMOD_OUT_OF_MEMORY;
        for (int64_t c = 0; c < nrhs; c++) memcpy(tmp + c * n, Bp + c * B->d, (size_t)n * sizeof(double));
        Bp = tmp;
    }
    double *Xw = X->d == (size_t)n ? Xp : (double *)malloc((size_t)(n * nrhs) * sizeof(double));
    if (!Xw) { free(tmp); return CHOLMOD_OUT_OF_MEMORY; }
    const int64_t *Perm = (const int64_t *)L->Perm;
    const int use_perm = L->ordering != CHOLMOD_NATURAL;
    int st = MXC_OK;
    switch (sys) {
        case CHOLMOD_A: case CHOLMOD_LDLt:
            st = mxc_solve(e->F, Bp, nrhs, Xw, MXC_HOST);
            break;
        case CHOLMOD_LD: case CHOLMOD_L: case CHOLMOD_Lt: case CHOLMOD_DLt:
            memmove(Xw, Bp, (size_t)(n * nrhs) * sizeof(double));
            st = (sys == CHOLMOD_LD || sys == CHOLMOD_L) ? mxc_super_lsolve(e->F, Xw, nrhs, MXC_HOST)
                                                          : mxc_super_ltsolve(e->F, Xw, nrhs, MXC_HOST);
            break;
        case CHOLMOD_D:
            memmove(Xw, Bp, (size_t)(n * nrhs) * sizeof(double));     /* D = I for LL' */
            break;
        case CHOLMOD_P: case CHOLMOD_Pt:
            for (int64_t c = 0; c < nrhs; c++)
                for (int64_t k = 0; k < n; k++) {
                    const int64_t pk = use_perm ? Perm[k] : k;
                    if (sys == CHOLMOD_P) Xw[c * n + k] = Bp[c * n + pk];
                    else Xw[c * n + pk] = Bp[c * n + k];
                }
            break;
        default:
            st = MXC_INVALID;
    }
    if (Xw != Xp) {
        for (int64_t c = 0; c < nrhs; c++) memcpy(Xp + c * X->d, Xw + c * n, (size_t)n * sizeof(double));
        free(Xw);
    }
    free(tmp);
    return st;
}

static int device_rhs(const cholmod_dense *B, const cholmod_factor *L) {
    return B && B->xtype == CHOLMOD_REAL && B->dtype == CHOLMOD_DOUBLE && B->nrow == L->n &&
           B->d >= B->nrow;
}

cholmod_dense *cholmod_l_solve(int sys, cholmod_factor *L, cholmod_dense *B, cholmod_common *cm) {
    if (!cm) return NULL;
    if (L && device_rhs(B, L) && sys >= CHOLMOD_A && sys <= CHOLMOD_Pt) {
        pthread_mutex_lock(&g_tab_mu);
        entry *e = fresh_entry(L);
        if (e) {
            cholmod_dense *X = cholmod_l_allocate_dense(L->n, B->ncol, L->n, CHOLMOD_REAL + CHOLMOD_DOUBLE, cm);
            if (!X) {
                pthread_mutex_unlock(&g_tab_mu);
                return NULL;
            }
            int st = device_solve(e, L, sys, B, X);
            if (st == MXC_OK) add_solve_stats(cm, e->F);
            pthread_mutex_unlock(&g_tab_mu);
            if (st != MXC_OK) {
                cholmod_l_free_dense(&X, cm);
                report(cm, st, mxc_last_error());
                return NULL;
            }
            cm->status = CHOLMOD_OK;
            return X;
        }
        pthread_mutex_unlock(&g_tab_mu);
    }
    return NEXT(solve_fn, "cholmod_l_solve")(sys, L, B, cm);
}

int cholmod_l_solve2(int sys, cholmod_factor *L, cholmod_dense *B, cholmod_sparse *Bset, cholmod_dense **X_Handle,
                     cholmod_sparse **Xset_Handle, cholmod_dense **Y_Handle, cholmod_dense **E_Handle,
                     cholmod_common *cm) {
    if (!cm) return 0;
    int device = L && X_Handle && device_rhs(B, L) && sys >= CHOLMOD_A && sys <= CHOLMOD_Pt;
    if (device && Bset) {
        /* sparse right-hand side: one column, int64 pattern, a place for Xset */
        device = Xset_Handle && B->ncol == 1 && Bset->ncol == 1 && Bset->nrow == L->n &&
                 Bset->itype == CHOLMOD_LONG && sys != CHOLMOD_P && sys != CHOLMOD_Pt;
    }
    if (device) {
        pthread_mutex_lock(&g_tab_mu);
        entry *e = fresh_entry(L);
        if (e) {
            const int64_t n = e->n;
            cholmod_dense *X = cholmod_l_ensure_dense(X_Handle, L->n, B->ncol, L->n,
                                                      CHOLMOD_REAL + CHOLMOD_DOUBLE, cm);
            if (!X) {
                pthread_mutex_unlock(&g_tab_mu);
                return 0;
            }
            int st;
            if (!Bset) {
                /* the reference's supernodal workspaces (cholmod_solve.c:663-676), kept for the caller */
                if (Y_Handle) cholmod_l_ensure_dense(Y_Handle, L->n, B->ncol, L->n, CHOLMOD_REAL + CHOLMOD_DOUBLE, cm);
                if (E_Handle)
                    cholmod_l_ensure_dense(E_Handle, B->ncol, L->maxesize, B->ncol, CHOLMOD_REAL + CHOLMOD_DOUBLE, cm);
                st = device_solve(e, L, sys, B, X);
            } else {
                /* Xset: n x 1, nzmax >= n, pattern, packed, unsorted (cholmod_solve.c:386-404) */
                cholmod_sparse *Xset = *Xset_Handle;
                if (!Xset || (int64_t)Xset->nrow != n || Xset->ncol != 1 || (int64_t)Xset->nzmax < n ||
                    Xset->xtype != CHOLMOD_PATTERN) {
                    cholmod_l_free_sparse(Xset_Handle, cm);
                    Xset = cholmod_l_allocate_sparse(L->n, 1, L->n, 0, 1, 0, CHOLMOD_PATTERN + L->dtype, cm);
                    *Xset_Handle = Xset;
                }
                if (Y_Handle) cholmod_l_ensure_dense(Y_Handle, 1, L->n, 1, CHOLMOD_REAL + CHOLMOD_DOUBLE, cm);
                if (!Xset || cm->status < CHOLMOD_OK) {
                    pthread_mutex_unlock(&g_tab_mu);
                    return 0;
                }
                Xset->sorted = 0;
                Xset->stype = 0;
                const int64_t *Bp = (const int64_t *)Bset->p;
                const int64_t blen = Bset->packed ? Bp[1] - Bp[0] : ((const int64_t *)Bset->nz)[0];
                int64_t xlen = 0;
                st = mxc_solve_bset(e->F, sys, (const double *)B->x, (const int64_t *)Bset->i + Bp[0], blen,
                                    (double *)X->x, (int64_t *)Xset->i, &xlen);
                ((int64_t *)Xset->p)[0] = 0;
                ((int64_t *)Xset->p)[1] = xlen;
            }
            if (st == MXC_OK) add_solve_stats(cm, e->F);
            pthread_mutex_unlock(&g_tab_mu);
            if (st != MXC_OK) {
                report(cm, st, mxc_last_error());
                return 0;
            }
            if (Bset && L->is_super) {
                /* the reference leaves L simplicial after a Bset solve (cholmod_solve.c:312-331);
                   the device factor keeps serving (the values do not change) */
                pthread_mutex_lock(&g_tab_mu);
                const char *keep = getenv("CHOLMOD_MI355X_KEEP_L_ON_DEVICE");
                entry *e1 = tab_find(L);
                if (keep && atoi(keep) != 0 && e1) mxc_factor_get_Lx(e1->F, (double *)L->x);
                cholmod_l_change_factor(CHOLMOD_REAL, 1, 0, 0, 0, L, cm);
                entry *e2 = tab_find(L);
                if (e2) {
                    e2->stamp_perm = L->Perm;
                    e2->stamp_super = L->super;
                }
                pthread_mutex_unlock(&g_tab_mu);
                if (cm->status < CHOLMOD_OK) return 0;
            }
            cm->status = CHOLMOD_OK;
            return 1;
        }
        pthread_mutex_unlock(&g_tab_mu);
    }
    return NEXT(solve2_fn, "cholmod_l_solve2")(sys, L, B, Bset, X_Handle, Xset_Handle, Y_Handle, E_Handle, cm);
}

/* ------------------------------------------------------------------------ */
/* update / downdate                                                        */
/* ------------------------------------------------------------------------ */

/* cholmod_l_updown on the device factor when L's device copy is current and C
 * is real, double and sorted: mxc_updown modifies it in place (L stays a
 * supernodal LL' factor) and L->x is refreshed.  The host library takes the
 * rest from L->x as it was before the call, and turns L into a simplicial
 * LDL' factor as the reference does: C with an entry outside the pattern of
 * L(:, first row of its column) (new fill), a downdate whose new pivot is not
 * positive (the reference keeps that indefinite LDL'; the device factor is
 * dropped), and every other input. */
int cholmod_l_updown(int update, cholmod_sparse *C, cholmod_factor *L, cholmod_common *cm) {
    if (!cm) return 0;
    if (!C || !L) {
        report(cm, CHOLMOD_INVALID, "argument missing");
        return 0;
    }
    const char *keep = getenv("CHOLMOD_MI355X_KEEP_L_ON_DEVICE");
    const int lx_on_host = !(keep && atoi(keep) != 0);
    pthread_mutex_lock(&g_tab_mu);
    entry *e = fresh_entry(L);
    if (e && L->is_super && C->xtype == CHOLMOD_REAL && C->dtype == CHOLMOD_DOUBLE && C->sorted &&
        C->nrow == L->n) {
        const int64_t ncol = (int64_t)C->ncol;
        const int64_t *Cp = (const int64_t *)C->p, *Ci = (const int64_t *)C->i;
        const double *Cx = (const double *)C->x;
        int64_t *pp = NULL, *pi = NULL;
        double *px = NULL;
        if (!C->packed) {                 /* a packed copy of an unpacked C */
            const int64_t *Cnz = (const int64_t *)C->nz;
            int64_t nz = 0;
            for (int64_t c = 0; c < ncol; c++) nz += Cnz[c];
            pp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncol + 1));
            pi = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nz > 0 ? nz : 1));
            px = (double *)malloc(sizeof(double) * (size_t)(nz > 0 ? nz : 1));
            if (!pp || !pi || !px) {
                free(pp), free(pi), free(px);
                pthread_mutex_unlock(&g_tab_mu);
                report(cm, CHOLMOD_OUT_OF_MEMORY, "out of memory");
                return 0;
            }
            pp[0] = 0;
            for (int64_t c = 0; c < ncol; c++) {
                memcpy(pi + pp[c], Ci + Cp[c], sizeof(int64_t) * (size_t)Cnz[c]);
                memcpy(px + pp[c], Cx + Cp[c], sizeof(double) * (size_t)Cnz[c]);
                pp[c + 1] = pp[c] + Cnz[c];
            }
            Cp = pp, Ci = pi, Cx = px;
        }
        double modfl = 0;
        int64_t info[2] = {0, 0};
        int st = mxc_updown(e->F, update, ncol, Cp, Ci, Cx, cm->dbound, &modfl, info);
        free(pp), free(pi), free(px);
        if (st == MXC_OK) {
            if (lx_on_host) st = mxc_factor_get_Lx(e->F, (double *)L->x);
            pthread_mutex_unlock(&g_tab_mu);
            if (st != MXC_OK) {
                report(cm, st, mxc_last_error());
                return 0;
            }
            cm->status = CHOLMOD_OK;
            cm->modfl = modfl;
            cm->ndbounds_hit += (double)info[1];
            if (info[1] > 0) report(cm, CHOLMOD_DSMALL, "diagonal entry is below threshold");
            return 1;
        }
        if (st != MXC_INVALID && st != MXC_NOT_POSDEF) {
            pthread_mutex_unlock(&g_tab_mu);
            report(cm, st, mxc_last_error());
            return 0;
        }
        if (!lx_on_host) {
            /* L->x was never filled: copy the device factor out first, if it still holds L */
            if (st == MXC_NOT_POSDEF || mxc_factor_get_Lx(e->F, (double *)L->x) != MXC_OK) {
                pthread_mutex_unlock(&g_tab_mu);
                tab_drop(L);
                report(cm, st == MXC_NOT_POSDEF ? CHOLMOD_NOT_POSDEF : CHOLMOD_GPU_PROBLEM,
                       "update/downdate needs the host factor, which was kept on the device only");
                return 0;
            }
        }
    }
    pthread_mutex_unlock(&g_tab_mu);
    tab_drop(L);                       /* the host path changes L */
    return NEXT(updown_fn, "cholmod_l_updown")(update, C, L, cm);
}

/* ------------------------------------------------------------------------ */
/* lifetime and statistics                                                  */
/* ------------------------------------------------------------------------ */

int cholmod_l_free_factor(cholmod_factor **LHandle, cholmod_common *cm) {
    if (LHandle && *LHandle) tab_drop(*LHandle);
    return NEXT(free_factor_fn, "cholmod_l_free_factor")(LHandle, cm);
}

int cholmod_l_finish(cholmod_common *cm) {
    if (cm) {
        pthread_mutex_lock(&g_tab_mu);
        for (size_t k = 0; k < g_ntab;) {
            if (g_tab[k].cm == cm) {
                entry_release(&g_tab[k]);
                g_tab[k] = g_tab[--g_ntab];
            } else {
                k++;
            }
        }
        pthread_mutex_unlock(&g_tab_mu);
        if (cm->dev_mempool) {
            mxc_pool_free((mxc_pool *)cm->dev_mempool);
            cm->dev_mempool = NULL;
            cm->dev_mempool_size = 0;
        }
    }
    return NEXT(finish_fn, "cholmod_l_finish")(cm);
}

int cholmod_l_gpu_stats(cholmod_common *cm) {
    if (!cm) return 0;
    if (cm->print >= 2) {
        printf("\nCHOLMOD MI355X statistics:\n");
        printf("descendant updates (SYRK+GEMM) launches %12.0f\n", (double)cm->cholmod_gpu_syrk_calls);
        printf("trailing updates (GEMM)        launches %12.0f\n", (double)cm->cholmod_gpu_gemm_calls);
        printf("diagonal POTRF                 launches %12.0f\n", (double)cm->cholmod_gpu_potrf_calls);
        printf("panel TRSM                     launches %12.0f\n", (double)cm->cholmod_gpu_trsm_calls);
        printf("kernel launches %d, device time %12.4e s, flops %12.4e\n", cm->gpuNumKernelLaunches,
               cm->gpuKernelTime, (double)cm->gpuFlops);
        if (cm->dev_mempool) {
            int64_t info[4];
            mxc_pool_info((const mxc_pool *)cm->dev_mempool, info);
            printf("device pool %12.0f bytes, %12.0f in use, peak %12.0f\n", (double)info[0], (double)info[1],
                   (double)info[2]);
        }
    }
    return NEXT(gpu_stats_fn, "cholmod_l_gpu_stats")(cm);
}
