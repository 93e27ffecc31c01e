// ref_driver.c -- test-side glue around the reference CHOLMOD (CPU path).
//
// ORACLE / TEST INFRASTRUCTURE ONLY.  Linked into oracle/_ref/libssref.so
// together with the reference sources (oracle/ref.mk).  Only tests/, the
// smoke check and the CPU-baseline leg of bench.py load it; the product never
// does.
//
// It exposes plain-pointer entry points over the reference's own API:
//   cholmod_l_start / finish                 CHOLMOD/Include/cholmod.h:1028,1035
//   cholmod_l_analyze / cholmod_l_analyze_p  cholmod.h:2714, :2734
//   cholmod_l_factorize                      cholmod.h:2775
//   cholmod_l_solve                          cholmod.h:2823
//   cholmod_l_solve2 (sparse b, Bset)        cholmod.h:2845
//   cholmod_l_sdmult (residual)              cholmod.h (MatrixOps)
//   cholmod_l_read_sparse2 (file reader)     CHOLMOD/Check/cholmod_read.c
//   cholmod_l_updown (rank-k modification)   CHOLMOD/Modify/cholmod_updown.c:74
// Input matrices are the lower triangle of a symmetric matrix in CSC form
// (stype = -1), int64 indices, as CHOLMOD's int64 ("_l_") interface takes.

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>
#include "cholmod.h"

extern int MKL_Set_Threading_Layer(int);   // exported by libmkl_rt
extern void MKL_Set_Num_Threads(int);

typedef struct {
    cholmod_common cm;
    cholmod_sparse *A;
    cholmod_factor *L;
    int dtype;             // CHOLMOD_DOUBLE or CHOLMOD_SINGLE for the factor
    // supernodal structure kept by ref_updown (the reference turns L simplicial)
    int64_t nsuper, *Super, *Lpi, *Lpx, *Ls;
} ref_handle;

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int g_threads_set = 0;

// Pin MKL to the GNU OpenMP layer (MKL_THREADING_GNU == 3) before any BLAS
// call, and size both OpenMP and MKL thread pools.
int ref_set_threads(int nthreads) {
    setenv("MKL_THREADING_LAYER", "GNU", 1);
    MKL_Set_Threading_Layer(3);
    if (nthreads > 0) {
        omp_set_num_threads(nthreads);
        MKL_Set_Num_Threads(nthreads);
    }
    g_threads_set = 1;
    return omp_get_max_threads();
}

// ordering: 0 = given (UserPerm), 1 = AMD, 2 = METIS, 3 = natural,
//           4 = CHOLMOD's default strategy (nmethods = 0; UserPerm used if not NULL)
// dtype: 0 = double factor, 1 = single factor (values converted on input)
void *ref_analyze(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax,
                  int ordering, const int64_t *UserPerm, int single, double *t_analyze) {
    if (!g_threads_set) ref_set_threads(0);
    ref_handle *h = (ref_handle *)calloc(1, sizeof(ref_handle));
    if (!h) return NULL;
    cholmod_common *cm = &h->cm;
    cholmod_l_start(cm);
    cm->useGPU = 0;
    cm->supernodal = CHOLMOD_SUPERNODAL;
    cm->final_ll = 0;
    cm->nmethods = 1;
    cm->postorder = 1;
    switch (ordering) {
        case 0: cm->method[0].ordering = CHOLMOD_GIVEN; break;
        case 1: cm->method[0].ordering = CHOLMOD_AMD; break;
        case 2: cm->method[0].ordering = CHOLMOD_METIS; break;
        case 4: cm->nmethods = 0; break;
        default: cm->method[0].ordering = CHOLMOD_NATURAL; break;
    }
    int64_t nnz = Ap[n];
    h->A = cholmod_l_allocate_sparse(n, n, nnz, 1, 1, -1, CHOLMOD_REAL + CHOLMOD_DOUBLE, cm);
    if (!h->A) { cholmod_l_finish(cm); free(h); return NULL; }
    memcpy(h->A->p, Ap, sizeof(int64_t) * (n + 1));
    memcpy(h->A->i, Ai, sizeof(int64_t) * nnz);
    if (Ax) memcpy(h->A->x, Ax, sizeof(double) * nnz);
    else memset(h->A->x, 0, sizeof(double) * nnz);
    if (single) {
        cholmod_l_sparse_xtype(CHOLMOD_REAL + CHOLMOD_SINGLE, h->A, cm);
        h->dtype = CHOLMOD_SINGLE;
    }
    double t0 = now_s();
    if (ordering == 0 || (ordering == 4 && UserPerm))
        h->L = cholmod_l_analyze_p(h->A, (int64_t *)UserPerm, NULL, 0, cm);
    else
        h->L = cholmod_l_analyze(h->A, cm);
    if (t_analyze) *t_analyze = now_s() - t0;
    if (!h->L) {
        cholmod_l_free_sparse(&h->A, cm);
        cholmod_l_finish(cm);
        free(h);
        return NULL;
    }
    return h;
}

// Common->selected, method[0..2].fl / lnz and anz after ref_analyze
int ref_strategy(void *vh, int64_t *selected, double *method_fl, double *method_lnz, double *anz) {
    ref_handle *h = (ref_handle *)vh;
    *selected = h->cm.selected;
    for (int m = 0; m < 3; m++) {
        method_fl[m] = h->cm.method[m].fl;
        method_lnz[m] = h->cm.method[m].lnz;
    }
    *anz = h->cm.anz;
    return 0;
}

// info[0..9] = n, nsuper, ssize, xsize, maxcsize, maxesize, is_super, ordering, status, minor
// dinfo[0..1] = fl, lnz
int ref_info(void *vh, int64_t *info, double *dinfo) {
    ref_handle *h = (ref_handle *)vh;
    cholmod_factor *L = h->L;
    info[0] = (int64_t)L->n;
    info[1] = (int64_t)L->nsuper;
    info[2] = (int64_t)L->ssize;
    info[3] = (int64_t)L->xsize;
    info[4] = (int64_t)L->maxcsize;
    info[5] = (int64_t)L->maxesize;
    info[6] = (int64_t)L->is_super;
    info[7] = (int64_t)L->ordering;
    info[8] = (int64_t)h->cm.status;
    info[9] = (int64_t)L->minor;
    dinfo[0] = h->cm.fl;
    dinfo[1] = h->cm.lnz;
    return 0;
}

int ref_get_symbolic(void *vh, int64_t *Perm, int64_t *ColCount, int64_t *Super,
                     int64_t *Lpi, int64_t *Lpx, int64_t *Ls) {
    ref_handle *h = (ref_handle *)vh;
    cholmod_factor *L = h->L;
    size_t n = L->n, ns = L->nsuper;
    if (Perm) memcpy(Perm, L->Perm, sizeof(int64_t) * n);
    if (ColCount) memcpy(ColCount, L->ColCount, sizeof(int64_t) * n);
    if (!L->is_super) return -1;
    if (Super) memcpy(Super, L->super, sizeof(int64_t) * (ns + 1));
    if (Lpi) memcpy(Lpi, L->pi, sizeof(int64_t) * (ns + 1));
    if (Lpx) memcpy(Lpx, L->px, sizeof(int64_t) * (ns + 1));
    if (Ls) memcpy(Ls, L->s, sizeof(int64_t) * L->ssize);
    return 0;
}

// Numeric factorization.  Ax_new (optional) replaces the values of A (same
// pattern, refactorization).  Returns cholmod status; L->minor in *minor.
int ref_factorize(void *vh, const double *Ax_new, double *t_factor, int64_t *minor) {
    ref_handle *h = (ref_handle *)vh;
    cholmod_common *cm = &h->cm;
    if (Ax_new) {
        int64_t nnz = ((int64_t *)h->A->p)[h->A->ncol];
        if (h->dtype == CHOLMOD_SINGLE) {
            float *x = (float *)h->A->x;
            for (int64_t k = 0; k < nnz; k++) x[k] = (float)Ax_new[k];
        } else {
            memcpy(h->A->x, Ax_new, sizeof(double) * nnz);
        }
    }
    double t0 = now_s();
    int ok = cholmod_l_factorize(h->A, h->L, cm);
    if (t_factor) *t_factor = now_s() - t0;
    if (minor) *minor = (int64_t)h->L->minor;
    if (!ok) return cm->status != 0 ? cm->status : -99;
    return cm->status;
}

// cholmod_l_factorize_p (cholmod.h:2794) with beta: factors A + beta[0]*I
// (fset = NULL, the symmetric case of CHOLMOD/Cholesky/cholmod_factorize.c:93-418).
int ref_factorize_p(void *vh, const double *Ax_new, double beta, double *t_factor, int64_t *minor) {
    ref_handle *h = (ref_handle *)vh;
    cholmod_common *cm = &h->cm;
    if (Ax_new) {
        int64_t nnz = ((int64_t *)h->A->p)[h->A->ncol];
        if (h->dtype == CHOLMOD_SINGLE) {
            float *x = (float *)h->A->x;
            for (int64_t k = 0; k < nnz; k++) x[k] = (float)Ax_new[k];
        } else {
            memcpy(h->A->x, Ax_new, sizeof(double) * nnz);
        }
    }
    double b[2] = {beta, 0};
    double t0 = now_s();
    int ok = cholmod_l_factorize_p(h->A, b, NULL, 0, h->L, cm);
    if (t_factor) *t_factor = now_s() - t0;
    if (minor) *minor = (int64_t)h->L->minor;
    if (!ok) return cm->status != 0 ? cm->status : -99;
    return cm->status;
}

// Copy L->x (supernodal values) out as double.
int ref_get_Lx(void *vh, double *Lx) {
    ref_handle *h = (ref_handle *)vh;
    cholmod_factor *L = h->L;
    if (!L->is_super || L->xtype == CHOLMOD_PATTERN) return -1;
    if (h->dtype == CHOLMOD_SINGLE) {
        float *x = (float *)L->x;
        for (size_t k = 0; k < L->xsize; k++) Lx[k] = (double)x[k];
    } else {
        memcpy(Lx, L->x, sizeof(double) * L->xsize);
    }
    return 0;
}

// X = A \ B with the current factor (sys = CHOLMOD_A).  B, X column-major n x nrhs.
int ref_solve(void *vh, const double *B, int64_t nrhs, double *X, double *t_solve) {
    ref_handle *h = (ref_handle *)vh;
    cholmod_common *cm = &h->cm;
    int64_t n = (int64_t)h->L->n;
    int xt = CHOLMOD_REAL + (h->dtype == CHOLMOD_SINGLE ? CHOLMOD_SINGLE : CHOLMOD_DOUBLE);
    cholmod_dense *Bd = cholmod_l_allocate_dense(n, nrhs, n, xt, cm);
    if (!Bd) return -2;
    if (h->dtype == CHOLMOD_SINGLE) {
        float *b = (float *)Bd->x;
        for (int64_t k = 0; k < n * nrhs; k++) b[k] = (float)B[k];
    } else {
        memcpy(Bd->x, B, sizeof(double) * n * nrhs);
    }
    double t0 = now_s();
    cholmod_dense *Xd = cholmod_l_solve(CHOLMOD_A, h->L, Bd, cm);
    if (t_solve) *t_solve = now_s() - t0;
    cholmod_l_free_dense(&Bd, cm);
    if (!Xd) return cm->status != 0 ? cm->status : -99;
    if (h->dtype == CHOLMOD_SINGLE) {
        float *x = (float *)Xd->x;
        for (int64_t k = 0; k < n * nrhs; k++) X[k] = (double)x[k];
    } else {
        memcpy(X, Xd->x, sizeof(double) * n * nrhs);
    }
    cholmod_l_free_dense(&Xd, cm);
    return 0;
}

// cholmod_l_solve2 (cholmod.h:2845) with a sparse right-hand side: b dense
// (n entries), Bset = rows bset[0..blen).  xset[0..*xlen) receives Xset in the
// reference's order and x the solution at those rows (other entries are not
// written).  The reference converts the factor to simplicial here, so this is
// the last call to make on a handle.
int ref_solve2_bset(void *vh, int sys, const double *b, const int64_t *bset, int64_t blen,
                    double *x, int64_t *xset, int64_t *xlen) {
    ref_handle *h = (ref_handle *)vh;
    cholmod_common *cm = &h->cm;
    if (h->dtype == CHOLMOD_SINGLE) return -1;
    int64_t n = (int64_t)h->L->n;
    cholmod_dense *Bd = cholmod_l_allocate_dense(n, 1, n, CHOLMOD_REAL, cm);
    cholmod_sparse *Bs = cholmod_l_allocate_sparse(n, 1, blen > 0 ? blen : 1, 0, 1, 0, CHOLMOD_PATTERN, cm);
    if (!Bd || !Bs) return -2;
    memcpy(Bd->x, b, sizeof(double) * n);
    int64_t *bp = (int64_t *)Bs->p, *bi = (int64_t *)Bs->i;
    bp[0] = 0;
    bp[1] = blen;
    if (blen > 0) memcpy(bi, bset, sizeof(int64_t) * blen);
    cholmod_dense *X = NULL, *Y = NULL, *E = NULL;
    cholmod_sparse *Xs = NULL;
    int ok = cholmod_l_solve2(sys, h->L, Bd, Bs, &X, &Xs, &Y, &E, cm);
    int rc = ok ? 0 : (cm->status != 0 ? cm->status : -99);
    if (ok) {
        const int64_t *xp = (const int64_t *)Xs->p, *xi = (const int64_t *)Xs->i;
        const double *xv = (const double *)X->x;
        *xlen = xp[1];
        for (int64_t p = 0; p < xp[1]; p++) {
            xset[p] = xi[p];
            x[xi[p]] = xv[xi[p]];
        }
    }
    cholmod_l_free_dense(&Bd, cm);
    cholmod_l_free_sparse(&Bs, cm);
    cholmod_l_free_dense(&X, cm);
    cholmod_l_free_dense(&Y, cm);
    cholmod_l_free_dense(&E, cm);
    cholmod_l_free_sparse(&Xs, cm);
    return rc;
}

// R = B - A*X using the reference's cholmod_l_sdmult on the double-valued A.
int ref_residual(void *vh, const double *B, const double *X, int64_t nrhs, double *R) {
    ref_handle *h = (ref_handle *)vh;
    cholmod_common *cm = &h->cm;
    if (h->dtype == CHOLMOD_SINGLE) return -1;
    int64_t n = (int64_t)h->L->n;
    cholmod_dense *Xd = cholmod_l_allocate_dense(n, nrhs, n, CHOLMOD_REAL, cm);
    cholmod_dense *Rd = cholmod_l_allocate_dense(n, nrhs, n, CHOLMOD_REAL, cm);
    if (!Xd || !Rd) return -2;
    memcpy(Xd->x, X, sizeof(double) * n * nrhs);
    memcpy(Rd->x, B, sizeof(double) * n * nrhs);
    double one[2] = {1, 0}, minusone[2] = {-1, 0};
    cholmod_l_sdmult(h->A, 0, minusone, one, Xd, Rd, cm);
    memcpy(R, Rd->x, sizeof(double) * n * nrhs);
    cholmod_l_free_dense(&Xd, cm);
    cholmod_l_free_dense(&Rd, cm);
    return 0;
}

// cholmod_l_updown (Modify/cholmod_updown.c:74) on the handle's factor with
// C (n x ncol, CSC, sorted rows in the factor's order).  The first call needs
// the supernodal factor of ref_factorize; the reference turns it into a
// simplicial LDL' one (cholmod_updown.c:478-491), which later calls keep
// modifying.  Lx receives the result as LL' values in the supernodal layout
// the factor had: l_jj = sqrt(d_j), l_ij * l_jj (as the reference's
// LDL' -> LL' conversion computes them).  modfl = Common->modfl,
// ndbounds = Common->ndbounds_hit, t = seconds in cholmod_l_updown.  Returns
// the cholmod status, -3 if a column's pattern no longer matches the
// supernodal one (new fill), -4 without a supernodal structure.
int ref_updown(void *vh, int update, int64_t ncol, const int64_t *Cp, const int64_t *Ci, const double *Cx,
               double dbound, double *Lx, double *modfl, int64_t *ndbounds, double *t) {
    ref_handle *h = (ref_handle *)vh;
    cholmod_common *cm = &h->cm;
    cholmod_factor *L = h->L;
    int64_t n = (int64_t)L->n;
    if (!h->Super) {
        if (!L->is_super) return -4;
        int64_t ns = (int64_t)L->nsuper;
        h->nsuper = ns;
        h->Super = malloc(sizeof(int64_t) * (ns + 1));
        h->Lpi = malloc(sizeof(int64_t) * (ns + 1));
        h->Lpx = malloc(sizeof(int64_t) * (ns + 1));
        h->Ls = malloc(sizeof(int64_t) * (L->ssize > 0 ? L->ssize : 1));
        memcpy(h->Super, L->super, sizeof(int64_t) * (ns + 1));
        memcpy(h->Lpi, L->pi, sizeof(int64_t) * (ns + 1));
        memcpy(h->Lpx, L->px, sizeof(int64_t) * (ns + 1));
        memcpy(h->Ls, L->s, sizeof(int64_t) * L->ssize);
    }
    int xt = CHOLMOD_REAL + (h->dtype == CHOLMOD_SINGLE ? CHOLMOD_SINGLE : CHOLMOD_DOUBLE);
    int64_t nz = Cp[ncol];
    cholmod_sparse *C = cholmod_l_allocate_sparse(n, ncol, nz > 0 ? nz : 1, 1, 1, 0, xt, cm);
    if (!C) return -2;
    memcpy(C->p, Cp, sizeof(int64_t) * (ncol + 1));
    memcpy(C->i, Ci, sizeof(int64_t) * nz);
    if (h->dtype == CHOLMOD_SINGLE) {
        float *x = (float *)C->x;
        for (int64_t p = 0; p < nz; p++) x[p] = (float)Cx[p];
    } else {
        memcpy(C->x, Cx, sizeof(double) * nz);
    }
    cm->dbound = dbound;
    cm->sbound = (float)dbound;
    cm->ndbounds_hit = 0;
    double t0 = now_s();
    int ok = cholmod_l_updown(update, C, L, cm);
    if (t) *t = now_s() - t0;
    cholmod_l_free_sparse(&C, cm);
    if (modfl) *modfl = cm->modfl;
    if (ndbounds) *ndbounds = (int64_t)cm->ndbounds_hit;
    if (!ok) return cm->status != 0 ? cm->status : -99;
    const int64_t *Lp = (const int64_t *)L->p, *Lnz = (const int64_t *)L->nz, *Li = (const int64_t *)L->i;
    for (int64_t s = 0; s < h->nsuper; s++) {
        int64_t k1 = h->Super[s], nscol = h->Super[s + 1] - k1;
        int64_t psi = h->Lpi[s], nsrow = h->Lpi[s + 1] - psi;
        for (int64_t jj = 0; jj < nscol; jj++) {
            int64_t j = k1 + jj, p = Lp[j];
            if (Lnz[j] != nsrow - jj || Li[p] != j) return -3;
            double d, l;
            if (h->dtype == CHOLMOD_SINGLE) {
                const float *x = (const float *)L->x;
                float ljj = sqrtf(x[p]);
                Lx[h->Lpx[s] + jj * nsrow + jj] = ljj;
                for (int64_t q = 1; q < nsrow - jj; q++) {
                    if (Li[p + q] != h->Ls[psi + jj + q]) return -3;
                    Lx[h->Lpx[s] + jj * nsrow + jj + q] = (double)(x[p + q] * ljj);
                }
            } else {
                const double *x = (const double *)L->x;
                d = x[p];
                double ljj = sqrt(d);
                Lx[h->Lpx[s] + jj * nsrow + jj] = ljj;
                for (int64_t q = 1; q < nsrow - jj; q++) {
                    if (Li[p + q] != h->Ls[psi + jj + q]) return -3;
                    l = x[p + q];
                    Lx[h->Lpx[s] + jj * nsrow + jj + q] = l * ljj;
                }
            }
        }
    }
    return cm->status;
}

void ref_free(void *vh) {
    ref_handle *h = (ref_handle *)vh;
    if (!h) return;
    free(h->Super), free(h->Lpi), free(h->Lpx), free(h->Ls);
    cholmod_l_free_factor(&h->L, &h->cm);
    cholmod_l_free_sparse(&h->A, &h->cm);
    cholmod_l_finish(&h->cm);
    free(h);
}

// cholmod_l_read_sparse2 (CHOLMOD/Check/cholmod_read.c, cholmod.h) on a file,
// with the reader's defaults (prefer_binary FALSE: symmetric pattern-only
// matrices get diag = 1 + degree and off-diagonals -1).  dims receives nrow,
// ncol, nnz, stype, xtype (CHOLMOD_PATTERN 0 / REAL 1 / COMPLEX 2).  With Ap
// NULL only dims is filled; otherwise the packed, sorted CSC arrays are copied
// (Ax only for a real matrix).  Returns 0, or -1 if the reader fails.
int ref_read_sparse(const char *path, int64_t *dims, int64_t *Ap, int64_t *Ai, double *Ax) {
    cholmod_common cm;
    cholmod_l_start(&cm);
    cm.print = 0;
    FILE *f = fopen(path, "r");
    if (!f) {
        cholmod_l_finish(&cm);
        return -1;
    }
    cholmod_sparse *A = cholmod_l_read_sparse2(f, CHOLMOD_DOUBLE, &cm);
    fclose(f);
    if (!A) {
        cholmod_l_finish(&cm);
        return -1;
    }
    if (!A->packed || !A->sorted) cholmod_l_sort(A, &cm);
    const int64_t *P = (const int64_t *)A->p, *I = (const int64_t *)A->i;
    const int64_t nnz = P[A->ncol];
    dims[0] = (int64_t)A->nrow;
    dims[1] = (int64_t)A->ncol;
    dims[2] = nnz;
    dims[3] = A->stype;
    dims[4] = A->xtype;
    if (Ap) {
        memcpy(Ap, P, (A->ncol + 1) * sizeof(int64_t));
        memcpy(Ai, I, nnz * sizeof(int64_t));
        if (Ax && A->xtype == CHOLMOD_REAL) memcpy(Ax, A->x, nnz * sizeof(double));
    }
    cholmod_l_free_sparse(&A, &cm);
    cholmod_l_finish(&cm);
    return 0;
}

// The raw orderings CHOLMOD's analyze computes before its postorder:
// which = 1: cholmod_l_amd, 2: cholmod_l_metis (postorder FALSE), on tril(A).
// stats[0..2] = Common->fl, Common->lnz, Common->anz after the call (AMD
// sets all three; METIS only anz).
int ref_order(int64_t n, const int64_t *Ap, const int64_t *Ai, int which, int64_t *Perm,
              double *stats) {
    cholmod_common cm;
    cholmod_l_start(&cm);
    int64_t nnz = Ap[n];
    cholmod_sparse *A = cholmod_l_allocate_sparse(n, n, nnz, 1, 1, -1, CHOLMOD_PATTERN, &cm);
    if (!A) { cholmod_l_finish(&cm); return -1; }
    memcpy(A->p, Ap, sizeof(int64_t) * (n + 1));
    memcpy(A->i, Ai, sizeof(int64_t) * nnz);
    cm.fl = cm.lnz = cm.anz = -1;
    int ok = which == 1 ? cholmod_l_amd(A, NULL, 0, Perm, &cm)
                        : cholmod_l_metis(A, NULL, 0, 0, Perm, &cm);
    if (stats) { stats[0] = cm.fl; stats[1] = cm.lnz; stats[2] = cm.anz; }
    cholmod_l_free_sparse(&A, &cm);
    cholmod_l_finish(&cm);
    return ok ? 0 : -2;
}
