/* dropin_flow.c -- a CHOLMOD caller, written against cholmod.h only, that
 * runs the sequence of CHOLMOD/Demo/cholmod_dl_demo.c:287-759 on one matrix
 * file twice side by side: once through the MI355X drop-in (Common->useGPU =
 * 1) and once on the host CHOLMOD (useGPU = 0), and prints what the test
 * compares, one "key value" per line.
 *
 * Linked as: dropin_flow -> libcholmod_mi355x_dropin.so (analyze, factorize,
 * solve, solve2, free_factor, finish, gpu_stats) -> the host CHOLMOD (here the
 * reference compiled from its sources) for everything else.
 *
 *   dropin_flow <matrix file> [supernodal] [gpu|host]
 *
 * After the demo's sequence, a fresh factor takes a rank-2 update with C
 * inside L's pattern (cholmod_l_updown, Modify/cholmod_updown.c:74), a solve
 * against A + P'CC'P, the same downdate (back to A), and an update whose C
 * needs new fill (which the drop-in hands to the host library).
 *
 * supernodal: Common->supernodal = CHOLMOD_SUPERNODAL (default: CHOLMOD_AUTO).
 * gpu / host: run only that side (default: both, and compare).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cholmod.h"

typedef struct {
    cholmod_common c;
    cholmod_sparse *A;
    cholmod_factor *L;
    cholmod_dense *B, *X;
    double Lx_copy_n;
    double *Lx;            /* L->x right after the factorization (before any Bset solve) */
    int is_super, status, minor, nsuper;
    double rcond, resid, resid_ir, fl, lnz;
    double bset_err, solve2_err;   /* max |difference| / max |x| */
    int ordering, is_super_after_bset;
    int selected, called_nd;             /* Common->selected, called_nd after analyze */
    double anz, method_lnz[3];           /* Common->anz, method[0..2].lnz */
    /* update / downdate */
    int ud_done, ud_status, ud_is_super, fill_status, fill_is_super;
    double ud_modfl, ud_resid, back_err, fill_modfl, fill_resid;
    cholmod_dense *Xup, *Xfill;
} side;

static double norm_inf_dense(cholmod_dense *X) {
    double m = 0;
    double *x = (double *)X->x;
    for (size_t j = 0; j < X->ncol; j++)
        for (size_t i = 0; i < X->nrow; i++) m = fmax(m, fabs(x[i + j * X->d]));
    return m;
}

/* |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf), the demo's residual (cholmod_dl_demo.c:614-617) */
static double resid_of(side *s, cholmod_dense *X) {
    double one[2] = {1, 0}, minusone[2] = {-1, 0};
    cholmod_dense *R = cholmod_l_copy_dense(s->B, &s->c);
    cholmod_l_sdmult(s->A, 0, minusone, one, X, R, &s->c);
    double r = cholmod_l_norm_dense(R, 0, &s->c);
    double anorm = cholmod_l_norm_sparse(s->A, 0, &s->c);
    double v = r / (anorm * norm_inf_dense(X) + norm_inf_dense(s->B) + (s->A->nrow == 0));
    cholmod_l_free_dense(&R, &s->c);
    return v;
}

/* |b - (A + P'CC'P) x|_inf / (|A|_inf |x|_inf + |b|_inf) */
static double resid_mod(side *s, cholmod_sparse *C, cholmod_dense *X) {
    const size_t n = s->A->nrow;
    double one[2] = {1, 0}, minusone[2] = {-1, 0};
    cholmod_dense *R = cholmod_l_copy_dense(s->B, &s->c);
    cholmod_l_sdmult(s->A, 0, minusone, one, X, R, &s->c);
    const int64_t *Perm = (const int64_t *)s->L->Perm, *Cp = (const int64_t *)C->p, *Ci = (const int64_t *)C->i;
    const double *Cx = (const double *)C->x, *x = (const double *)X->x;
    double *r = (double *)R->x;
    for (size_t c = 0; c < C->ncol; c++) {
        double t = 0;                      /* C(:,c)' P x */
        for (int64_t p = Cp[c]; p < Cp[c + 1]; p++) t += Cx[p] * x[Perm[Ci[p]]];
        for (int64_t p = Cp[c]; p < Cp[c + 1]; p++) r[Perm[Ci[p]]] -= Cx[p] * t;
    }
    double v = cholmod_l_norm_dense(R, 0, &s->c) /
               (cholmod_l_norm_sparse(s->A, 0, &s->c) * norm_inf_dense(X) + norm_inf_dense(s->B) + (n == 0));
    cholmod_l_free_dense(&R, &s->c);
    return v;
}

/* C (n x ncol) with column c starting at row ks[c] and taking every other row
 * of the pattern of L(:, ks[c]) (supernodal L); the values are fixed numbers. */
static cholmod_sparse *pattern_c(side *s, const int64_t *ks, int ncol, double scale) {
    const cholmod_factor *L = s->L;
    const int64_t *Super = (const int64_t *)L->super, *Lpi = (const int64_t *)L->pi, *Ls = (const int64_t *)L->s;
    cholmod_sparse *C = cholmod_l_allocate_sparse(L->n, ncol, L->n * ncol + 1, 1, 1, 0,
                                                  CHOLMOD_REAL + CHOLMOD_DOUBLE, &s->c);
    int64_t *Cp = (int64_t *)C->p, *Ci = (int64_t *)C->i;
    double *Cx = (double *)C->x;
    Cp[0] = 0;
    for (int c = 0; c < ncol; c++) {
        const int64_t k = ks[c];
        int64_t sn = 0;
        while (Super[sn + 1] <= k) sn++;
        int64_t q = Cp[c];
        for (int64_t p = Lpi[sn] + (k - Super[sn]); p < Lpi[sn + 1]; p += 2) {
            Ci[q] = Ls[p];
            Cx[q] = scale * (0.3 + 0.1 * (double)(q % 7));
            q++;
        }
        Cp[c + 1] = q;
    }
    return C;
}

static void updown_step(side *s) {
    const int64_t n = (int64_t)s->A->nrow;
    /* a row outside the pattern of L(:, 0), for the update that needs fill
       (found now: the host library leaves L simplicial after its first updown) */
    int64_t out = -1;
    {
        const int64_t *Lpi = (const int64_t *)s->L->pi, *Ls = (const int64_t *)s->L->s;
        for (int64_t r = 1, p = Lpi[0] + 1; r < n; r++) {
            while (p < Lpi[1] && Ls[p] < r) p++;
            if (p == Lpi[1] || Ls[p] != r) {
                out = r;
                break;
            }
        }
    }
    int64_t ks[2] = {n / 3, (2 * n) / 3};
    cholmod_sparse *C = pattern_c(s, ks, 2, 1.0);
    s->ud_status = cholmod_l_updown(1, C, s->L, &s->c) ? s->c.status : -100 + s->c.status;
    s->ud_modfl = s->c.modfl;
    s->ud_is_super = s->L->is_super;
    s->Xup = cholmod_l_solve(CHOLMOD_A, s->L, s->B, &s->c);
    s->ud_resid = s->Xup ? resid_mod(s, C, s->Xup) : -1;
    /* the same downdate: back to A */
    cholmod_l_updown(0, C, s->L, &s->c);
    cholmod_dense *Xb = cholmod_l_solve(CHOLMOD_A, s->L, s->B, &s->c);
    s->back_err = 0;
    const double xmax = norm_inf_dense(s->X) + (n == 0);
    for (int64_t i = 0; Xb && i < n; i++)
        s->back_err = fmax(s->back_err, fabs(((double *)Xb->x)[i] - ((double *)s->X->x)[i]) / xmax);
    cholmod_l_free_dense(&Xb, &s->c);
    cholmod_l_free_sparse(&C, &s->c);
    /* an update that needs new fill: row 0 and a row outside the pattern of L(:, 0) */
    s->fill_status = -1000;
    {
        if (out > 0) {
            cholmod_sparse *F = cholmod_l_allocate_sparse(n, 1, 2, 1, 1, 0, CHOLMOD_REAL + CHOLMOD_DOUBLE, &s->c);
            ((int64_t *)F->p)[0] = 0;
            ((int64_t *)F->p)[1] = 2;
            ((int64_t *)F->i)[0] = 0;
            ((int64_t *)F->i)[1] = out;
            ((double *)F->x)[0] = 0.7;
            ((double *)F->x)[1] = -0.4;
            s->fill_status = cholmod_l_updown(1, F, s->L, &s->c) ? s->c.status : -100 + s->c.status;
            s->fill_modfl = s->c.modfl;
            s->fill_is_super = s->L->is_super;
            s->Xfill = cholmod_l_solve(CHOLMOD_A, s->L, s->B, &s->c);
            s->fill_resid = s->Xfill ? resid_mod(s, F, s->Xfill) : -1;
            cholmod_l_free_sparse(&F, &s->c);
        }
    }
    s->ud_done = 1;
}

static int run(side *s, const char *path, int gpu, int supernodal) {
    cholmod_l_start(&s->c);
    s->c.print = 1;
    s->c.useGPU = gpu;
    s->c.maxGpuMemBytes = (size_t)1 << 30;     /* a 1 GiB pool, not the whole device */
    if (supernodal) s->c.supernodal = CHOLMOD_SUPERNODAL;
    FILE *f = fopen(path, "r");
    if (!f) return 1;
    s->A = cholmod_l_read_sparse(f, &s->c);
    fclose(f);
    if (!s->A) return 2;
    const size_t n = s->A->nrow;
    s->B = cholmod_l_zeros(n, 1, CHOLMOD_REAL + CHOLMOD_DOUBLE, &s->c);
    double *b = (double *)s->B->x;
    for (size_t i = 0; i < n; i++) b[i] = 1 + (double)i / (double)n;

    s->L = cholmod_l_analyze(s->A, &s->c);
    if (!s->L) return 3;
    s->fl = s->c.fl;
    s->lnz = s->c.lnz;
    s->ordering = s->L->ordering;
    s->selected = s->c.selected;
    s->called_nd = s->c.called_nd;
    s->anz = s->c.anz;
    for (int m = 0; m < 3; m++) s->method_lnz[m] = s->c.method[m].lnz;
    for (int t = 0; t < 3; t++) cholmod_l_factorize(s->A, s->L, &s->c);
    s->status = s->c.status;
    s->minor = (int)s->L->minor;
    s->is_super = s->L->is_super;
    s->nsuper = (int)s->L->nsuper;
    if (s->L->is_super && s->L->x) {
        s->Lx_copy_n = (double)s->L->xsize;
        s->Lx = (double *)malloc(s->L->xsize * sizeof(double));
        memcpy(s->Lx, s->L->x, s->L->xsize * sizeof(double));
    }
    s->rcond = cholmod_l_rcond(s->L, &s->c);
    if (s->status != CHOLMOD_OK) return 0;

    /* solve, and solve again with a reused X and workspaces (solve2) */
    s->X = cholmod_l_solve(CHOLMOD_A, s->L, s->B, &s->c);
    s->resid = resid_of(s, s->X);
    cholmod_dense *X2 = NULL, *Y = NULL, *E = NULL;
    for (int t = 0; t < 3; t++) cholmod_l_solve2(CHOLMOD_A, s->L, s->B, NULL, &X2, NULL, &Y, &E, &s->c);
    s->solve2_err = 0;
    const double xmax = norm_inf_dense(s->X) + (n == 0);
    for (size_t i = 0; i < n; i++)
        s->solve2_err = fmax(s->solve2_err, fabs(((double *)X2->x)[i] - ((double *)s->X->x)[i]) / xmax);

    /* sparse right-hand sides: x(Xset) for b = e_i * (3.1 i + 0.9) */
    cholmod_dense *B2 = cholmod_l_zeros(n, 1, CHOLMOD_REAL + CHOLMOD_DOUBLE, &s->c), *X3 = NULL, *Xfull = NULL;
    cholmod_sparse *Bset = cholmod_l_allocate_sparse(n, 1, 1, 0, 1, 0, CHOLMOD_PATTERN + CHOLMOD_DOUBLE, &s->c);
    cholmod_sparse *Xset = NULL;
    ((int64_t *)Bset->p)[0] = 0;
    ((int64_t *)Bset->p)[1] = 1;
    s->bset_err = 0;
    for (size_t i = 0; i < (n < 20 ? n : 20); i++) {
        ((int64_t *)Bset->i)[0] = (int64_t)i;
        ((double *)B2->x)[i] = 3.1 * (double)i + 0.9;
        cholmod_l_solve2(CHOLMOD_A, s->L, B2, NULL, &Xfull, NULL, &Y, &E, &s->c);
        cholmod_l_solve2(CHOLMOD_A, s->L, B2, Bset, &X3, &Xset, &Y, &E, &s->c);
        const int64_t xlen = ((int64_t *)Xset->p)[1];
        const double fmaxv = norm_inf_dense(Xfull) + (n == 0);
        for (int64_t k = 0; k < xlen; k++) {
            const int64_t j = ((int64_t *)Xset->i)[k];
            s->bset_err = fmax(s->bset_err, fabs(((double *)X3->x)[j] - ((double *)Xfull->x)[j]) / fmaxv);
        }
        ((double *)B2->x)[i] = 0;
    }
    s->is_super_after_bset = s->L->is_super;
    cholmod_l_free_dense(&B2, &s->c);
    cholmod_l_free_dense(&X2, &s->c);
    cholmod_l_free_dense(&X3, &s->c);
    cholmod_l_free_dense(&Xfull, &s->c);
    cholmod_l_free_dense(&Y, &s->c);
    cholmod_l_free_dense(&E, &s->c);
    cholmod_l_free_sparse(&Bset, &s->c);
    cholmod_l_free_sparse(&Xset, &s->c);

    /* one step of iterative refinement (cholmod_dl_demo.c:627-663) */
    double one[2] = {1, 0}, minusone[2] = {-1, 0};
    cholmod_dense *R = cholmod_l_copy_dense(s->B, &s->c);
    cholmod_l_sdmult(s->A, 0, minusone, one, s->X, R, &s->c);
    cholmod_dense *R2 = cholmod_l_solve(CHOLMOD_A, s->L, R, &s->c);
    cholmod_dense *Xir = cholmod_l_copy_dense(s->X, &s->c);
    for (size_t i = 0; i < n; i++) ((double *)Xir->x)[i] += ((double *)R2->x)[i];
    s->resid_ir = resid_of(s, Xir);
    cholmod_l_free_dense(&R, &s->c);
    cholmod_l_free_dense(&R2, &s->c);
    cholmod_l_free_dense(&Xir, &s->c);

    /* a fresh factor for the update / downdate (the Bset solves left L simplicial) */
    cholmod_l_free_factor(&s->L, &s->c);
    s->L = cholmod_l_analyze(s->A, &s->c);
    if (!s->L) return 4;
    cholmod_l_factorize(s->A, s->L, &s->c);
    if (s->c.status == CHOLMOD_OK && s->L->is_super && n > 0) updown_step(s);
    return 0;
}

static void report(const char *tag, side *s) {
    printf("%s_status %d\n%s_minor %d\n%s_is_super %d\n%s_nsuper %d\n%s_ordering %d\n", tag, s->status, tag,
           s->minor, tag, s->is_super, tag, s->nsuper, tag, s->ordering);
    printf("%s_fl %.17g\n%s_lnz %.17g\n%s_rcond %.17g\n%s_resid %.6e\n%s_resid_ir %.6e\n", tag, s->fl, tag,
           s->lnz, tag, s->rcond, tag, s->resid, tag, s->resid_ir);
    printf("%s_bset_err %.6e\n%s_solve2_err %.6e\n%s_is_super_after_bset %d\n", tag, s->bset_err, tag,
           s->solve2_err, tag, s->is_super_after_bset);
    printf("%s_selected %d\n%s_called_nd %d\n%s_anz %.17g\n", tag, s->selected, tag, s->called_nd, tag, s->anz);
    printf("%s_method_lnz %.17g %.17g %.17g\n", tag, s->method_lnz[0], s->method_lnz[1], s->method_lnz[2]);
    printf("%s_gpu_launches %d\n%s_gpu_time %.6e\n%s_pool_bytes %.0f\n", tag, s->c.gpuNumKernelLaunches, tag,
           s->c.gpuKernelTime, tag, (double)s->c.dev_mempool_size);
    printf("%s_gpu_calls %.0f %.0f %.0f %.0f\n", tag, (double)s->c.cholmod_gpu_syrk_calls,
           (double)s->c.cholmod_gpu_gemm_calls, (double)s->c.cholmod_gpu_potrf_calls,
           (double)s->c.cholmod_gpu_trsm_calls);
    if (s->ud_done) {
        printf("%s_ud_status %d\n%s_ud_modfl %.17g\n%s_ud_is_super %d\n%s_ud_resid %.6e\n%s_back_err %.6e\n", tag,
               s->ud_status, tag, s->ud_modfl, tag, s->ud_is_super, tag, s->ud_resid, tag, s->back_err);
        printf("%s_fill_status %d\n%s_fill_modfl %.17g\n%s_fill_is_super %d\n%s_fill_resid %.6e\n", tag,
               s->fill_status, tag, s->fill_modfl, tag, s->fill_is_super, tag, s->fill_resid);
    }
}

static void finish(side *s) {
    cholmod_l_free_dense(&s->Xup, &s->c);
    cholmod_l_free_dense(&s->Xfill, &s->c);
    cholmod_l_free_factor(&s->L, &s->c);
    cholmod_l_free_dense(&s->X, &s->c);
    cholmod_l_free_dense(&s->B, &s->c);
    cholmod_l_free_sparse(&s->A, &s->c);
    cholmod_l_finish(&s->c);
    free(s->Lx);
}

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <matrix file> [supernodal] [gpu|host]\n", argv[0]);
        return 2;
    }
    int supernodal = 0, do_gpu = 1, do_host = 1;
    for (int k = 2; k < argc; k++) {
        if (!strcmp(argv[k], "supernodal")) supernodal = 1;
        if (!strcmp(argv[k], "gpu")) do_host = 0;
        if (!strcmp(argv[k], "host")) do_gpu = 0;
    }
    static side g, h;
    if (do_gpu) {
        int rc = run(&g, argv[1], 1, supernodal);
        printf("gpu_rc %d\n", rc);
        if (rc) return 1;
        report("gpu", &g);
    }
    if (do_host) {
        int rc = run(&h, argv[1], 0, supernodal);
        printf("host_rc %d\n", rc);
        if (rc) return 1;
        report("host", &h);
    }
    if (do_gpu && do_host) {
        /* the two factors side by side (before either side's Bset solves) */
        double d = 0, m = 0;
        if (g.Lx && h.Lx && g.Lx_copy_n == h.Lx_copy_n) {
            for (size_t k = 0; k < (size_t)g.Lx_copy_n; k++) {
                d = fmax(d, fabs(g.Lx[k] - h.Lx[k]));
                m = fmax(m, fabs(h.Lx[k]));
            }
            printf("L_rel_diff %.6e\n", m > 0 ? d / m : d);
        } else {
            printf("L_rel_diff -1\n");
        }
        if (g.X && h.X) {
            double num = 0, den = 0;
            for (size_t i = 0; i < g.X->nrow; i++) {
                const double a = ((double *)g.X->x)[i], b = ((double *)h.X->x)[i];
                num += (a - b) * (a - b);
                den += b * b;
            }
            printf("x_rel_diff %.6e\n", den > 0 ? sqrt(num / den) : sqrt(num));
        }
        if (g.Xup && h.Xup) {
            double num = 0, den = 0, fnum = 0, fden = 0;
            for (size_t i = 0; i < g.Xup->nrow; i++) {
                const double a = ((double *)g.Xup->x)[i], b = ((double *)h.Xup->x)[i];
                num += (a - b) * (a - b);
                den += b * b;
                if (g.Xfill && h.Xfill) {
                    const double fa = ((double *)g.Xfill->x)[i], fb = ((double *)h.Xfill->x)[i];
                    fnum += (fa - fb) * (fa - fb);
                    fden += fb * fb;
                }
            }
            printf("xup_rel_diff %.6e\n", den > 0 ? sqrt(num / den) : sqrt(num));
            if (g.Xfill && h.Xfill) printf("xfill_rel_diff %.6e\n", fden > 0 ? sqrt(fnum / fden) : sqrt(fnum));
        }
        printf("perm_equal %d\n", g.L && h.L && g.L->n == h.L->n &&
               !memcmp(g.L->Perm, h.L->Perm, g.L->n * sizeof(int64_t)));
    }
    if (do_gpu) {
        cholmod_l_gpu_stats(&g.c);
        finish(&g);
    }
    if (do_host) finish(&h);
    return 0;
}
