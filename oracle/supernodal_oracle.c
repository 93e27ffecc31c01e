// supernodal_oracle.c -- CPU restatement of CHOLMOD's supernodal numeric
// factorization and supernodal triangular solves, in plain C with naive dense
// loops in place of BLAS/LAPACK.
//
// ORACLE / TEST INFRASTRUCTURE ONLY.  Built into oracle/_build/liboracle.so by
// oracle/Makefile; loaded by tests/, the smoke check and nothing else.  It is
// the checker for the HIP path, never part of it.
//
// Followed reference code (paths relative to the SuiteSparse root):
//   numeric: CHOLMOD/Supernodal/t_cholmod_super_numeric_worker.c
//     :335-379  per supernode: zero Lx(s), Map[Ls[psi+k]] = k
//     :438-507  scatter tril(A(:,k1:k2-1)) into s (assignment, rows i >= k)
//     :510-520  beta added on the diagonal
//     :588-705  descendants d of s (Head/Next lists, Lpos): ndrow1, ndrow2
//     :769-834  C1 = L1 L1' (dsyrk lower), C2 = L2 L1' (dgemm N,C)
//     :896-930  RelativeMap; Lx[psx + Rel[j]*nsrow + Rel[i]] -= C[i + ndrow2*j], i >= j
//     :950-967  move d to the list of its next ancestor
//     :1023-1027 dpotrf("L", nscol2, Lx+psx, nsrow)
//     :1084-1146 not positive definite: minor, zero tail, repeat s with info-1 cols
//     :1175-1181 dtrsm("R","L","C","N", nsrow2, nscol2): L2 = S2 L1^{-T}
//     :1222-1235 link s into its parent's list; :1269 L->minor = n
//   solve: CHOLMOD/Supernodal/t_cholmod_super_solve_worker.c
//     :20-303   lsolve: gather E, x1 = L1 \ x1, E -= L2 x1, scatter
//     :309-580  ltsolve: gather E, x1 -= L2' E, x1 = L1' \ x1
//   permutation: CHOLMOD/Cholesky/t_cholmod_psolve_worker.c:40-110 (Y = P B, X = P' Y)
//   update/downdate: CHOLMOD/Modify/t_cholmod_updown_numkr.c:294-372 on the
//     paths of Modify/t_cholmod_updown_worker.c (oracle_updown, below)
//
// Every integer array is int64 (CHOLMOD's "_l_" interface).

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t Int;

// Naive lower Cholesky of the leading n x n block of column-major A (lda).
// Right-looking, column by column.  Returns LAPACK-style info: 0 on success,
// else the 1-based column whose pivot was not positive.
static Int potrf_lower(Int n, double *A, Int lda) {
    for (Int j = 0; j < n; j++) {
        double d = A[j + j * lda];
        for (Int k = 0; k < j; k++) d -= A[j + k * lda] * A[j + k * lda];
        if (!(d > 0)) {     // also catches NaN
            A[j + j * lda] = d;
            return j + 1;
        }
        d = sqrt(d);
        A[j + j * lda] = d;
        for (Int i = j + 1; i < n; i++) {
            double v = A[i + j * lda];
            for (Int k = 0; k < j; k++) v -= A[i + k * lda] * A[j + k * lda];
            A[i + j * lda] = v / d;
        }
    }
    return 0;
}

// B (m x n, ldb) = B * L^{-T}, L lower n x n (lda), non-unit.
static void trsm_rlcn(Int m, Int n, const double *L, Int lda, double *B, Int ldb) {
    for (Int j = 0; j < n; j++) {
        double djj = L[j + j * lda];
        for (Int i = 0; i < m; i++) {
            double v = B[i + j * ldb];
            for (Int k = 0; k < j; k++) v -= B[i + k * ldb] * L[j + k * lda];
            B[i + j * ldb] = v / djj;
        }
    }
}

// Supernodal numeric factorization.
//   S = tril(P A P') in CSC (Sp, Si, Sx), column k holds rows i >= k.
//   beta: added to every diagonal entry (cholmod_factorize_p's beta[0]).
//   Lx (xsize) receives the factor; *minor = n on success.
// Returns 0 on success, 1 if not positive definite (CHOLMOD_NOT_POSDEF).
int oracle_super_numeric(Int n, Int nsuper, const Int *Super, const Int *Lpi,
                         const Int *Lpx, const Int *Ls, const Int *Sp, const Int *Si,
                         const double *Sx, double beta, double *Lx, Int *minor) {
    Int xsize = Lpx[nsuper];
    Int *Map = (Int *)malloc(sizeof(Int) * (n > 0 ? n : 1));
    Int *SuperMap = (Int *)malloc(sizeof(Int) * (n > 0 ? n : 1));
    Int *Head = (Int *)malloc(sizeof(Int) * (nsuper + 1));
    Int *Next = (Int *)malloc(sizeof(Int) * (nsuper + 1));
    Int *Lpos = (Int *)malloc(sizeof(Int) * (nsuper + 1));
    Int *Next_save = (Int *)malloc(sizeof(Int) * (nsuper + 1));
    Int *Lpos_save = (Int *)malloc(sizeof(Int) * (nsuper + 1));
    Int *Rel = (Int *)malloc(sizeof(Int) * (n > 0 ? n : 1));
    Int maxc = 1;
    for (Int s = 0; s < nsuper; s++) {
        Int nsr = Lpi[s + 1] - Lpi[s];
        if (nsr * nsr > maxc) maxc = nsr * nsr;
    }
    double *C = (double *)malloc(sizeof(double) * maxc);
    for (Int s = 0; s < nsuper; s++) {
        for (Int k = Super[s]; k < Super[s + 1]; k++) SuperMap[k] = s;
        Head[s] = -1;
        Next[s] = -1;
        Lpos[s] = 0;
    }
    for (Int i = 0; i < n; i++) Map[i] = -1;
    int status = 0;
    int repeat = 0;
    Int nscol_new = 0;
    *minor = n;

    for (Int s = 0; s < nsuper; s++) {
        Int k1 = Super[s], k2 = Super[s + 1], nscol = k2 - k1;
        Int psi = Lpi[s], psx = Lpx[s], nsrow = Lpi[s + 1] - psi;
        for (Int p = psx; p < psx + nsrow * nscol; p++) Lx[p] = 0;
        for (Int k = 0; k < nsrow; k++) Map[Ls[psi + k]] = k;
        for (Int k = k1; k < k2; k++) {
            for (Int p = Sp[k]; p < Sp[k + 1]; p++) {
                Int i = Si[p];
                if (i >= k) {
                    Int imap = Map[i];
                    if (imap >= 0 && imap < nsrow) Lx[psx + imap + (k - k1) * nsrow] = Sx[p];
                }
            }
        }
        if (beta != 0) {
            for (Int k = 0; k < nscol; k++) Lx[psx + k + k * nsrow] += beta;
        }
        if (!repeat) {
            for (Int d = Head[s]; d != -1; d = Next[d]) { Lpos_save[d] = Lpos[d]; Next_save[d] = Next[d]; }
        } else {
            for (Int d = Head[s]; d != -1; d = Next[d]) { Lpos[d] = Lpos_save[d]; Next[d] = Next_save[d]; }
        }
        Int dnext = Head[s];
        while (dnext != -1) {
            Int d = dnext;
            Int ndcol = Super[d + 1] - Super[d];
            Int pdi = Lpi[d], pdx = Lpx[d], pdend = Lpi[d + 1], ndrow = pdend - pdi;
            Int p = Lpos[d], pdi1 = pdi + p, pdx1 = pdx + p, pdi2;
            for (pdi2 = pdi1; pdi2 < pdend && Ls[pdi2] < k2; pdi2++) ;
            Int ndrow1 = pdi2 - pdi1, ndrow2 = pdend - pdi1;
            // C (ndrow2 x ndrow1, ld ndrow2) = L(d rows pdi1.., :) * L(d rows pdi1..pdi2-1, :)'
            // lower part of the top ndrow1 x ndrow1 block (dsyrk "L") and all of the rest (dgemm)
            for (Int j = 0; j < ndrow1; j++) {
                for (Int i = j; i < ndrow2; i++) {
                    double c = 0;
                    for (Int k = 0; k < ndcol; k++)
                        c += Lx[pdx1 + i + k * ndrow] * Lx[pdx1 + j + k * ndrow];
                    C[i + ndrow2 * j] = c;
                }
            }
            for (Int i = 0; i < ndrow2; i++) Rel[i] = Map[Ls[pdi1 + i]];
            for (Int j = 0; j < ndrow1; j++) {
                Int px = psx + Rel[j] * nsrow;
                for (Int i = j; i < ndrow2; i++) Lx[px + Rel[i]] -= C[i + ndrow2 * j];
            }
            dnext = Next[d];
            if (!repeat) {
                Lpos[d] = pdi2 - pdi;
                if (Lpos[d] < ndrow) {
                    Int anc = SuperMap[Ls[pdi2]];
                    Next[d] = Head[anc];
                    Head[anc] = d;
                }
            }
        }
        Int nscol2 = repeat ? nscol_new : nscol;
        Int info = potrf_lower(nscol2, Lx + psx, nsrow);
        if (repeat) {
            info = 0;
            for (Int p = psx + nsrow * nscol_new; p < psx + nsrow * nscol; p++) Lx[p] = 0;
        }
        if (info != 0) {
            status = 1;
            *minor = k1 + info - 1;
            for (Int ss = s + 1; ss < nsuper; ss++) Head[ss] = -1;
            for (Int p = psx; p < xsize; p++) Lx[p] = 0;
            if (info == 1) break;
            repeat = 1;
            nscol_new = info - 1;
            s--;
            continue;
        }
        Int nsrow2 = nsrow - nscol2;
        if (nsrow2 > 0) {
            trsm_rlcn(nsrow2, nscol2, Lx + psx, nsrow, Lx + psx + nscol2, nsrow);
            if (!repeat) {
                Lpos[s] = nscol;
                Int sp = SuperMap[Ls[psi + nscol]];
                Next[s] = Head[sp];
                Head[sp] = s;
            }
        }
        Head[s] = -1;
        if (repeat) break;
    }
    free(Map); free(SuperMap); free(Head); free(Next); free(Lpos);
    free(Next_save); free(Lpos_save); free(Rel); free(C);
    return status;
}

// Forward solve L X = B in place (X is n x nrhs, column-major, ld n).
void oracle_super_lsolve(Int n, Int nsuper, const Int *Super, const Int *Lpi, const Int *Lpx,
                         const Int *Ls, const double *Lx, double *X, Int nrhs) {
    Int maxe = 1;
    for (Int s = 0; s < nsuper; s++) {
        Int e = (Lpi[s + 1] - Lpi[s]) - (Super[s + 1] - Super[s]);
        if (e > maxe) maxe = e;
    }
    double *E = (double *)malloc(sizeof(double) * maxe);
    for (Int r = 0; r < nrhs; r++) {
        double *x = X + r * n;
        for (Int s = 0; s < nsuper; s++) {
            Int k1 = Super[s], nscol = Super[s + 1] - k1;
            Int psi = Lpi[s], psx = Lpx[s], nsrow = Lpi[s + 1] - psi, nsrow2 = nsrow - nscol;
            const double *L1 = Lx + psx;
            for (Int ii = 0; ii < nsrow2; ii++) E[ii] = x[Ls[psi + nscol + ii]];
            for (Int j = 0; j < nscol; j++) {            // x1 = L1 \ x1
                double v = x[k1 + j] / L1[j + j * nsrow];
                x[k1 + j] = v;
                for (Int i = j + 1; i < nscol; i++) x[k1 + i] -= L1[i + j * nsrow] * v;
            }
            for (Int j = 0; j < nscol; j++) {            // E -= L2 x1
                double v = x[k1 + j];
                for (Int ii = 0; ii < nsrow2; ii++) E[ii] -= L1[nscol + ii + j * nsrow] * v;
            }
            for (Int ii = 0; ii < nsrow2; ii++) x[Ls[psi + nscol + ii]] = E[ii];
        }
    }
    free(E);
}

// Backward solve L' X = B in place.
void oracle_super_ltsolve(Int n, Int nsuper, const Int *Super, const Int *Lpi, const Int *Lpx,
                          const Int *Ls, const double *Lx, double *X, Int nrhs) {
    Int maxe = 1;
    for (Int s = 0; s < nsuper; s++) {
        Int e = (Lpi[s + 1] - Lpi[s]) - (Super[s + 1] - Super[s]);
        if (e > maxe) maxe = e;
    }
    double *E = (double *)malloc(sizeof(double) * maxe);
    for (Int r = 0; r < nrhs; r++) {
        double *x = X + r * n;
        for (Int s = nsuper - 1; s >= 0; s--) {
            Int k1 = Super[s], nscol = Super[s + 1] - k1;
            Int psi = Lpi[s], psx = Lpx[s], nsrow = Lpi[s + 1] - psi, nsrow2 = nsrow - nscol;
            const double *L1 = Lx + psx;
            for (Int ii = 0; ii < nsrow2; ii++) E[ii] = x[Ls[psi + nscol + ii]];
            for (Int j = 0; j < nscol; j++) {            // x1 -= L2' E
                double v = 0;
                for (Int ii = 0; ii < nsrow2; ii++) v += L1[nscol + ii + j * nsrow] * E[ii];
                x[k1 + j] -= v;
            }
            for (Int j = nscol - 1; j >= 0; j--) {       // x1 = L1' \ x1
                double v = x[k1 + j];
                for (Int i = j + 1; i < nscol; i++) v -= L1[i + j * nsrow] * x[k1 + i];
                x[k1 + j] = v / L1[j + j * nsrow];
            }
        }
    }
    free(E);
}

// X = A \ B for A = P' L L' P: Y = P B, lsolve, ltsolve, X = P' Y.
// Perm[k] = original index of the k-th pivot.  B, X column-major n x nrhs.
void oracle_solve(Int n, Int nsuper, const Int *Super, const Int *Lpi, const Int *Lpx,
                  const Int *Ls, const double *Lx, const Int *Perm, const double *B,
                  double *X, Int nrhs) {
    double *Y = (double *)malloc(sizeof(double) * (n > 0 ? n : 1) * (nrhs > 0 ? nrhs : 1));
    for (Int r = 0; r < nrhs; r++)
        for (Int k = 0; k < n; k++) Y[k + r * n] = B[Perm[k] + r * n];
    oracle_super_lsolve(n, nsuper, Super, Lpi, Lpx, Ls, Lx, Y, nrhs);
    oracle_super_ltsolve(n, nsuper, Super, Lpi, Lpx, Ls, Lx, Y, nrhs);
    for (Int r = 0; r < nrhs; r++)
        for (Int k = 0; k < n; k++) X[Perm[k] + r * n] = Y[k + r * n];
    free(Y);
}

// S = tril(P A P') from tril(A) (both CSC, lower, int64): the permuted input
// cholmod_factorize hands to super_numeric (CHOLMOD/Cholesky/cholmod_factorize.c:220-250,
// ptranspose with values).  Rows in each column of S come out sorted.
// Sp has n+1 entries; Si/Sx have nnz = Ap[n] entries.  Returns 0.
int oracle_permute_lower(Int n, const Int *Ap, const Int *Ai, const double *Ax,
                         const Int *Perm, Int *Sp, Int *Si, double *Sx) {
    Int *Pinv = (Int *)malloc(sizeof(Int) * (n > 0 ? n : 1));
    Int *cnt = (Int *)calloc((size_t)(n + 1), sizeof(Int));
    for (Int k = 0; k < n; k++) Pinv[Perm[k]] = k;
    for (Int j = 0; j < n; j++)
        for (Int p = Ap[j]; p < Ap[j + 1]; p++) {
            Int a = Pinv[Ai[p]], b = Pinv[j];
            Int col = a < b ? a : b;
            cnt[col]++;
        }
    Sp[0] = 0;
    for (Int k = 0; k < n; k++) Sp[k + 1] = Sp[k] + cnt[k];
    for (Int k = 0; k < n; k++) cnt[k] = Sp[k];
    for (Int j = 0; j < n; j++)
        for (Int p = Ap[j]; p < Ap[j + 1]; p++) {
            Int a = Pinv[Ai[p]], b = Pinv[j];
            Int row = a > b ? a : b, col = a < b ? a : b;
            Int q = cnt[col]++;
            Si[q] = row;
            Sx[q] = Ax[p];
        }
    // insertion sort rows within each column (columns are short)
    for (Int k = 0; k < n; k++) {
        for (Int p = Sp[k] + 1; p < Sp[k + 1]; p++) {
            Int r = Si[p];
            double v = Sx[p];
            Int q = p - 1;
            while (q >= Sp[k] && Si[q] > r) { Si[q + 1] = Si[q]; Sx[q + 1] = Sx[q]; q--; }
            Si[q + 1] = r;
            Sx[q + 1] = v;
        }
    }
    free(Pinv);
    free(cnt);
    return 0;
}

// ---------------------------------------------------------------------------
// Rank-k update/downdate of a supernodal LL' factor in place:
//   Lnew Lnew' = L L' + C C' (update) or L L' - C C' (downdate)
// with C (n x ncol, CSC, sorted rows) in the factor's row order, as
// cholmod_updown takes it (CHOLMOD/Modify/cholmod_updown.c:26-43: L->Perm is
// not applied).
//   - C is taken 8 columns at a time (Common->maxrank,
//     Modify/t_cholmod_updown_worker.c:166-183); each column's path starts at
//     its first row, an empty column at n-1 (:196-209), and climbs the column
//     elimination tree of the factor's pattern: the next column of the
//     supernode, then the first row below it.  A supernodal factor keeps its
//     full pattern when the reference converts it (Utility/
//     t_cholmod_change_factor_3_template.c:41), so these are its parents too.
//   - W(:, r) = C(:, r), later entries of a column overwriting earlier ones
//     (Modify/t_cholmod_updown_wdim.c:118-137); Alpha[r] = 1.
//   - Each column j on the union of the paths, in increasing order, is turned
//     into its LDL' form (d = l_jj^2, l_ij / l_jj, as t_cholmod_change_factor_3
//     does), modified by the recurrence of Modify/t_cholmod_updown_numkr.c:
//     :294-324 (ALPHA_GAMMA) for the columns of W whose path holds j, then
//     :363-372 for each row below, and turned back (l_jj = sqrt(d), l_ij l_jj).
//     Ranks are applied in the order of C's columns; the reference orders
//     merged paths by a depth-first walk (t_cholmod_updown_worker.c:915-935),
//     which changes only the rounding.
//   - dbound > 0 bounds each new d as cholmod_dbound does
//     (Utility/t_cholmod_bound.c): counted in info[1].
//   - modfl = sum over the paths of (6 + 4 |L(:,j)|) per column
//     (t_cholmod_updown_worker.c:784).
// info[0] = first column whose new d is not positive (an LL' factor cannot
// hold it), else n.  Returns 0, or -1 if an entry of C lies outside the
// pattern of L(:, first row of its column) (the reference would add fill).
// ---------------------------------------------------------------------------
#define UD_RANK 8

int oracle_updown(Int n, Int nsuper, const Int *Super, const Int *Lpi, const Int *Lpx, const Int *Ls,
                  double *Lx, int update, Int ncol, const Int *Cp, const Int *Ci, const double *Cx,
                  double dbound, double *modfl, Int *info) {
    Int *colsuper = (Int *)malloc(sizeof(Int) * (n > 0 ? n : 1));
    double *W = (double *)malloc(sizeof(double) * UD_RANK * (n > 0 ? n : 1));
    unsigned char *act = (unsigned char *)malloc(n > 0 ? n : 1);
    if (!colsuper || !W || !act) {
        free(colsuper), free(W), free(act);
        return -2;
    }
    for (Int s = 0; s < nsuper; s++)
        for (Int j = Super[s]; j < Super[s + 1]; j++) colsuper[j] = s;
    info[0] = n;
    info[1] = 0;
    double fl = 0;
    int rc = 0;
    // pattern check: rows of C(:, r) inside {k} U pattern of L(:, k), k its first row
    for (Int r = 0; r < ncol && rc == 0; r++) {
        if (Cp[r + 1] == Cp[r]) continue;
        Int k = Ci[Cp[r]], s = colsuper[k], k2 = Super[s + 1];
        Int q = Lpi[s] + (k2 - Super[s]), qend = Lpi[s + 1];
        for (Int p = Cp[r]; p < Cp[r + 1]; p++) {
            Int i = Ci[p];
            if (i < k || i >= n) { rc = -1; break; }
            if (i < k2) continue;
            while (q < qend && Ls[q] < i) q++;
            if (q == qend || Ls[q] != i) { rc = -1; break; }
        }
    }
    for (Int r0 = 0; r0 < ncol && rc == 0; r0 += UD_RANK) {
        int k = (int)((ncol - r0) < UD_RANK ? (ncol - r0) : UD_RANK);
        memset(W, 0, sizeof(double) * UD_RANK * n);
        memset(act, 0, n);
        double alpha[UD_RANK];
        for (int r = 0; r < k; r++) {
            alpha[r] = 1.0;
            Int c = r0 + r;
            Int j = Cp[c + 1] > Cp[c] ? Ci[Cp[c]] : n - 1;
            const int numeric = Cp[c + 1] > Cp[c];
            for (Int p = Cp[c]; p < Cp[c + 1]; p++) W[Ci[p] * UD_RANK + r] = Cx[p];
            while (j >= 0) {
                Int s = colsuper[j], k1 = Super[s], nscol = Super[s + 1] - k1, nsrow = Lpi[s + 1] - Lpi[s];
                fl += 6.0 + 4.0 * (double)(nsrow - (j - k1));
                if (numeric) act[j] |= (unsigned char)(1u << r);
                j = (j + 1 < k1 + nscol) ? j + 1 : (nsrow > nscol ? Ls[Lpi[s] + nscol] : -1);
            }
        }
        for (Int j = 0; j < n; j++) {
            if (!act[j]) continue;
            Int s = colsuper[j], k1 = Super[s], nscol = Super[s + 1] - k1, nsrow = Lpi[s + 1] - Lpi[s];
            Int jl = j - k1, psi = Lpi[s];
            double *col = Lx + Lpx[s] + jl * nsrow;
            double lold = col[jl], dj = lold * lold;
            double wj[UD_RANK], gam[UD_RANK];
            for (int r = 0; r < k; r++) {
                if (!(act[j] & (1u << r))) continue;
                double w = W[j * UD_RANK + r], al = alpha[r];
                double t = w * w;
                t = t / dj;
                double a = update ? al + t : al - t;
                dj = dj * a;
                alpha[r] = a;
                gam[r] = update ? (-w) / dj : w / dj;
                dj = dj / al;
                wj[r] = w;
            }
            if (dbound > 0 && !isnan(dj)) {
                if (dj >= 0 ? dj < dbound : dj > -dbound) {
                    dj = dj >= 0 ? dbound : -dbound;
                    info[1]++;
                }
            }
            if (!(dj > 0) && info[0] == n) info[0] = j;
            double lnew = sqrt(dj);
            col[jl] = lnew;
            for (Int ii = jl + 1; ii < nsrow; ii++) {
                Int i = ii < nscol ? k1 + ii : Ls[psi + ii];
                double l = col[ii] / lold;
                for (int r = 0; r < k; r++) {
                    if (!(act[j] & (1u << r))) continue;
                    double *w = W + i * UD_RANK + r;
                    *w = *w - wj[r] * l;
                    l = l - gam[r] * *w;
                }
                col[ii] = l * lnew;
            }
        }
    }
    *modfl = fl;
    free(colsuper), free(W), free(act);
    return rc;
}
