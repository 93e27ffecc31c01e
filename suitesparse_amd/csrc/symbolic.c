// symbolic.c -- CPU symbolic analysis for the supernodal LL' factorization.
//
// Given tril(A) (CSC, int64) and a fill-reducing ordering, compute exactly the
// symbolic factor that CHOLMOD's cholmod_l_analyze_p2 builds for the
// supernodal case with default parameters (postorder on, nrelax = {4,16,48},
// zrelax = {0.8,0.1,0.05}, CPU-mode supernode sizes):
//
//   Perm, ColCount, Parent       CHOLMOD/Cholesky/cholmod_analyze.c:297-357 (analyze_ordering)
//                                 and :828-880 (weighted postorder folded into Perm)
//   etree                        CHOLMOD/Cholesky/cholmod_etree.c:81-221 (unique by definition)
//   postorder                    CHOLMOD/Cholesky/cholmod_postorder.c:137-286: children visited
//                                 in increasing (weight, index); roots in increasing index
//   column counts, fl, lnz       CHOLMOD/Cholesky/cholmod_rowcolcounts.c:182-530 (values are
//                                 unique; computed here with the first-descendant /
//                                 skeleton-leaf method of Gilbert, Ng and Peyton)
//   Super, Lpi, Lpx, Ls, sizes   CHOLMOD/Supernodal/cholmod_super_symbolic.c:391-928
//
// The supernode partition depends on tie-breaking details (the fundamental
// supernode test, the merge sweep order and the path-compressed Merged links,
// integer vs double arithmetic in the zero-count test), so those steps follow
// the reference's sequence exactly; everything else is written from the
// mathematics.

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/cholmod_mi355x.h"
#include "symbolic_internal.h"

typedef int64_t Int;
#define EMPTY (-1)

static void *xmalloc(size_t n, size_t sz) {
    if (n == 0) n = 1;
    return malloc(n * sz);
}

// ---------------------------------------------------------------------------
// Permuted patterns.  From tril(A) (rows i >= j in column j) and Pinv
// (Pinv[old] = new), build the upper pattern S of P A P' by columns (column k
// holds rows i <= k, sorted ascending) -- CHOLMOD's S = tril(A(p,p))' -- and
// its transpose F (column k holds rows i >= k, sorted).
// ---------------------------------------------------------------------------
static int permute_patterns(Int n, const Int *Ap, const Int *Ai, const Int *Pinv,
                            Int **Sp_out, Int **Si_out, Int **Fp_out, Int **Fi_out) {
    Int nnz = 0;
    for (Int j = 0; j < n; j++)
        for (Int p = Ap[j]; p < Ap[j + 1]; p++)
            if (Ai[p] >= j) nnz++;
    Int *Fp = (Int *)calloc((size_t)n + 1, sizeof(Int));
    Int *Fi = (Int *)xmalloc((size_t)nnz, sizeof(Int));
    Int *Sp = (Int *)calloc((size_t)n + 1, sizeof(Int));
    Int *Si = (Int *)xmalloc((size_t)nnz, sizeof(Int));
    Int *w = (Int *)xmalloc((size_t)n + 1, sizeof(Int));
    if (!Fp || !Fi || !Sp || !Si || !w) {
        free(Fp); free(Fi); free(Sp); free(Si); free(w);
        return MXC_OUT_OF_MEMORY;
    }
    // bucket the entries of F (lower) by their ROW = max(a,b): this makes a
    // row-major lower pattern, i.e. the column-major upper pattern S.
    for (Int j = 0; j < n; j++)
        for (Int p = Ap[j]; p < Ap[j + 1]; p++) {
            Int i = Ai[p];
            if (i < j) continue;
            Int a = Pinv[i], b = Pinv[j];
            Int hi = a > b ? a : b;
            Sp[hi + 1]++;
        }
    for (Int k = 0; k < n; k++) Sp[k + 1] += Sp[k];
    for (Int k = 0; k < n; k++) w[k] = Sp[k];
    // fill S column hi with rows lo, visiting lo in increasing order so that
    // each column comes out sorted: first bucket by lo into F, then sweep F.
    for (Int j = 0; j < n; j++)
        for (Int p = Ap[j]; p < Ap[j + 1]; p++) {
            Int i = Ai[p];
            if (i < j) continue;
            Int a = Pinv[i], b = Pinv[j];
            Int lo = a < b ? a : b;
            Fp[lo + 1]++;
        }
    for (Int k = 0; k < n; k++) Fp[k + 1] += Fp[k];
    Int *wf = (Int *)xmalloc((size_t)n + 1, sizeof(Int));
    if (!wf) { free(Fp); free(Fi); free(Sp); free(Si); free(w); return MXC_OUT_OF_MEMORY; }
    for (Int k = 0; k < n; k++) wf[k] = Fp[k];
    for (Int j = 0; j < n; j++)
        for (Int p = Ap[j]; p < Ap[j + 1]; p++) {
            Int i = Ai[p];
            if (i < j) continue;
            Int a = Pinv[i], b = Pinv[j];
            Int lo = a < b ? a : b, hi = a > b ? a : b;
            Fi[wf[lo]++] = hi;          // unsorted rows for now
        }
    // S = F' : sweep columns lo of F in increasing order -> sorted rows in S
    for (Int lo = 0; lo < n; lo++)
        for (Int p = Fp[lo]; p < Fp[lo + 1]; p++) Si[w[Fi[p]]++] = lo;
    // F = S' : sweep columns of S in increasing order -> sorted rows in F
    for (Int k = 0; k < n; k++) wf[k] = Fp[k];
    for (Int hi = 0; hi < n; hi++)
        for (Int p = Sp[hi]; p < Sp[hi + 1]; p++) Fi[wf[Si[p]]++] = hi;
    free(w);
    free(wf);
    *Sp_out = Sp; *Si_out = Si; *Fp_out = Fp; *Fi_out = Fi;
    return MXC_OK;
}

// Elimination tree of the upper pattern S (column j: rows i <= j): Liu's
// algorithm with path compression.  The tree is unique for a given pattern.
static void etree_upper(Int n, const Int *Sp, const Int *Si, Int *Parent, Int *Anc) {
    for (Int j = 0; j < n; j++) { Parent[j] = EMPTY; Anc[j] = EMPTY; }
    for (Int j = 0; j < n; j++) {
        for (Int p = Sp[j]; p < Sp[j + 1]; p++) {
            Int i = Si[p];
            if (i >= j) continue;
            // walk from i towards its root, compressing onto j
            while (i != EMPTY && i != j) {
                Int next = Anc[i];
                Anc[i] = j;
                if (next == EMPTY) Parent[i] = j;
                i = next;
            }
        }
    }
}

// Postorder of a forest.  Children of each node are visited in increasing
// (Weight, index) order (increasing index when Weight is NULL; weights are
// clamped to [0, n-1]); trees are taken in increasing order of their root.
// Returns the number of nodes ordered.
static Int postorder(Int n, const Int *Parent, const Int *Weight, Int *Post) {
    Int *first = (Int *)xmalloc((size_t)n, sizeof(Int));     // first child
    Int *sib = (Int *)xmalloc((size_t)n, sizeof(Int));       // next sibling
    Int *stack = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *key = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *order = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *cnt = (Int *)calloc((size_t)n + 1, sizeof(Int));
    Int k = 0;
    if (!first || !sib || !stack || !key || !order || !cnt) goto done;
    // order the nodes by (key, index) with a stable counting sort, then push
    // them onto their parents' child lists in reverse so lists come out ascending
    for (Int j = 0; j < n; j++) {
        Int w = 0;
        if (Weight) {
            w = Weight[j];
            if (w < 0) w = 0;
            if (w > n - 1) w = n - 1;
        }
        key[j] = w;
        cnt[w + 1]++;
    }
    for (Int w = 0; w < n; w++) cnt[w + 1] += cnt[w];
    for (Int j = 0; j < n; j++) order[cnt[key[j]]++] = j;
    for (Int j = 0; j < n; j++) { first[j] = EMPTY; sib[j] = EMPTY; }
    for (Int t = n - 1; t >= 0; t--) {
        Int j = order[t], p = Parent[j];
        if (p >= 0 && p < n) { sib[j] = first[p]; first[p] = j; }
    }
    for (Int r = 0; r < n; r++) {
        if (Parent[r] != EMPTY) continue;
        Int top = 0;
        stack[0] = r;
        while (top >= 0) {
            Int v = stack[top];
            Int c = first[v];
            if (c == EMPTY) {
                top--;
                Post[k++] = v;
            } else {
                first[v] = sib[c];
                stack[++top] = c;
            }
        }
    }
done:
    free(first); free(sib); free(stack); free(key); free(order); free(cnt);
    return k;
}

// Column counts of L (diagonal included) for the lower pattern F (column j:
// rows i >= j), given the etree and a postorder of it.  Skeleton-matrix /
// least-common-ancestor formulation (Gilbert, Ng, Peyton 1994).
static int column_counts(Int n, const Int *Fp, const Int *Fi, const Int *Parent,
                         const Int *Post, Int *ColCount) {
    Int *first = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *maxfirst = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *prevleaf = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *anc = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *delta = ColCount;
    if (!first || !maxfirst || !prevleaf || !anc) {
        free(first); free(maxfirst); free(prevleaf); free(anc);
        return MXC_OUT_OF_MEMORY;
    }
    for (Int j = 0; j < n; j++) { first[j] = EMPTY; maxfirst[j] = EMPTY; prevleaf[j] = EMPTY; anc[j] = j; }
    for (Int k = 0; k < n; k++) {
        Int j = Post[k];
        delta[j] = (first[j] == EMPTY) ? 1 : 0;       // leaf of the etree
        for (; j != EMPTY && first[j] == EMPTY; j = Parent[j]) first[j] = k;
    }
    for (Int k = 0; k < n; k++) {
        Int j = Post[k];
        if (Parent[j] != EMPTY) delta[Parent[j]]--;
        for (Int p = Fp[j]; p < Fp[j + 1]; p++) {
            Int i = Fi[p];
            if (i <= j || first[j] <= maxfirst[i]) continue;   // not a leaf of row subtree i
            maxfirst[i] = first[j];
            Int jprev = prevleaf[i];
            prevleaf[i] = j;
            delta[j]++;
            if (jprev != EMPTY) {
                // least common ancestor of jprev and j: root of jprev's set
                Int q = jprev;
                while (q != anc[q]) q = anc[q];
                for (Int s = jprev; s != q;) { Int nx = anc[s]; anc[s] = q; s = nx; }
                delta[q]--;
            }
        }
        if (Parent[j] != EMPTY) anc[j] = Parent[j];
    }
    for (Int k = 0; k < n; k++) {
        Int j = Post[k];
        if (Parent[j] != EMPTY) ColCount[Parent[j]] += ColCount[j];
    }
    free(first); free(maxfirst); free(prevleaf); free(anc);
    return MXC_OK;
}

// Relaxed supernodes and the supernodal pattern (cholmod_super_symbolic.c:391-928).
static int super_symbolic(Int n, const Int *Sp, const Int *Si, const Int *Parent,
                          const Int *ColCount, const Int nrelax[3], const double zrelax_in[3],
                          mxc_symbolic *L) {
    int rc = MXC_OUT_OF_MEMORY;
    Int *Nchild = (Int *)calloc((size_t)n + 1, sizeof(Int));
    Int *Super = (Int *)xmalloc((size_t)n + 1, sizeof(Int));
    Int *SuperMap = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Sparent = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Snz = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Merged = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Nscol = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Zeros = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Flag = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Lpi2 = (Int *)xmalloc((size_t)n + 1, sizeof(Int));
    double zrelax[3];
    if (!Nchild || !Super || !SuperMap || !Sparent || !Snz || !Merged || !Nscol || !Zeros ||
        !Flag || !Lpi2)
        goto done;
    for (int t = 0; t < 3; t++) zrelax[t] = isnan(zrelax_in[t]) ? 0 : zrelax_in[t];

    for (Int j = 0; j < n; j++)
        if (Parent[j] != EMPTY) Nchild[Parent[j]]++;

    // fundamental supernodes: j joins j-1 iff j is j-1's parent, its only child,
    // and column j-1's pattern is column j's plus the diagonal
    Int nfsuper = (n == 0) ? 0 : 1;
    Super[0] = 0;
    for (Int j = 1; j < n; j++) {
        if (Parent[j - 1] != j || ColCount[j - 1] != ColCount[j] + 1 || Nchild[j] > 1)
            Super[nfsuper++] = j;
    }
    Super[nfsuper] = n;
    for (Int s = 0; s < nfsuper; s++)
        for (Int k = Super[s]; k < Super[s + 1]; k++) SuperMap[k] = s;
    for (Int s = 0; s < nfsuper; s++) {
        Int par = Parent[Super[s + 1] - 1];
        Sparent[s] = (par == EMPTY) ? EMPTY : SuperMap[par];
    }
    for (Int s = 0; s < nfsuper; s++) {
        Merged[s] = EMPTY;
        Nscol[s] = Super[s + 1] - Super[s];
        Zeros[s] = 0;
        Snz[s] = ColCount[Super[s]];
    }
    // relaxed amalgamation: sweep s downward, trying to absorb s+1 when it is
    // (after earlier merges) s's parent
    for (Int s = nfsuper - 2; s >= 0; s--) {
        Int ss = Sparent[s];
        if (ss == EMPTY) continue;
        for (ss = Sparent[s]; Merged[ss] != EMPTY; ss = Merged[ss]) ;
        Int sparent = ss;
        for (ss = Sparent[s]; Merged[ss] != EMPTY;) {
            Int snext = Merged[ss];
            Merged[ss] = sparent;
            ss = snext;
        }
        if (sparent != s + 1) continue;
        Int nscol0 = Nscol[s], nscol1 = Nscol[s + 1];
        Int ns = nscol0 + nscol1;
        Int totzeros = Zeros[s + 1];
        double lnz1 = (double)Snz[s + 1];
        int merge;
        if (ns <= nrelax[0]) {
            merge = 1;
        } else {
            double lnz0 = (double)Snz[s];
            double xnewzeros = nscol0 * (lnz1 + nscol0 - lnz0);
            Int newzeros = nscol0 * (Snz[s + 1] + nscol0 - Snz[s]);
            if (xnewzeros == 0) {
                merge = 1;
            } else {
                double xtotzeros = ((double)totzeros) + xnewzeros;
                double xns = (double)ns;
                double xtotsize = (xns * (xns + 1) / 2) + xns * (lnz1 - nscol1);
                double z = xtotzeros / xtotsize;
                totzeros += newzeros;
                merge = ((ns <= nrelax[1] && z < zrelax[0]) || (ns <= nrelax[2] && z < zrelax[1]) ||
                         (z < zrelax[2])) &&
                        (xtotsize < ((double)INT64_MAX) / sizeof(double));
            }
        }
        if (merge) {
            Zeros[s] = totzeros;
            Merged[s + 1] = s;
            Snz[s] = nscol0 + Snz[s + 1];
            Nscol[s] += Nscol[s + 1];
        }
    }
    Int nsuper = 0;
    for (Int s = 0; s < nfsuper; s++) {
        if (Merged[s] == EMPTY) {
            Super[nsuper] = Super[s];
            Snz[nsuper] = Snz[s];
            nsuper++;
        }
    }
    Super[nsuper] = n;
    for (Int s = 0; s < nsuper; s++)
        for (Int k = Super[s]; k < Super[s + 1]; k++) SuperMap[k] = s;
    for (Int s = 0; s < nsuper; s++) {
        Int par = Parent[Super[s + 1] - 1];
        Sparent[s] = (par == EMPTY) ? EMPTY : SuperMap[par];
    }
    Int ssize = 0, xsize = 0;
    for (Int s = 0; s < nsuper; s++) {
        ssize += Snz[s];
        xsize += (Super[s + 1] - Super[s]) * Snz[s];
    }
    L->nsuper = nsuper;
    L->ssize = ssize > 1 ? ssize : 1;
    L->xsize = xsize > 1 ? xsize : 1;
    L->Super = (Int *)xmalloc((size_t)nsuper + 1, sizeof(Int));
    L->Lpi = (Int *)xmalloc((size_t)nsuper + 1, sizeof(Int));
    L->Lpx = (Int *)xmalloc((size_t)nsuper + 1, sizeof(Int));
    L->Ls = (Int *)calloc((size_t)L->ssize, sizeof(Int));
    if (!L->Super || !L->Lpi || !L->Lpx || !L->Ls) goto done;
    memcpy(L->Super, Super, sizeof(Int) * (size_t)(nsuper + 1));
    Int p = 0;
    for (Int s = 0; s < nsuper; s++) { L->Lpi[s] = p; p += Snz[s]; }
    L->Lpi[nsuper] = p;
    p = 0;
    for (Int s = 0; s < nsuper; s++) {
        L->Lpx[s] = p;
        p += (Super[s + 1] - Super[s]) * Snz[s];
    }
    L->Lpx[nsuper] = p;
    // row indices: each supernode's own columns first, then, for every later
    // column k, the supernodes on the path from SuperMap[i] (A(i,k) != 0, i < k1)
    // up to the first one already holding k
    for (Int s = 0; s < nsuper; s++) Lpi2[s] = L->Lpi[s];
    for (Int s = 0; s < nsuper; s++) Flag[s] = EMPTY;
    Int mark = 0;
    Int *Ls = L->Ls;
    for (Int s = 0; s < nsuper; s++) {
        Int k1 = Super[s], k2 = Super[s + 1];
        for (Int k = k1; k < k2; k++) Ls[Lpi2[s]++] = k;
        for (Int k = k1; k < k2; k++) {
            mark++;
            Flag[s] = mark;
            for (Int q = Sp[k]; q < Sp[k + 1]; q++) {
                Int i = Si[q];
                if (i >= k1) break;      // S columns are sorted
                for (Int si = SuperMap[i]; Flag[si] < mark; si = Sparent[si]) {
                    Ls[Lpi2[si]++] = k;
                    Flag[si] = mark;
                }
            }
        }
    }
    // largest update block and largest off-diagonal row count
    Int maxcsize = 1, maxesize = 1;
    for (Int d = 0; d < nsuper; d++) {
        Int nscol = Super[d + 1] - Super[d];
        Int q = L->Lpi[d] + nscol, plast = q, pend = L->Lpi[d + 1];
        Int esize = pend - q;
        if (esize > maxesize) maxesize = esize;
        Int slast = (q == pend) ? EMPTY : SuperMap[Ls[q]];
        for (; q <= pend; q++) {
            Int s = (q == pend) ? EMPTY : SuperMap[Ls[q]];
            if (s != slast) {
                Int ndrow1 = q - plast, ndrow2 = pend - plast;
                Int csize = ndrow2 * ndrow1;
                if (csize > maxcsize) maxcsize = csize;
                plast = q;
                slast = s;
            }
        }
    }
    L->maxcsize = maxcsize;
    L->maxesize = maxesize;
    rc = MXC_OK;
done:
    free(Nchild); free(Super); free(SuperMap); free(Sparent); free(Snz); free(Merged);
    free(Nscol); free(Zeros); free(Flag); free(Lpi2);
    return rc;
}

void mxc_symbolic_free(mxc_symbolic *L) {
    if (!L) return;
    free(L->Perm); free(L->ColCount); free(L->Parent);
    free(L->Super); free(L->Lpi); free(L->Lpx); free(L->Ls);
    free(L);
}

// Analyze tril(A) under the given ordering.  ordering: MXC_ORDER_GIVEN (UserPerm),
// MXC_ORDER_NATURAL or MXC_ORDER_AMD.
int mxc_analyze(int64_t n, const int64_t *Ap, const int64_t *Ai, int ordering,
                const int64_t *UserPerm, mxc_symbolic **out) {
    if (!out || n < 0 || (n > 0 && (!Ap || !Ai))) return MXC_INVALID;
    *out = NULL;
    mxc_symbolic *L = (mxc_symbolic *)calloc(1, sizeof(mxc_symbolic));
    if (!L) return MXC_OUT_OF_MEMORY;
    L->n = n;
    L->ordering = ordering;
    Int *Perm = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Pinv = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Parent = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Post = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *ColCount = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Wk = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Sp = NULL, *Si = NULL, *Fp = NULL, *Fi = NULL;
    int rc = MXC_OUT_OF_MEMORY;
    if (!Perm || !Pinv || !Parent || !Post || !ColCount || !Wk) goto fail;

    if (ordering == MXC_ORDER_GIVEN) {
        if (!UserPerm) { rc = MXC_INVALID; goto fail; }
        for (Int k = 0; k < n; k++) Pinv[k] = EMPTY;
        for (Int k = 0; k < n; k++) {
            Int j = UserPerm[k];
            if (j < 0 || j >= n || Pinv[j] != EMPTY) { rc = MXC_INVALID; goto fail; }
            Pinv[j] = k;
            Perm[k] = j;
        }
    } else if (ordering == MXC_ORDER_NATURAL) {
        for (Int k = 0; k < n; k++) Perm[k] = k;
    } else if (ordering == MXC_ORDER_AMD) {
        rc = mxc_amd_order(n, Ap, Ai, Perm);
        if (rc != MXC_OK) goto fail;
    } else {
        rc = MXC_INVALID;
        goto fail;
    }
    for (Int k = 0; k < n; k++) Pinv[Perm[k]] = k;

    // analyze the ordering: etree, postorder, column counts
    rc = permute_patterns(n, Ap, Ai, Pinv, &Sp, &Si, &Fp, &Fi);
    if (rc != MXC_OK) goto fail;
    etree_upper(n, Sp, Si, Parent, Wk);
    if (postorder(n, Parent, NULL, Post) != n) { rc = MXC_INVALID; goto fail; }
    rc = column_counts(n, Fp, Fi, Parent, Post, ColCount);
    if (rc != MXC_OK) goto fail;
    double fl = 0, lnz = 0;
    for (Int j = 0; j < n; j++) {
        double f = (double)ColCount[j];
        lnz += f;
        fl += f * f;
    }
    L->fl = fl;
    L->lnz = lnz;

    // weighted postorder (children by increasing column count), folded into Perm
    if (postorder(n, Parent, ColCount, Post) == n) {
        for (Int k = 0; k < n; k++) Wk[k] = Perm[Post[k]];
        memcpy(Perm, Wk, sizeof(Int) * (size_t)n);
        for (Int k = 0; k < n; k++) Wk[k] = ColCount[Post[k]];
        memcpy(ColCount, Wk, sizeof(Int) * (size_t)n);
        for (Int k = 0; k < n; k++) Pinv[Post[k]] = k;          // inverse postorder
        for (Int k = 0; k < n; k++) {
            Int op = Parent[Post[k]];
            Wk[k] = (op == EMPTY) ? EMPTY : Pinv[op];
        }
        memcpy(Parent, Wk, sizeof(Int) * (size_t)n);
    }
    for (Int k = 0; k < n; k++) Pinv[Perm[k]] = k;
    free(Sp); free(Si); free(Fp); free(Fi);
    Sp = Si = Fp = Fi = NULL;

    // supernodal pattern of the final ordering
    rc = permute_patterns(n, Ap, Ai, Pinv, &Sp, &Si, &Fp, &Fi);
    if (rc != MXC_OK) goto fail;
    {
        const Int nrelax[3] = {4, 16, 48};
        const double zrelax[3] = {0.8, 0.1, 0.05};
        rc = super_symbolic(n, Sp, Si, Parent, ColCount, nrelax, zrelax, L);
        if (rc != MXC_OK) goto fail;
    }
    L->Perm = Perm;
    L->ColCount = ColCount;
    L->Parent = Parent;
    free(Pinv); free(Post); free(Wk);
    free(Sp); free(Si); free(Fp); free(Fi);
    *out = L;
    return MXC_OK;
fail:
    free(Perm); free(Pinv); free(Parent); free(Post); free(ColCount); free(Wk);
    free(Sp); free(Si); free(Fp); free(Fi);
    L->Perm = NULL; L->ColCount = NULL; L->Parent = NULL;
    mxc_symbolic_free(L);
    return rc;
}

int mxc_symbolic_info(const mxc_symbolic *L, int64_t *info, double *dinfo) {
    if (!L || !info) return MXC_INVALID;
    info[0] = L->n;
    info[1] = L->nsuper;
    info[2] = L->ssize;
    info[3] = L->xsize;
    info[4] = L->maxcsize;
    info[5] = L->maxesize;
    info[6] = L->ordering;
    if (dinfo) { dinfo[0] = L->fl; dinfo[1] = L->lnz; }
    return MXC_OK;
}

int mxc_symbolic_export(const mxc_symbolic *L, int64_t *Perm, int64_t *ColCount, int64_t *Super,
                        int64_t *Lpi, int64_t *Lpx, int64_t *Ls) {
    if (!L) return MXC_INVALID;
    size_t n = (size_t)L->n, ns = (size_t)L->nsuper;
    if (Perm && L->Perm) memcpy(Perm, L->Perm, sizeof(Int) * n);
    if (ColCount && L->ColCount) memcpy(ColCount, L->ColCount, sizeof(Int) * n);
    if (Super) memcpy(Super, L->Super, sizeof(Int) * (ns + 1));
    if (Lpi) memcpy(Lpi, L->Lpi, sizeof(Int) * (ns + 1));
    if (Lpx) memcpy(Lpx, L->Lpx, sizeof(Int) * (ns + 1));
    if (Ls) memcpy(Ls, L->Ls, sizeof(Int) * (size_t)L->ssize);
    return MXC_OK;
}

// The columns of L reached from the columns C[0..blen) (the pattern of L \ b
// for b nonzero at C), in the order cholmod_l_lsolve_pattern returns them
// (cholmod_rowfac.c:383-398 -> row_lsubtree :127-153, :563-568): L taken as
// the reference's simplicial conversion of the supernodal factor (explicit
// zeros kept), whose parent of column j is its first off-diagonal row
// (cholmod_rowfac.c:539), i.e. j + 1 inside j's supernode, else the
// supernode's first row below its columns, else none.  Each new path from a
// C entry up to the first column already reached is placed, bottom-up, in
// front of the earlier paths.  Yset needs room for n entries.
int mxc_reach(int64_t n, int64_t nsuper, const int64_t *Super, const int64_t *Lpi, const int64_t *Ls,
              const int64_t *C, int64_t blen, int64_t *Yset, int64_t *ylen) {
    int64_t *col_super = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t *path = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    char *seen = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    if (!col_super || !path || !seen) {
        free(col_super); free(path); free(seen);
        return MXC_OUT_OF_MEMORY;
    }
    for (int64_t s = 0; s < nsuper; s++)
        for (int64_t k = Super[s]; k < Super[s + 1]; k++) col_super[k] = s;
    int64_t top = n;
    for (int64_t p = 0; p < blen; p++) {
        int64_t len = 0;
        for (int64_t i = C[p]; i != -1 && !seen[i];) {
            path[len++] = i;
            seen[i] = 1;
            const int64_t s = col_super[i];
            if (i + 1 < Super[s + 1]) {
                i = i + 1;
            } else {
                const int64_t nscol = Super[s + 1] - Super[s], nsrow = Lpi[s + 1] - Lpi[s];
                i = nsrow > nscol ? Ls[Lpi[s] + nscol] : -1;
            }
        }
        while (len > 0) Yset[--top] = path[--len];
    }
    *ylen = n - top;
    memmove(Yset, Yset + top, sizeof(int64_t) * (size_t)(n - top));
    free(col_super); free(path); free(seen);
    return MXC_OK;
}

int mxc_lsolve_pattern(const mxc_symbolic *S, int sys, const int64_t *Bset, int64_t blen, int64_t *Xset,
                       int64_t *xlen) {
    if (!S || !Xset || !xlen || blen < 0 || (blen > 0 && !Bset)) return MXC_INVALID;
    if (sys < MXC_SYS_A || sys > MXC_SYS_D) return MXC_INVALID;
    const int64_t n = S->n;
    for (int64_t p = 0; p < blen; p++)
        if (Bset[p] < 0 || Bset[p] >= n) return MXC_INVALID;
    const int perm = sys == MXC_SYS_A && S->Perm != NULL;
    int64_t *C = (int64_t *)malloc(sizeof(int64_t) * (size_t)(blen > 0 ? blen : 1));
    int64_t *Pinv = perm ? (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1)) : NULL;
    if (!C || (perm && !Pinv)) {
        free(C); free(Pinv);
        return MXC_OUT_OF_MEMORY;
    }
    if (perm)
        for (int64_t k = 0; k < n; k++) Pinv[S->Perm[k]] = k;
    for (int64_t p = 0; p < blen; p++) C[p] = perm ? Pinv[Bset[p]] : Bset[p];
    int rc = MXC_OK;
    if (sys == MXC_SYS_D) {
        for (int64_t p = 0; p < blen; p++) Xset[p] = C[p];
        *xlen = blen;
    } else {
        rc = mxc_reach(n, S->nsuper, S->Super, S->Lpi, S->Ls, C, blen, Xset, xlen);
    }
    if (rc == MXC_OK && perm)
        for (int64_t p = 0; p < *xlen; p++) Xset[p] = S->Perm[Xset[p]];
    free(C); free(Pinv);
    return rc;
}

// Wrap a supernodal symbolic factor built elsewhere (e.g. the fields of a
// cholmod_factor after cholmod_l_analyze: L->Perm, L->super, L->pi, L->px, L->s).
int mxc_symbolic_import(int64_t n, int64_t nsuper, const int64_t *Perm, const int64_t *Super,
                        const int64_t *Lpi, const int64_t *Lpx, const int64_t *Ls,
                        mxc_symbolic **out) {
    if (!out || n < 0 || nsuper < 0 || !Perm || !Super || !Lpi || !Lpx || !Ls) return MXC_INVALID;
    *out = NULL;
    if (Super[0] != 0 || Super[nsuper] != n || Lpi[0] != 0 || Lpx[0] != 0) return MXC_INVALID;
    mxc_symbolic *L = (mxc_symbolic *)calloc(1, sizeof(mxc_symbolic));
    if (!L) return MXC_OUT_OF_MEMORY;
    L->n = n;
    L->nsuper = nsuper;
    L->ordering = MXC_ORDER_GIVEN;
    L->ssize = Lpi[nsuper] > 1 ? Lpi[nsuper] : 1;
    L->xsize = Lpx[nsuper] > 1 ? Lpx[nsuper] : 1;
    L->Perm = (Int *)xmalloc((size_t)n, sizeof(Int));
    L->Super = (Int *)xmalloc((size_t)nsuper + 1, sizeof(Int));
    L->Lpi = (Int *)xmalloc((size_t)nsuper + 1, sizeof(Int));
    L->Lpx = (Int *)xmalloc((size_t)nsuper + 1, sizeof(Int));
    L->Ls = (Int *)xmalloc((size_t)L->ssize, sizeof(Int));
    if (!L->Perm || !L->Super || !L->Lpi || !L->Lpx || !L->Ls) { mxc_symbolic_free(L); return MXC_OUT_OF_MEMORY; }
    memcpy(L->Perm, Perm, sizeof(Int) * (size_t)n);
    memcpy(L->Super, Super, sizeof(Int) * (size_t)(nsuper + 1));
    memcpy(L->Lpi, Lpi, sizeof(Int) * (size_t)(nsuper + 1));
    memcpy(L->Lpx, Lpx, sizeof(Int) * (size_t)(nsuper + 1));
    memcpy(L->Ls, Ls, sizeof(Int) * (size_t)Lpi[nsuper]);
    // validate: supernode sizes, own columns leading each row list, sorted rows
    Int maxe = 1;
    for (Int s = 0; s < nsuper; s++) {
        Int nscol = Super[s + 1] - Super[s], nsrow = Lpi[s + 1] - Lpi[s];
        if (nscol <= 0 || nsrow < nscol || Lpx[s + 1] - Lpx[s] != nscol * nsrow) {
            mxc_symbolic_free(L);
            return MXC_INVALID;
        }
        for (Int k = 0; k < nsrow; k++) {
            Int r = Ls[Lpi[s] + k];
            if (r < 0 || r >= n || (k < nscol && r != Super[s] + k) ||
                (k > 0 && r <= Ls[Lpi[s] + k - 1])) {
                mxc_symbolic_free(L);
                return MXC_INVALID;
            }
        }
        if (nsrow - nscol > maxe) maxe = nsrow - nscol;
    }
    L->maxesize = maxe;
    L->maxcsize = 1;
    double lnz = 0;
    for (Int s = 0; s < nsuper; s++) {
        Int nscol = Super[s + 1] - Super[s], nsrow = Lpi[s + 1] - Lpi[s];
        for (Int k = 0; k < nscol; k++) lnz += (double)(nsrow - k);
    }
    L->lnz = lnz;
    L->fl = 0;
    *out = L;
    return MXC_OK;
}
