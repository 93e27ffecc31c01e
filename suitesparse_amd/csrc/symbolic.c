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
// The supernode partition depends on the exact rules CHOLMOD applies (the
// fundamental-supernode test, the order of the relaxed-merge sweep and the
// integer / double arithmetic of its zero-count test), so those two rules are
// stated as CHOLMOD states them; the row structure, the sizes and everything
// else are written from their definitions.

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

// ---------------------------------------------------------------------------
// Supernodes.  Three steps, each in its own function:
//   1. amalgamate(): the column partition.  Two rules decide it and are kept
//      as CHOLMOD states them, because the partition (hence Super, Lpx and
//      everything the numeric phase does) must be identical:
//        - a column joins the run of the column before it when it is that
//          column's etree parent, that column's only child, and the column
//          before has exactly one more entry
//          (cholmod_super_symbolic.c:391-438);
//        - relaxed merging visits the runs from the second-to-last down to
//          the first and lets a run absorb the run right after it when that
//          run is (after earlier absorptions) its parent and the zero count
//          of the union passes the nrelax/zrelax test, evaluated with the
//          same mixed integer / double arithmetic (:470-600).
//   2. supernode_rows(): the row structure of every supernode, sorted.  It is
//      defined by the partition alone: the supernode's own columns, then every
//      row k beyond them with L(k, j) != 0 for some column j of the supernode
//      or of a supernode below it in the supernodal etree.  Built bottom-up as
//      a union (A's entries below the supernode plus its children's rows).
//   3. update_extents(): maxesize (most rows below a supernode's columns) and
//      maxcsize (largest ndrow1 x ndrow2 update block, ndrow1 the rows of a
//      descendant that fall in one ancestor, ndrow2 those rows and all below).
// ---------------------------------------------------------------------------

// One run of columns during amalgamation.
typedef struct {
    Int width;    // columns in the run (grows as parents are absorbed)
    Int rows;     // entries of the run's first column of L, diagonal included
    Int zeros;    // explicit zeros that absorbing has introduced
    Int up;       // the run holding the etree parent of the run's last column
    Int absorber; // EMPTY while the run is alive, else the run that absorbed it
} run_t;

static Int run_alive(run_t *r, Int t) {
    Int top = t;
    while (r[top].absorber != EMPTY) top = r[top].absorber;
    while (r[t].absorber != EMPTY) {           // shorten the chain for later finds
        Int nx = r[t].absorber;
        r[t].absorber = top;
        t = nx;
    }
    return top;
}

// May run c absorb the run right after it (p)?  *zeros_out receives the zero
// count of the union.
static int relaxed_ok(const run_t *c, const run_t *p, const Int nrelax[3], const double zrelax[3],
                      Int *zeros_out) {
    const Int w = c->width + p->width;
    *zeros_out = p->zeros;
    if (w <= nrelax[0]) return 1;
    // c's columns fill up to p's first-column height: the new zeros, counted
    // both in integers (kept) and in doubles (tested)
    const Int add = c->width * (p->rows + c->width - c->rows);
    const double add_d = c->width * ((double)p->rows + c->width - (double)c->rows);
    if (add_d == 0) return 1;
    *zeros_out = p->zeros + add;
    const double wd = (double)w;
    const double size_d = wd * (wd + 1) / 2 + wd * ((double)p->rows - p->width);
    const double frac = ((double)p->zeros + add_d) / size_d;
    const int small = (w <= nrelax[1] && frac < zrelax[0]) || (w <= nrelax[2] && frac < zrelax[1]) ||
                      frac < zrelax[2];
    return small && size_d < ((double)INT64_MAX) / sizeof(double);
}

// Column partition: on return Super[0..*nsuper] and Srows[0..*nsuper) (rows of
// each supernode, own columns included).
static int amalgamate(Int n, const Int *Parent, const Int *ColCount, const Int nrelax[3],
                      const double zrelax_in[3], Int *Super, Int *Srows, Int *nsuper_out) {
    double zrelax[3];
    for (int t = 0; t < 3; t++) zrelax[t] = isnan(zrelax_in[t]) ? 0 : zrelax_in[t];
    Int *kids = (Int *)calloc((size_t)n + 1, sizeof(Int));
    Int *run_of = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *first = (Int *)xmalloc((size_t)n + 1, sizeof(Int));
    run_t *run = (run_t *)xmalloc((size_t)n, sizeof(run_t));
    if (!kids || !run_of || !first || !run) {
        free(kids); free(run_of); free(first); free(run);
        return MXC_OUT_OF_MEMORY;
    }
    for (Int j = 0; j < n; j++)
        if (Parent[j] != EMPTY) kids[Parent[j]]++;
    // fundamental runs
    Int nrun = 0;
    for (Int j = 0; j < n; j++) {
        const int chain = j > 0 && Parent[j - 1] == j && ColCount[j - 1] == ColCount[j] + 1 && kids[j] <= 1;
        if (!chain) {
            first[nrun] = j;
            run[nrun].width = 0;
            run[nrun].rows = ColCount[j];
            run[nrun].zeros = 0;
            run[nrun].absorber = EMPTY;
            nrun++;
        }
        run[nrun - 1].width++;
        run_of[j] = nrun - 1;
    }
    first[nrun] = n;
    for (Int t = 0; t < nrun; t++) {
        const Int par = Parent[first[t + 1] - 1];
        run[t].up = par == EMPTY ? EMPTY : run_of[par];
    }
    // relaxed merging, last-but-one run first
    for (Int t = nrun - 2; t >= 0; t--) {
        if (run[t].up == EMPTY || run_alive(run, run[t].up) != t + 1) continue;
        Int z;
        if (!relaxed_ok(&run[t], &run[t + 1], nrelax, zrelax, &z)) continue;
        run[t].zeros = z;
        run[t].rows = run[t].width + run[t + 1].rows;
        run[t].width += run[t + 1].width;
        run[t + 1].absorber = t;
    }
    Int ns = 0;
    for (Int t = 0; t < nrun; t++) {
        if (run[t].absorber != EMPTY) continue;
        Super[ns] = first[t];
        Srows[ns] = run[t].rows;
        ns++;
    }
    Super[ns] = n;
    *nsuper_out = ns;
    free(kids); free(run_of); free(first); free(run);
    return MXC_OK;
}

static int cmp_int64(const void *a, const void *b) {
    const Int x = *(const Int *)a, y = *(const Int *)b;
    return (x > y) - (x < y);
}

// Row lists of the supernodes (see the section comment).  Fp/Fi: the lower
// pattern of the permuted matrix (column j: rows i >= j).  Lpi is set from
// Srows; the union must reproduce those counts.
static int supernode_rows(Int n, Int nsuper, const Int *Super, const Int *Srows, const Int *Fp,
                          const Int *Fi, const Int *Parent, const Int *Lpi, Int *Ls) {
    Int *owner = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *seen = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *head = (Int *)xmalloc((size_t)nsuper, sizeof(Int));
    Int *next = (Int *)xmalloc((size_t)nsuper, sizeof(Int));
    int rc = MXC_OUT_OF_MEMORY;
    if (!owner || !seen || !head || !next) goto out;
    for (Int s = 0; s < nsuper; s++) {
        head[s] = EMPTY;
        for (Int k = Super[s]; k < Super[s + 1]; k++) owner[k] = s;
    }
    for (Int k = 0; k < n; k++) seen[k] = EMPTY;
    // child lists (a child's columns precede its parent's in a postorder)
    for (Int s = nsuper - 1; s >= 0; s--) {
        const Int par = Parent[Super[s + 1] - 1];
        if (par == EMPTY) continue;
        next[s] = head[owner[par]];
        head[owner[par]] = s;
    }
    rc = MXC_INVALID;
    for (Int s = 0; s < nsuper; s++) {
        const Int c0 = Super[s], c1 = Super[s + 1];
        Int *out = Ls + Lpi[s];
        const Int cap = Srows[s];
        Int cnt = 0;
        for (Int k = c0; k < c1; k++) {
            out[cnt++] = k;
            seen[k] = s;
        }
        for (Int j = c0; j < c1; j++)
            for (Int p = Fp[j]; p < Fp[j + 1]; p++) {
                const Int i = Fi[p];
                if (i < c1 || seen[i] == s) continue;
                if (cnt == cap) goto out;
                seen[i] = s;
                out[cnt++] = i;
            }
        for (Int c = head[s]; c != EMPTY; c = next[c]) {
            const Int own = Super[c + 1] - Super[c];
            for (Int q = Lpi[c] + own; q < Lpi[c + 1]; q++) {
                const Int i = Ls[q];
                if (i < c1 || seen[i] == s) continue;
                if (cnt == cap) goto out;
                seen[i] = s;
                out[cnt++] = i;
            }
        }
        if (cnt != cap) goto out;
        qsort(out + (c1 - c0), (size_t)(cnt - (c1 - c0)), sizeof(Int), cmp_int64);
    }
    rc = MXC_OK;
out:
    free(owner); free(seen); free(head); free(next);
    return rc;
}

static int update_extents(Int n, Int nsuper, const Int *Super, const Int *Lpi, const Int *Ls,
                          Int *maxcsize, Int *maxesize) {
    Int *owner = (Int *)xmalloc((size_t)n, sizeof(Int));
    if (!owner) return MXC_OUT_OF_MEMORY;
    for (Int s = 0; s < nsuper; s++)
        for (Int k = Super[s]; k < Super[s + 1]; k++) owner[k] = s;
    Int cmax = 1, emax = 1;
    for (Int d = 0; d < nsuper; d++) {
        const Int end = Lpi[d + 1];
        Int q = Lpi[d] + (Super[d + 1] - Super[d]);
        if (end - q > emax) emax = end - q;
        while (q < end) {                       // one ancestor's rows per pass
            const Int anc = owner[Ls[q]];
            Int r = q + 1;
            while (r < end && owner[Ls[r]] == anc) r++;
            const Int c = (r - q) * (end - q);
            if (c > cmax) cmax = c;
            q = r;
        }
    }
    *maxcsize = cmax;
    *maxesize = emax;
    free(owner);
    return MXC_OK;
}

// Supernodal symbolic factor of the permuted matrix (lower pattern Fp/Fi),
// given its etree and column counts: the fields CHOLMOD's super_symbolic2
// produces (cholmod_super_symbolic.c:147-938, CPU mode).
static int super_symbolic(Int n, const Int *Fp, const Int *Fi, const Int *Parent, const Int *ColCount,
                          const Int nrelax[3], const double zrelax[3], mxc_symbolic *L) {
    Int *Super = (Int *)xmalloc((size_t)n + 1, sizeof(Int));
    Int *Srows = (Int *)xmalloc((size_t)n + 1, sizeof(Int));
    Int nsuper = 0;
    int rc = MXC_OUT_OF_MEMORY;
    if (!Super || !Srows) goto done;
    rc = amalgamate(n, Parent, ColCount, nrelax, zrelax, Super, Srows, &nsuper);
    if (rc != MXC_OK) goto done;
    rc = MXC_OUT_OF_MEMORY;
    L->nsuper = nsuper;
    L->Super = (Int *)xmalloc((size_t)nsuper + 1, sizeof(Int));
    L->Lpi = (Int *)xmalloc((size_t)nsuper + 1, sizeof(Int));
    L->Lpx = (Int *)xmalloc((size_t)nsuper + 1, sizeof(Int));
    if (!L->Super || !L->Lpi || !L->Lpx) goto done;
    memcpy(L->Super, Super, sizeof(Int) * (size_t)(nsuper + 1));
    L->Lpi[0] = 0;
    L->Lpx[0] = 0;
    for (Int s = 0; s < nsuper; s++) {
        L->Lpi[s + 1] = L->Lpi[s] + Srows[s];
        L->Lpx[s + 1] = L->Lpx[s] + (Super[s + 1] - Super[s]) * Srows[s];
    }
    L->ssize = L->Lpi[nsuper] > 1 ? L->Lpi[nsuper] : 1;
    L->xsize = L->Lpx[nsuper] > 1 ? L->Lpx[nsuper] : 1;
    L->Ls = (Int *)calloc((size_t)L->ssize, sizeof(Int));
    if (!L->Ls) goto done;
    rc = supernode_rows(n, nsuper, Super, Srows, Fp, Fi, Parent, L->Lpi, L->Ls);
    if (rc != MXC_OK) goto done;
    rc = update_extents(n, nsuper, Super, L->Lpi, L->Ls, &L->maxcsize, &L->maxesize);
done:
    free(Super); free(Srows);
    return rc;
}

void mxc_symbolic_free(mxc_symbolic *L) {
    if (!L) return;
    free(L->Perm); free(L->ColCount); free(L->Parent);
    free(L->Super); free(L->Lpi); free(L->Lpx); free(L->Ls);
    free(L);
}

// lnz of the factor under Perm (Common->lnz after cholmod_l_analyze_ordering)
static int ordering_lnz(Int n, const Int *Ap, const Int *Ai, const Int *Perm, double *lnz) {
    Int *Pinv = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Parent = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Post = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *ColCount = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Wk = (Int *)xmalloc((size_t)n, sizeof(Int));
    Int *Sp = NULL, *Si = NULL, *Fp = NULL, *Fi = NULL;
    int rc = MXC_OUT_OF_MEMORY;
    if (!Pinv || !Parent || !Post || !ColCount || !Wk) goto done;
    for (Int k = 0; k < n; k++) Pinv[Perm[k]] = k;
    rc = permute_patterns(n, Ap, Ai, Pinv, &Sp, &Si, &Fp, &Fi);
    if (rc != MXC_OK) goto done;
    etree_upper(n, Sp, Si, Parent, Wk);
    if (postorder(n, Parent, NULL, Post) != n) { rc = MXC_INVALID; goto done; }
    rc = column_counts(n, Fp, Fi, Parent, Post, ColCount);
    if (rc != MXC_OK) goto done;
    double s = 0;
    for (Int j = 0; j < n; j++) s += (double)ColCount[j];
    *lnz = s;
done:
    free(Pinv); free(Parent); free(Post); free(ColCount); free(Wk);
    free(Sp); free(Si); free(Fp); free(Fi);
    return rc;
}

// Perm from UserPerm, rejecting anything that is not a permutation of 0..n-1
static int given_perm(Int n, const Int *UserPerm, Int *Perm, Int *Pinv) {
    if (!UserPerm) return MXC_INVALID;
    for (Int k = 0; k < n; k++) Pinv[k] = EMPTY;
    for (Int k = 0; k < n; k++) {
        Int j = UserPerm[k];
        if (j < 0 || j >= n || Pinv[j] != EMPTY) return MXC_INVALID;
        Pinv[j] = k;
        Perm[k] = j;
    }
    return MXC_OK;
}

// CHOLMOD's default strategy (CHOLMOD/Cholesky/cholmod_analyze.c:444-462,
// :554-782): GIVEN when UserPerm is passed (lnz of its analysis), AMD (lnz
// as AMD predicts it), then METIS (lnz of its analysis) unless AMD predicted
// fewer than 500 flops per entry of L or nnz(L) < 5 nnz(A); the first
// strictly smaller lnz wins.  Leaves the winner in Perm and its ordering in
// L->ordering.
static int default_strategy(Int n, const Int *Ap, const Int *Ai, const Int *UserPerm, Int *Perm,
                            Int *Wk, mxc_symbolic *L) {
    int rc;
    double best = -1;
    L->selected = -1;
    Int *cand = (Int *)xmalloc((size_t)n, sizeof(Int));
    if (!cand) return MXC_OUT_OF_MEMORY;
    if (UserPerm) {
        rc = given_perm(n, UserPerm, cand, Wk);
        if (rc != MXC_OK) goto done;
        double lnz;
        rc = ordering_lnz(n, Ap, Ai, cand, &lnz);
        if (rc != MXC_OK) goto done;
        L->method_lnz[0] = lnz;
        best = lnz;
        L->selected = 0;
        L->ordering = MXC_ORDER_GIVEN;
        memcpy(Perm, cand, sizeof(Int) * (size_t)n);
    }
    double st[3];
    rc = mxc_amd_order_stats(n, Ap, Ai, cand, st);
    if (rc != MXC_OK) goto done;
    L->method_fl[1] = st[0];
    L->method_lnz[1] = st[1];
    L->anz = st[2];
    if (L->selected == -1 || st[1] < best) {
        best = st[1];
        L->selected = 1;
        L->ordering = MXC_ORDER_AMD;
        memcpy(Perm, cand, sizeof(Int) * (size_t)n);
    }
    if (!(st[0] < 500 * st[1] || st[1] < 5 * st[2])) {
        rc = mxc_metis_order(n, Ap, Ai, 3000, 0.66, cand);
        if (rc != MXC_OK) goto done;
        double lnz;
        rc = ordering_lnz(n, Ap, Ai, cand, &lnz);
        if (rc != MXC_OK) goto done;
        L->method_lnz[2] = lnz;
        if (lnz < best) {
            L->selected = 2;
            L->ordering = MXC_ORDER_METIS;
            memcpy(Perm, cand, sizeof(Int) * (size_t)n);
        }
    }
    rc = MXC_OK;
done:
    free(cand);
    return rc;
}

// Analyze tril(A) under an ordering: MXC_ORDER_GIVEN (UserPerm),
// MXC_ORDER_NATURAL, MXC_ORDER_AMD, MXC_ORDER_METIS or MXC_ORDER_DEFAULT.
int mxc_analyze(int64_t n, const int64_t *Ap, const int64_t *Ai, int ordering,
                const int64_t *UserPerm, mxc_symbolic **out) {
    if (!out || n < 0 || (n > 0 && (!Ap || !Ai))) return MXC_INVALID;
    *out = NULL;
    mxc_symbolic *L = (mxc_symbolic *)calloc(1, sizeof(mxc_symbolic));
    if (!L) return MXC_OUT_OF_MEMORY;
    L->n = n;
    L->ordering = ordering;
    L->selected = 0;
    for (int m = 0; m < 3; m++) L->method_fl[m] = L->method_lnz[m] = -1;
    L->anz = -1;
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
        rc = given_perm(n, UserPerm, Perm, Pinv);
        if (rc != MXC_OK) goto fail;
    } else if (ordering == MXC_ORDER_NATURAL) {
        for (Int k = 0; k < n; k++) Perm[k] = k;
    } else if (ordering == MXC_ORDER_AMD) {
        double st[3];
        rc = mxc_amd_order_stats(n, Ap, Ai, Perm, st);
        if (rc != MXC_OK) goto fail;
        L->anz = st[2];
    } else if (ordering == MXC_ORDER_METIS) {
        rc = mxc_metis_order(n, Ap, Ai, 3000, 0.66, Perm);
        if (rc != MXC_OK) goto fail;
    } else if (ordering == MXC_ORDER_DEFAULT) {
        rc = default_strategy(n, Ap, Ai, UserPerm, Perm, Wk, L);
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
        rc = super_symbolic(n, Fp, Fi, Parent, ColCount, nrelax, zrelax, L);
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

int mxc_symbolic_strategy(const mxc_symbolic *L, int64_t *selected, double *method_fl,
                          double *method_lnz, double *anz) {
    if (!L) return MXC_INVALID;
    if (selected) *selected = L->selected;
    for (int m = 0; m < 3; m++) {
        if (method_fl) method_fl[m] = L->method_fl[m];
        if (method_lnz) method_lnz[m] = L->method_lnz[m];
    }
    if (anz) *anz = L->anz;
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
    for (int m = 0; m < 3; m++) L->method_fl[m] = L->method_lnz[m] = -1;
    L->anz = -1;
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
