// amd.c -- approximate minimum degree ordering, as CHOLMOD applies it when
// cholmod_l_analyze runs with CHOLMOD_AMD (CHOLMOD/Cholesky/cholmod_amd.c:40-194).
//
// The input to the ordering is the pattern of A + A' without its diagonal,
// laid out as cholmod_l_copy(A, 0, -2) lays it out (each column: the entries
// from earlier columns first, then the column's own lower entries; ~50%
// elbow room; CHOLMOD/Utility/t_cholmod_copy.c:195-260), and the defaults
// dense = 10*sqrt(n) (at least 16) and aggressive absorption on.
//
// The algorithm is approximate minimum degree (Amestoy, Davis, Duff, SIAM J.
// Matrix Anal. Appl. 17(4), 1996).  The ordering it returns depends on every
// tie-break -- the order variables enter and leave the degree lists, the
// hash used to find indistinguishable variables and the order of each hash
// bucket, the garbage-collection layout, the postorder of the assembly tree
// with the largest child last -- so the elimination loop below performs the
// same operations in the same order as AMD/Source/amd_2.c:25-1842 and
// AMD/Source/amd_postorder.c, amd_post_tree.c.  The code is written as a set
// of small steps over one workspace struct.

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/cholmod_mi355x.h"

typedef int64_t Int;
#define NONE (-1)
#define TAG(i) (-(i) - 2)          /* involution marking "absorbed into i" / "element" */

typedef struct {
    Int n;
    Int *pe, *len, *nv, *next, *last, *head, *elen, *deg, *w;
    Int *iw;
    Int iwlen, pfree;
    Int wflg, wbig;
    Int mindeg, nel, lemax;
} amd_state;

// W[x] < flag for every x after this (W entries that are nonzero drop to 1)
static Int marks_reset(Int flag, Int limit, Int *W, Int n) {
    if (flag < 2 || flag >= limit) {
        for (Int x = 0; x < n; x++)
            if (W[x] != 0) W[x] = 1;
        flag = 2;
    }
    return flag;
}

static void degree_list_unlink(amd_state *S, Int i) {
    Int prev = S->last[i], nxt = S->next[i];
    if (nxt != NONE) S->last[nxt] = prev;
    if (prev != NONE) S->next[prev] = nxt;
    else S->head[S->deg[i]] = nxt;
}

static void degree_list_push(amd_state *S, Int i, Int d) {
    Int first = S->head[d];
    if (first != NONE) S->last[first] = i;
    S->next[i] = first;
    S->head[d] = i;
}

// Compact Iw: keep every live list (each marked by tagging its first entry),
// then append the partially built element Iw[*pme1 .. pfree-1].
static void compact(amd_state *S, Int *pme1) {
    Int n = S->n, *pe = S->pe, *iw = S->iw, *len = S->len;
    for (Int j = 0; j < n; j++) {
        Int pn = pe[j];
        if (pn >= 0) {
            pe[j] = iw[pn];
            iw[pn] = TAG(j);
        }
    }
    Int src = 0, dst = 0, stop = *pme1 - 1;
    while (src <= stop) {
        Int j = TAG(iw[src++]);
        if (j >= 0) {
            iw[dst] = pe[j];
            pe[j] = dst++;
            for (Int k = 0; k <= len[j] - 2; k++) iw[dst++] = iw[src++];
        }
    }
    Int start = dst;
    for (src = *pme1; src <= S->pfree - 1; src++) iw[dst++] = iw[src];
    *pme1 = start;
    S->pfree = dst;
}

// Form element `me` from its variables and absorbed elements.  Returns the
// external degree and sets [*pme1, *pme2] to the element's pattern.
static Int build_element(amd_state *S, Int me, Int elenme, Int *pme1_out, Int *pme2_out) {
    Int *pe = S->pe, *iw = S->iw, *len = S->len, *nv = S->nv;
    Int degme = 0, pme1, pme2;
    if (elenme == 0) {
        // in place: the list of `me` holds only variables
        pme1 = pe[me];
        pme2 = pme1 - 1;
        for (Int p = pme1; p <= pme1 + len[me] - 1; p++) {
            Int i = iw[p], nvi = nv[i];
            if (nvi > 0) {
                degme += nvi;
                nv[i] = -nvi;
                iw[++pme2] = i;
                degree_list_unlink(S, i);
            }
        }
    } else {
        // in the free space, merging the element lists and me's own variables
        Int p = pe[me];
        pme1 = S->pfree;
        Int slenme = len[me] - elenme;
        for (Int knt1 = 1; knt1 <= elenme + 1; knt1++) {
            Int e, pj, ln;
            if (knt1 > elenme) { e = me; pj = p; ln = slenme; }
            else { e = iw[p++]; pj = pe[e]; ln = len[e]; }
            for (Int knt2 = 1; knt2 <= ln; knt2++) {
                Int i = iw[pj++], nvi = nv[i];
                if (nvi <= 0) continue;
                if (S->pfree >= S->iwlen) {
                    // trim the lists being walked to what is left, then compact
                    pe[me] = p;
                    len[me] -= knt1;
                    if (len[me] == 0) pe[me] = NONE;
                    pe[e] = pj;
                    len[e] = ln - knt2;
                    if (len[e] == 0) pe[e] = NONE;
                    compact(S, &pme1);
                    pj = pe[e];
                    p = pe[me];
                }
                degme += nvi;
                nv[i] = -nvi;
                iw[S->pfree++] = i;
                degree_list_unlink(S, i);
            }
            if (e != me) {
                pe[e] = TAG(me);       // e absorbed into me
                S->w[e] = 0;
            }
        }
        pme2 = S->pfree - 1;
    }
    *pme1_out = pme1;
    *pme2_out = pme2;
    return degme;
}

// stats (optional): the fill and work AMD predicts, as cholmod_l_amd stores
// them in Common (CHOLMOD/Cholesky/cholmod_amd.c:145-180, from the Info
// counts of AMD/Source/amd_2.c:1515-1590): stats[0] = fl = divisions + 2 *
// LDL' multiply-subtracts + n, stats[1] = lnz = n + nnz(L) below the
// diagonal, stats[2] = anz = n + nnz(strictly lower A).
static int amd_order(int64_t n, const int64_t *Ap, const int64_t *Ai, int64_t *Perm, double *stats) {
    if (n < 0 || !Perm || (n > 0 && (!Ap || !Ai))) return MXC_INVALID;
    if (stats) stats[0] = stats[1] = stats[2] = 0;
    if (n == 0) return MXC_OK;

    // ---- pattern of A + A' without the diagonal (column layout of cholmod_l_copy)
    Int *cnt = (Int *)calloc((size_t)n + 1, sizeof(Int));
    if (!cnt) return MXC_OUT_OF_MEMORY;
    Int cnz = 0;
    for (Int j = 0; j < n; j++)
        for (Int p = Ap[j]; p < Ap[j + 1]; p++) {
            Int i = Ai[p];
            if (i > j) { cnt[j]++; cnt[i]++; cnz += 2; }
        }
    if (stats) stats[2] = (double)(cnz / 2 + n);
    Int iwlen = cnz + cnz / 2 + n;
    amd_state S;
    memset(&S, 0, sizeof(S));
    S.n = n;
    S.iwlen = iwlen;
    S.pfree = cnz;
    Int *mem = (Int *)malloc(sizeof(Int) * (size_t)(10 * n + (iwlen > 0 ? iwlen : 1)));
    if (!mem) { free(cnt); return MXC_OUT_OF_MEMORY; }
    S.pe = mem; S.len = mem + n; S.nv = mem + 2 * n; S.next = mem + 3 * n; S.last = mem + 4 * n;
    S.head = mem + 5 * n; S.elen = mem + 6 * n; S.deg = mem + 7 * n; S.w = mem + 8 * n;
    Int *fill = mem + 9 * n;
    S.iw = mem + 10 * n;
    {
        Int acc = 0;
        for (Int j = 0; j < n; j++) { S.pe[j] = acc; fill[j] = acc; acc += cnt[j]; }
        for (Int j = 0; j < n; j++)
            for (Int p = Ap[j]; p < Ap[j + 1]; p++) {
                Int i = Ai[p];
                if (i > j) {
                    S.iw[fill[j]++] = i;     // C(i, j)
                    S.iw[fill[i]++] = j;     // C(j, i)
                }
            }
        for (Int j = 0; j < n; j++) S.len[j] = cnt[j];
    }
    free(cnt);

    Int *pe = S.pe, *iw = S.iw, *len = S.len, *nv = S.nv, *next = S.next, *last = S.last;
    Int *head = S.head, *elen = S.elen, *deg = S.deg, *W = S.w;

    Int dense = (Int)(10.0 * sqrt((double)n));
    if (dense < 16) dense = 16;
    if (dense > n) dense = n;

    for (Int i = 0; i < n; i++) {
        last[i] = NONE; head[i] = NONE; next[i] = NONE;
        nv[i] = 1; W[i] = 1; elen[i] = 0; deg[i] = len[i];
    }
    S.wbig = INT64_MAX - n;
    S.wflg = marks_reset(0, S.wbig, W, n);
    S.mindeg = 0;
    S.nel = 0;
    S.lemax = 0;
    Int ndense = 0;
    // predicted factor statistics, accumulated in the order amd_2 does
    double st_lnz = 0, st_ndiv = 0, st_nms_ldl = 0;

    // empty rows are eliminated at once; dense rows are set aside (ordered last)
    for (Int i = 0; i < n; i++) {
        Int d = deg[i];
        if (d == 0) {
            elen[i] = TAG(1);
            S.nel++;
            pe[i] = NONE;
            W[i] = 0;
        } else if (d > dense) {
            ndense++;
            nv[i] = 0;
            elen[i] = NONE;
            S.nel++;
            pe[i] = NONE;
        } else {
            degree_list_push(&S, i, d);
        }
    }

    while (S.nel < n) {
        // ---- pivot: head of the lowest non-empty degree list
        Int d, me = NONE;
        for (d = S.mindeg; d < n; d++) {
            me = head[d];
            if (me != NONE) break;
        }
        S.mindeg = d;
        {
            Int nx = next[me];
            if (nx != NONE) last[nx] = NONE;
            head[d] = nx;
        }
        Int elenme = elen[me];
        Int nvpiv = nv[me];
        S.nel += nvpiv;
        nv[me] = -nvpiv;

        Int pme1, pme2;
        Int degme = build_element(&S, me, elenme, &pme1, &pme2);
        deg[me] = degme;
        pe[me] = pme1;
        len[me] = pme2 - pme1 + 1;
        elen[me] = TAG(nvpiv + degme);
        S.wflg = marks_reset(S.wflg, S.wbig, W, n);

        // ---- scan 1: W[e] - wflg = |Le \ Lme| for elements e touching Lme
        for (Int pme = pme1; pme <= pme2; pme++) {
            Int i = iw[pme], eln = elen[i];
            if (eln <= 0) continue;
            Int nvi = -nv[i], wnvi = S.wflg - nvi;
            for (Int p = pe[i]; p <= pe[i] + eln - 1; p++) {
                Int e = iw[p], we = W[e];
                if (we >= S.wflg) we -= nvi;
                else if (we != 0) we = deg[e] + wnvi;
                W[e] = we;
            }
        }

        // ---- scan 2: approximate degrees, absorption, mass elimination, hashing
        for (Int pme = pme1; pme <= pme2; pme++) {
            Int i = iw[pme];
            Int p1 = pe[i], p2 = p1 + elen[i] - 1, pn = p1;
            uint64_t hash = 0;
            Int dg = 0;
            for (Int p = p1; p <= p2; p++) {
                Int e = iw[p], we = W[e];
                if (we == 0) continue;
                Int dext = we - S.wflg;
                if (dext > 0) {
                    dg += dext;
                    iw[pn++] = e;
                    hash += (uint64_t)e;
                } else {
                    pe[e] = TAG(me);      // aggressive absorption into me
                    W[e] = 0;
                }
            }
            elen[i] = pn - p1 + 1;
            Int p3 = pn, p4 = p1 + len[i];
            for (Int p = p2 + 1; p < p4; p++) {
                Int j = iw[p], nvj = nv[j];
                if (nvj > 0) {
                    dg += nvj;
                    iw[pn++] = j;
                    hash += (uint64_t)j;
                }
            }
            if (elen[i] == 1 && p3 == pn) {
                // only me is left: eliminate i together with me
                pe[i] = TAG(me);
                Int nvi = -nv[i];
                degme -= nvi;
                nvpiv += nvi;
                S.nel += nvi;
                nv[i] = 0;
                elen[i] = NONE;
            } else {
                if (dg < deg[i]) deg[i] = dg;
                iw[pn] = iw[p3];
                iw[p3] = iw[p1];
                iw[p1] = me;
                len[i] = pn - p1 + 1;
                Int hkey = (Int)(hash % (uint64_t)n);
                Int j = head[hkey];
                if (j <= NONE) {            // degree list empty: bucket head kept in head[]
                    next[i] = TAG(j);
                    head[hkey] = TAG(i);
                } else {                    // else kept in last[head]
                    next[i] = last[j];
                    last[j] = i;
                }
                last[i] = hkey;
            }
        }
        deg[me] = degme;
        if (degme > S.lemax) S.lemax = degme;
        S.wflg += S.lemax;
        S.wflg = marks_reset(S.wflg, S.wbig, W, n);

        // ---- indistinguishable supervariables: compare within each hash bucket
        for (Int pme = pme1; pme <= pme2; pme++) {
            Int i = iw[pme];
            if (nv[i] >= 0) continue;
            Int hkey = last[i];
            Int j = head[hkey];
            if (j == NONE) {
                i = NONE;
            } else if (j < NONE) {
                i = TAG(j);
                head[hkey] = NONE;
            } else {
                i = last[j];
                last[j] = NONE;
            }
            while (i != NONE && next[i] != NONE) {
                Int ln = len[i], eln = elen[i];
                for (Int p = pe[i] + 1; p <= pe[i] + ln - 1; p++) W[iw[p]] = S.wflg;
                Int jlast = i;
                j = next[i];
                while (j != NONE) {
                    int same = (len[j] == ln) && (elen[j] == eln);
                    for (Int p = pe[j] + 1; same && p <= pe[j] + ln - 1; p++)
                        if (W[iw[p]] != S.wflg) same = 0;
                    if (same) {
                        pe[j] = TAG(i);
                        nv[i] += nv[j];
                        nv[j] = 0;
                        elen[j] = NONE;
                        j = next[j];
                        next[jlast] = j;
                    } else {
                        jlast = j;
                        j = next[j];
                    }
                }
                S.wflg++;
                i = next[i];
            }
        }

        // ---- finalize degrees, reinsert, and keep only principal variables in Lme
        Int p = pme1;
        Int nleft = n - S.nel;
        for (Int pme = pme1; pme <= pme2; pme++) {
            Int i = iw[pme];
            Int nvi = -nv[i];
            if (nvi <= 0) continue;
            nv[i] = nvi;
            Int dg = deg[i] + degme - nvi;
            if (dg > nleft - nvi) dg = nleft - nvi;
            Int first = head[dg];
            if (first != NONE) last[first] = i;
            next[i] = first;
            last[i] = NONE;
            head[dg] = i;
            if (dg < S.mindeg) S.mindeg = dg;
            deg[i] = dg;
            iw[p++] = i;
        }
        nv[me] = nvpiv;
        {
            const double f = (double)nvpiv, r = (double)(degme + ndense);
            const double lnzme = f * r + (f - 1) * f / 2;
            st_lnz += lnzme;
            st_ndiv += lnzme;
            const double s = f * r * r + r * (f - 1) * f + (f - 1) * f * (2 * f - 1) / 6;
            st_nms_ldl += (s + lnzme) / 2;
        }
        len[me] = p - pme1;
        if (len[me] == 0) {
            pe[me] = NONE;
            W[me] = 0;
        }
        if (elenme != 0) S.pfree = p;
    }

    if (stats) {
        // the dense rows form one final dense block
        const double f = (double)ndense;
        const double lnzme = (f - 1) * f / 2;
        st_lnz += lnzme;
        st_ndiv += lnzme;
        const double s = (f - 1) * f * (2 * f - 1) / 6;
        st_nms_ldl += (s + lnzme) / 2;
        stats[0] = st_ndiv + 2 * st_nms_ldl + (double)n;
        stats[1] = (double)n + st_lnz;
    }

    // ---- assembly tree: parents in pe, front sizes in elen
    for (Int i = 0; i < n; i++) pe[i] = TAG(pe[i]);
    for (Int i = 0; i < n; i++) elen[i] = TAG(elen[i]);
    for (Int i = 0; i < n; i++) {
        if (nv[i] != 0) continue;
        Int j = pe[i];
        if (j == NONE) continue;          // dense variable: no parent
        while (nv[j] == 0) j = pe[j];
        Int e = j;
        j = i;
        while (nv[j] == 0) {
            Int jn = pe[j];
            pe[j] = e;
            j = jn;
        }
    }

    // ---- postorder of the elements, largest front (last on ties) visited last
    Int *child = head, *sibling = next, *stack = last, *order = W;
    for (Int j = 0; j < n; j++) { child[j] = NONE; sibling[j] = NONE; }
    for (Int j = n - 1; j >= 0; j--) {
        if (nv[j] > 0 && pe[j] != NONE) {
            sibling[j] = child[pe[j]];
            child[pe[j]] = j;
        }
    }
    for (Int i = 0; i < n; i++) {
        if (nv[i] <= 0 || child[i] == NONE) continue;
        Int fprev = NONE, maxf = NONE, bigprev = NONE, big = NONE;
        for (Int f = child[i]; f != NONE; f = sibling[f]) {
            if (elen[f] >= maxf) { maxf = elen[f]; bigprev = fprev; big = f; }
            fprev = f;
        }
        Int fnext = sibling[big];
        if (fnext != NONE) {
            if (bigprev == NONE) child[i] = fnext;
            else sibling[bigprev] = fnext;
            sibling[big] = NONE;
            sibling[fprev] = big;
        }
    }
    for (Int i = 0; i < n; i++) order[i] = NONE;
    Int k = 0;
    for (Int r = 0; r < n; r++) {
        if (pe[r] != NONE || nv[r] <= 0) continue;
        Int top = 0;
        stack[0] = r;
        while (top >= 0) {
            Int i = stack[top];
            if (child[i] != NONE) {
                Int c = 0;
                for (Int f = child[i]; f != NONE; f = sibling[f]) c++;
                Int h = top + c;
                top = h;
                for (Int f = child[i]; f != NONE; f = sibling[f]) stack[h--] = f;
                child[i] = NONE;
            } else {
                top--;
                order[i] = k++;
            }
        }
    }

    // ---- permutation: elements in postorder, each preceded by the variables
    // merged into it; dense variables last
    Int *byk = head, *pos = next;
    for (Int t = 0; t < n; t++) { byk[t] = NONE; pos[t] = NONE; }
    for (Int e = 0; e < n; e++)
        if (order[e] != NONE) byk[order[e]] = e;
    Int nel = 0;
    for (Int t = 0; t < n; t++) {
        Int e = byk[t];
        if (e == NONE) break;
        pos[e] = nel;
        nel += nv[e];
    }
    for (Int i = 0; i < n; i++) {
        if (nv[i] != 0) continue;
        Int e = pe[i];
        if (e != NONE) {
            pos[i] = pos[e];
            pos[e]++;
        } else {
            pos[i] = nel++;
        }
    }
    for (Int i = 0; i < n; i++) Perm[pos[i]] = i;
    free(mem);
    return MXC_OK;
}

int mxc_amd_order(int64_t n, const int64_t *Ap, const int64_t *Ai, int64_t *Perm) {
    return amd_order(n, Ap, Ai, Perm, NULL);
}

int mxc_amd_order_stats(int64_t n, const int64_t *Ap, const int64_t *Ai, int64_t *Perm, double *stats) {
    return amd_order(n, Ap, Ai, Perm, stats);
}
