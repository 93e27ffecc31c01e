// metis_nd.c -- nested-dissection ordering, as CHOLMOD applies it when
// cholmod_l_analyze runs with CHOLMOD_METIS (CHOLMOD/Partition/cholmod_metis.c:557-846).
//
// CHOLMOD hands METIS_NodeND the pattern of A + A' without its diagonal, in
// the column layout of cholmod_l_copy(A, 0, -1) (each column: the entries
// from earlier columns first, then the column's own lower entries;
// CHOLMOD/Utility/t_cholmod_copy_worker.c), with default options.  The
// vendored METIS is 5.1.0 with 64-bit indices, 32-bit reals and its own
// portable generator (CHOLMOD/Partition/cholmod_metis_wrapper.h:15-16,
// SuiteSparse_metis/include/metis.h:34,44), so its output is a deterministic
// function of the graph.  The ordering below reproduces it exactly, which
// pins every step that consumes a random number or breaks a tie:
//
//   generator      MT19937-64 seeded 4321, values masked to 63 bits, draws in
//                  [0, m) by remainder; array shuffles as GKlib/gk_mkrandom.h:23-110
//   sorting        the (key, value) quicksort of GKlib/gk_mksort.h:120-270
//                  (median of three, insertion sort below 5) -- its order
//                  among equal keys decides compression and 2-hop matching
//   queues         binary max-heaps with float keys, GKlib/gk_mkpqueue.h:15-330
//   arithmetic     the float and double expressions of the balance tests in
//                  the types METIS evaluates them in (real_t = float)
//
// The steps (SuiteSparse_metis/libmetis):
//   compression of identical closed neighbourhoods     compress.c:24-147, ometis.c:100-118
//   recursive bisection, separator first              ometis.c:186-227
//   best of several multilevel bisections             ometis.c:300-427
//   coarsening by random / heavy-edge matching + 2-hop coarsen.c:20-420, 430-640
//   contraction                                       coarsen.c:650-850
//   initial edge bisection by region growing + FM     initpart.c:67-320, balance.c, fm.c:30-200, refine.c
//   separator from the edge cut, node FM refinement   separator.c:23-54, srefine.c, sfm.c
//   split into the two sides                          ometis.c:440-540
//   multiple minimum degree on the leaves (<=120)     ometis.c:640-700, mmd.c
//
// The code is organised as one context (options, generator, workspace) and
// one graph record per level.  Host-only C.

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/cholmod_mi355x.h"

typedef int64_t Int;
typedef float real;

#define ND_MMD_LEAF       120      /* leaves at or below this size go to MMD */
#define ND_INIT_TRIES     7        /* region-growing starts per initial bisection */
#define ND_COARSEN_RATIO  0.85     /* stop coarsening when a level shrinks less */
#define ND_COMPRESS_RATIO 0.85     /* compress only if it removes >= 15% */
#define ND_TWO_HOP        0.10     /* unmatched fraction that triggers 2-hop matching */
#define ND_UNMATCHED      (-1)

/* ======================================================================== */
/* generator: MT19937-64                                                     */
/* ======================================================================== */

#define MT_N 312
#define MT_M 156

typedef struct {
    uint64_t s[MT_N];
    int i;
} mt64;

static void mt_seed(mt64 *g, uint64_t seed) {
    g->s[0] = seed;
    for (int k = 1; k < MT_N; k++)
        g->s[k] = 6364136223846793005ULL * (g->s[k - 1] ^ (g->s[k - 1] >> 62)) + (uint64_t)k;
    g->i = MT_N;
}

static void mt_refill(mt64 *g) {
    static const uint64_t upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL;
    static const uint64_t twist = 0xB5026F5AA96619E9ULL;
    for (int k = 0; k < MT_N; k++) {
        uint64_t y = (g->s[k] & upper) | (g->s[(k + 1) % MT_N] & lower);
        g->s[k] = g->s[(k + MT_M) % MT_N] ^ (y >> 1) ^ ((y & 1) ? twist : 0);
    }
    g->i = 0;
}

/* next draw, 63 bits */
static Int mt_draw(mt64 *g) {
    if (g->i >= MT_N) mt_refill(g);
    uint64_t y = g->s[g->i++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= y >> 43;
    return (Int)(y & 0x7FFFFFFFFFFFFFFFULL);
}

static Int draw_below(mt64 *g, Int m) { return mt_draw(g) % m; }

/* identity (if init) followed by the shuffle: n random pair swaps for short
 * arrays, else `rounds` exchanges of two overlapping 4-blocks */
static void shuffle(mt64 *g, Int n, Int *p, Int rounds, int init) {
    if (init)
        for (Int i = 0; i < n; i++) p[i] = i;
    Int t;
    if (n < 10) {
        for (Int r = 0; r < n; r++) {
            Int a = draw_below(g, n), b = draw_below(g, n);
            t = p[a]; p[a] = p[b]; p[b] = t;
        }
        return;
    }
    for (Int r = 0; r < rounds; r++) {
        Int a = draw_below(g, n - 3), b = draw_below(g, n - 3);
        t = p[a];     p[a] = p[b + 2];     p[b + 2] = t;
        t = p[a + 1]; p[a + 1] = p[b + 3]; p[b + 3] = t;
        t = p[a + 2]; p[a + 2] = p[b];     p[b] = t;
        t = p[a + 3]; p[a + 3] = p[b + 1]; p[b + 1] = t;
    }
}

/* ======================================================================== */
/* sort of (key, value) pairs by key, the introsort-free quicksort GKlib uses */
/* ======================================================================== */

typedef struct { Int key, val; } kv;

static void kv_sort(kv *a, Int n) {
    if (n <= 0) return;
    kv t;
    if (n > 4) {
        struct { kv *lo, *hi; } stack[64];
        int top = 0;
        kv *lo = a, *hi = a + n - 1;
        stack[top].lo = NULL; stack[top].hi = NULL; top++;   /* sentinel frame */
        while (top > 0) {
            kv *mid = lo + ((hi - lo) >> 1);
            /* order lo, mid, hi */
            if (mid->key < lo->key) { t = *mid; *mid = *lo; *lo = t; }
            if (hi->key < mid->key) {
                t = *mid; *mid = *hi; *hi = t;
                if (mid->key < lo->key) { t = *mid; *mid = *lo; *lo = t; }
            }
            kv *l = lo + 1, *r = hi - 1;
            do {
                while (l->key < mid->key) l++;
                while (mid->key < r->key) r--;
                if (l < r) {
                    t = *l; *l = *r; *r = t;
                    if (mid == l) mid = r;
                    else if (mid == r) mid = l;
                    l++; r--;
                } else if (l == r) {
                    l++; r--;
                    break;
                }
            } while (l <= r);
            /* recurse into the larger side later; small sides are left to the
             * final insertion pass */
            const int small_left = (r - lo) <= 4, small_right = (hi - l) <= 4;
            if (small_left && small_right) {
                top--;
                lo = stack[top].lo; hi = stack[top].hi;
            } else if (small_left) {
                lo = l;
            } else if (small_right) {
                hi = r;
            } else if ((r - lo) > (hi - l)) {
                stack[top].lo = lo; stack[top].hi = r; top++;
                lo = l;
            } else {
                stack[top].lo = l; stack[top].hi = hi; top++;
                hi = r;
            }
        }
    }
    /* insertion sort, smallest of the first five moved to the front first */
    kv *end = a + n - 1, *lim = a + 4 < end ? a + 4 : end, *m = a;
    for (kv *q = a + 1; q <= lim; q++)
        if (q->key < m->key) m = q;
    if (m != a) { t = *m; *m = *a; *a = t; }
    for (kv *q = a + 2; q <= end; q++) {
        kv *s = q - 1;
        while (q->key < s->key) s--;
        s++;
        if (s != q) {
            t = *q;
            memmove(s + 1, s, (size_t)(q - s) * sizeof(kv));
            *s = t;
        }
    }
}

/* ======================================================================== */
/* max-heap of vertices keyed by float gain                                   */
/* ======================================================================== */

typedef struct {
    Int n;
    real *key;
    Int *val;
    Int *where;          /* heap slot of each vertex, -1 if absent */
} heap;

static int heap_init(heap *h, Int cap) {
    h->n = 0;
    h->key = (real *)malloc(sizeof(real) * (size_t)(cap > 0 ? cap : 1));
    h->val = (Int *)malloc(sizeof(Int) * (size_t)(cap > 0 ? cap : 1));
    h->where = (Int *)malloc(sizeof(Int) * (size_t)(cap > 0 ? cap : 1));
    if (!h->key || !h->val || !h->where) return 0;
    for (Int i = 0; i < cap; i++) h->where[i] = -1;
    return 1;
}

static void heap_free(heap *h) {
    free(h->key); free(h->val); free(h->where);
    h->key = NULL; h->val = NULL; h->where = NULL;
}

static void heap_clear(heap *h) {
    for (Int i = h->n - 1; i >= 0; i--) h->where[h->val[i]] = -1;
    h->n = 0;
}

static inline void heap_put(heap *h, Int i, real k, Int v) {
    h->key[i] = k;
    h->val[i] = v;
    h->where[v] = i;
}

/* move the hole at slot i up while the parent is smaller than k */
static Int sift_up(heap *h, Int i, real k) {
    while (i > 0) {
        Int p = (i - 1) >> 1;
        if (!(k > h->key[p])) break;
        heap_put(h, i, h->key[p], h->val[p]);
        i = p;
    }
    return i;
}

/* move the hole at slot i down while a child is larger than k */
static Int sift_down(heap *h, Int i, real k) {
    const Int n = h->n;
    Int c;
    while ((c = 2 * i + 1) < n) {
        if (h->key[c] > k) {
            if (c + 1 < n && h->key[c + 1] > h->key[c]) c++;
        } else if (c + 1 < n && h->key[c + 1] > k) {
            c++;
        } else {
            break;
        }
        heap_put(h, i, h->key[c], h->val[c]);
        i = c;
    }
    return i;
}

static void heap_insert(heap *h, Int v, real k) {
    Int i = sift_up(h, h->n++, k);
    heap_put(h, i, k, v);
}

static void heap_delete(heap *h, Int v) {
    Int i = h->where[v];
    h->where[v] = -1;
    if (--h->n > 0 && h->val[h->n] != v) {
        Int u = h->val[h->n];
        real k = h->key[h->n];
        i = (k > h->key[i]) ? sift_up(h, i, k) : sift_down(h, i, k);
        heap_put(h, i, k, u);
    }
}

static void heap_update(heap *h, Int v, real k) {
    Int i = h->where[v];
    i = (k > h->key[i]) ? sift_up(h, i, k) : sift_down(h, i, k);
    heap_put(h, i, k, v);
}

static Int heap_top(const heap *h) { return h->n == 0 ? -1 : h->val[0]; }

static Int heap_pop(heap *h) {
    if (h->n == 0) return -1;
    Int v = h->val[0];
    h->where[v] = -1;
    if (--h->n > 0) {
        Int u = h->val[h->n];
        real k = h->key[h->n];
        Int i = sift_down(h, 0, k);
        heap_put(h, i, k, u);
    }
    return v;
}

/* ======================================================================== */
/* graphs                                                                    */
/* ======================================================================== */

typedef struct graph {
    Int nvtxs, nedges, tvwgt;
    real invtvwgt;
    Int *xadj, *adjncy, *adjwgt, *vwgt;
    Int *label;                 /* original vertex of each vertex (split graphs) */
    Int *cmap;                  /* vertex -> coarse vertex */
    /* partition state */
    Int *where, *bndptr, *bndind, nbnd, mincut;
    Int pwgts[3];
    Int *id, *ed;               /* edge bisection: internal / external degree */
    Int *edeg;                  /* separator: weight of neighbours on side 0, 1 */
    struct graph *coarser, *finer;
} graph;

typedef struct {
    mt64 rng;
    real ub;                    /* imbalance factor, float as METIS stores it */
    real pijbm[2];              /* balance multipliers of the current graph */
    Int coarsen_to, maxvwgt;
    int compress;               /* the input graph was compressed */
    Int nseps;
    int oom;
} nd_ctx;

static void *nd_alloc(nd_ctx *c, Int count, size_t elt) {
    void *p = malloc((size_t)(count > 0 ? count : 1) * elt);
    if (!p) c->oom = 1;
    return p;
}
#define IALLOC(c, n) ((Int *)nd_alloc((c), (n), sizeof(Int)))

static void graph_free_partition(graph *g) {
    free(g->where); free(g->bndptr); free(g->bndind); free(g->id); free(g->ed); free(g->edeg);
    g->where = g->bndptr = g->bndind = g->id = g->ed = g->edeg = NULL;
}

static void graph_free(graph *g) {
    if (!g) return;
    free(g->xadj); free(g->adjncy); free(g->adjwgt); free(g->vwgt);
    free(g->label); free(g->cmap);
    graph_free_partition(g);
    free(g);
}

static void graph_set_total(graph *g) {
    Int t = 0;
    for (Int i = 0; i < g->nvtxs; i++) t += g->vwgt[i];
    g->tvwgt = t;
    g->invtvwgt = (real)(1.0 / (t > 0 ? t : 1));
}

/* the partition arrays (allocated once per graph, reused by every bisection) */
static int graph_partition_mem(nd_ctx *c, graph *g) {
    if (g->where) return 1;
    Int n = g->nvtxs;
    g->where = IALLOC(c, n); g->bndptr = IALLOC(c, n); g->bndind = IALLOC(c, n);
    g->id = IALLOC(c, n); g->ed = IALLOC(c, n); g->edeg = IALLOC(c, 2 * n);
    return !c->oom;
}

#define BND_ADD(g, v) do { (g)->bndind[(g)->nbnd] = (v); (g)->bndptr[(v)] = (g)->nbnd++; } while (0)
#define BND_DEL(g, v) do { Int s_ = (g)->bndptr[(v)]; Int l_ = (g)->bndind[--(g)->nbnd]; \
                           (g)->bndind[s_] = l_; (g)->bndptr[l_] = s_; (g)->bndptr[(v)] = -1; } while (0)

static Int iabs64(Int x) { return x < 0 ? -x : x; }

/* ======================================================================== */
/* coarsening                                                                */
/* ======================================================================== */

/* Contract matched pairs.  Coarse adjacency lists keep the first-seen order
 * of the pair's neighbours (v's list, then u's), the pair's own edge removed
 * by moving the last entry into its slot. */
static graph *contract(nd_ctx *c, graph *g, Int cn, const Int *match) {
    const Int n = g->nvtxs;
    const Int *xadj = g->xadj, *adj = g->adjncy, *w = g->adjwgt, *cmap = g->cmap;
    graph *cg = (graph *)calloc(1, sizeof(graph));
    if (!cg) { c->oom = 1; return NULL; }
    cg->nvtxs = cn;
    cg->xadj = IALLOC(c, cn + 1);
    cg->vwgt = IALLOC(c, cn);
    cg->adjncy = IALLOC(c, g->nedges);
    cg->adjwgt = IALLOC(c, g->nedges);
    Int *slot = IALLOC(c, cn);
    if (c->oom) { free(slot); graph_free(cg); return NULL; }
    for (Int k = 0; k < cn; k++) slot[k] = -1;
    Int *cadj = cg->adjncy, *cw = cg->adjwgt;
    Int ne = 0, k = 0;
    cg->xadj[0] = 0;
    for (Int v = 0; v < n; v++) {
        Int u = match[v];
        if (u < v) continue;
        Int base = ne, cnt = 0;
        cg->vwgt[k] = g->vwgt[v] + (u != v ? g->vwgt[u] : 0);
        for (int side = 0; side < (u != v ? 2 : 1); side++) {
            Int x = side ? u : v;
            for (Int p = xadj[x]; p < xadj[x + 1]; p++) {
                Int t = cmap[adj[p]];
                Int s = slot[t];
                if (s == -1) {
                    cadj[base + cnt] = t;
                    cw[base + cnt] = w[p];
                    slot[t] = cnt++;
                } else {
                    cw[base + s] += w[p];
                }
            }
        }
        if (u != v && slot[k] != -1) {
            Int s = slot[k];
            cnt--;
            cadj[base + s] = cadj[base + cnt];
            cw[base + s] = cw[base + cnt];
            slot[k] = -1;
        }
        for (Int q = 0; q < cnt; q++) slot[cadj[base + q]] = -1;
        ne += cnt;
        cg->xadj[++k] = ne;
    }
    free(slot);
    cg->nedges = ne;
    graph_set_total(cg);
    cg->finer = g;
    g->coarser = cg;
    return cg;
}

/* 2-hop matching of unmatched low-degree vertices that share a neighbour */
static Int match_two_hop_any(nd_ctx *c, graph *g, const Int *perm, Int *match, Int cn,
                             size_t *unmatched, size_t maxdeg) {
    const Int n = g->nvtxs, *xadj = g->xadj, *adj = g->adjncy;
    Int *cmap = g->cmap;
    Int *ptr = IALLOC(c, n + 1);
    if (!ptr) return cn;
    for (Int i = 0; i < n; i++) ptr[i] = 0;
    for (Int i = 0; i < n; i++)
        if (match[i] == ND_UNMATCHED && (size_t)(xadj[i + 1] - xadj[i]) < maxdeg)
            for (Int p = xadj[i]; p < xadj[i + 1]; p++) ptr[adj[p]]++;
    for (Int i = 1; i < n; i++) ptr[i] += ptr[i - 1];
    for (Int i = n; i > 0; i--) ptr[i] = ptr[i - 1];
    ptr[0] = 0;
    Int *who = IALLOC(c, ptr[n]);
    if (!who) { free(ptr); return cn; }
    for (Int pi = 0; pi < n; pi++) {
        Int i = perm[pi];
        if (match[i] == ND_UNMATCHED && (size_t)(xadj[i + 1] - xadj[i]) < maxdeg)
            for (Int p = xadj[i]; p < xadj[i + 1]; p++) who[ptr[adj[p]]++] = i;
    }
    for (Int i = n; i > 0; i--) ptr[i] = ptr[i - 1];
    ptr[0] = 0;
    size_t left = *unmatched;
    for (Int pi = 0; pi < n; pi++) {
        Int i = perm[pi];
        if (ptr[i + 1] - ptr[i] < 2) continue;
        Int hi = ptr[i + 1];
        for (Int lo = ptr[i]; lo < hi; lo++) {
            if (match[who[lo]] != ND_UNMATCHED) continue;
            for (hi--; hi > lo; hi--) {
                if (match[who[hi]] == ND_UNMATCHED) {
                    Int a = who[lo], b = who[hi];
                    cmap[a] = cmap[b] = cn++;
                    match[a] = b;
                    match[b] = a;
                    left -= 2;
                    break;
                }
            }
        }
    }
    free(ptr); free(who);
    *unmatched = left;
    return cn;
}

/* 2-hop matching of unmatched vertices with identical adjacency lists */
static Int match_two_hop_all(nd_ctx *c, graph *g, const Int *perm, Int *match, Int cn,
                             size_t *unmatched, size_t maxdeg) {
    const Int n = g->nvtxs, *xadj = g->xadj, *adj = g->adjncy;
    Int *cmap = g->cmap;
    const Int mask = (Int)((uint64_t)INT64_MAX / maxdeg);
    kv *keys = (kv *)nd_alloc(c, (Int)*unmatched, sizeof(kv));
    Int *mark = IALLOC(c, n);
    if (c->oom) { free(keys); free(mark); return cn; }
    Int ncand = 0;
    for (Int pi = 0; pi < n; pi++) {
        Int i = perm[pi];
        Int d = xadj[i + 1] - xadj[i];
        if (match[i] == ND_UNMATCHED && d > 1 && (size_t)d < maxdeg) {
            Int k = 0;
            for (Int p = xadj[i]; p < xadj[i + 1]; p++) k += adj[p] % mask;
            keys[ncand].val = i;
            keys[ncand].key = (Int)((uint64_t)(k % mask) * (uint64_t)maxdeg + (uint64_t)d);
            ncand++;
        }
    }
    kv_sort(keys, ncand);
    for (Int i = 0; i < n; i++) mark[i] = 0;
    size_t left = *unmatched;
    for (Int a = 0; a < ncand; a++) {
        Int i = keys[a].val;
        if (match[i] != ND_UNMATCHED) continue;
        for (Int p = xadj[i]; p < xadj[i + 1]; p++) mark[adj[p]] = i;
        for (Int b = a + 1; b < ncand; b++) {
            Int k = keys[b].val;
            if (match[k] != ND_UNMATCHED) continue;
            if (keys[a].key != keys[b].key) break;
            if (xadj[i + 1] - xadj[i] != xadj[k + 1] - xadj[k]) break;
            Int p = xadj[k];
            while (p < xadj[k + 1] && mark[adj[p]] == i) p++;
            if (p == xadj[k + 1]) {
                cmap[i] = cmap[k] = cn++;
                match[i] = k;
                match[k] = i;
                left -= 2;
                break;
            }
        }
    }
    free(keys); free(mark);
    *unmatched = left;
    return cn;
}

/* One level of coarsening.  heavy = 0: first eligible neighbour in list
 * order (unit edge weights); heavy = 1: heaviest-edge neighbour, vertices
 * visited by increasing (capped) degree. */
static graph *coarsen_level(nd_ctx *c, graph *g, int heavy) {
    const Int n = g->nvtxs, *xadj = g->xadj, *adj = g->adjncy, *vw = g->vwgt, *ew = g->adjwgt;
    const Int maxvw = c->maxvwgt;
    if (!g->cmap) g->cmap = IALLOC(c, n);
    Int *match = IALLOC(c, n), *perm = IALLOC(c, n);
    if (c->oom) { free(match); free(perm); return NULL; }
    Int *cmap = g->cmap;
    for (Int i = 0; i < n; i++) match[i] = ND_UNMATCHED;
    if (!heavy) {
        shuffle(&c->rng, n, perm, n / 8, 1);
    } else {
        Int *tperm = IALLOC(c, n), *deg = IALLOC(c, n);
        if (c->oom) { free(tperm); free(deg); free(match); free(perm); return NULL; }
        shuffle(&c->rng, n, tperm, n / 8, 1);
        Int cap = (Int)(0.7 * (double)(xadj[n] / n));
        for (Int i = 0; i < n; i++) {
            Int d = xadj[i + 1] - xadj[i];
            deg[i] = d > cap ? cap : d;
        }
        /* counting sort of tperm by capped degree */
        Int *cnt = IALLOC(c, cap + 2);
        if (!cnt) { free(tperm); free(deg); free(match); free(perm); return NULL; }
        for (Int k = 0; k < cap + 2; k++) cnt[k] = 0;
        for (Int i = 0; i < n; i++) cnt[deg[i]]++;
        for (Int k = 1; k < cap + 1; k++) cnt[k] += cnt[k - 1];
        for (Int k = cap + 1; k > 0; k--) cnt[k] = cnt[k - 1];
        cnt[0] = 0;
        for (Int t = 0; t < n; t++) {
            Int i = tperm[t];
            perm[cnt[deg[i]]++] = i;
        }
        free(cnt); free(tperm); free(deg);
    }

    Int cn = 0, last_unmatched = 0;
    size_t unmatched = 0;
    for (Int pi = 0; pi < n; pi++) {
        Int i = perm[pi];
        if (match[i] != ND_UNMATCHED) continue;
        Int mate = i;
        if (vw[i] < maxvw) {
            if (xadj[i] == xadj[i + 1]) {
                /* island: pair it with the next unmatched vertex in visit order */
                last_unmatched = (pi > last_unmatched ? pi : last_unmatched) + 1;
                for (; last_unmatched < n; last_unmatched++) {
                    Int j = perm[last_unmatched];
                    if (match[j] == ND_UNMATCHED) { mate = j; break; }
                }
            } else {
                Int best = -1;
                for (Int p = xadj[i]; p < xadj[i + 1]; p++) {
                    Int k = adj[p];
                    if (match[k] != ND_UNMATCHED || vw[i] + vw[k] > maxvw) continue;
                    if (!heavy) { mate = k; break; }
                    if (best < ew[p]) { mate = k; best = ew[p]; }
                }
                if (mate == i && 3 * vw[i] < maxvw) {
                    unmatched++;
                    mate = ND_UNMATCHED;
                }
            }
        }
        if (mate != ND_UNMATCHED) {
            cmap[i] = cmap[mate] = cn++;
            match[i] = mate;
            match[mate] = i;
        }
    }
    if ((double)unmatched > ND_TWO_HOP * (double)n) {
        cn = match_two_hop_any(c, g, perm, match, cn, &unmatched, 2);
        cn = match_two_hop_all(c, g, perm, match, cn, &unmatched, 64);
        if ((double)unmatched > 1.5 * ND_TWO_HOP * (double)n)
            cn = match_two_hop_any(c, g, perm, match, cn, &unmatched, 3);
        if ((double)unmatched > 2.0 * ND_TWO_HOP * (double)n)
            cn = match_two_hop_any(c, g, perm, match, cn, &unmatched, (size_t)n);
    }
    /* number the coarse vertices in order of their smaller member */
    cn = 0;
    for (Int i = 0; i < n; i++) {
        if (match[i] == ND_UNMATCHED) {
            match[i] = i;
            cmap[i] = cn++;
        } else if (i <= match[i]) {
            cmap[i] = cmap[match[i]] = cn++;
        }
    }
    graph *cg = c->oom ? NULL : contract(c, g, cn, match);
    free(match); free(perm);
    return cg;
}

static int uniform_edges(const graph *g) {
    for (Int p = 1; p < g->nedges; p++)
        if (g->adjwgt[0] != g->adjwgt[p]) return 0;
    return 1;
}

/* coarsen until small (coarsen_to), stalled, or sparse */
static graph *coarsen(nd_ctx *c, graph *g) {
    int heavy = !uniform_edges(g);
    c->maxvwgt = (Int)(1.5 * (double)g->tvwgt / (double)c->coarsen_to);
    do {
        graph *cg = coarsen_level(c, g, heavy && g->nedges != 0);
        if (!cg) return NULL;
        g = cg;
        heavy = 1;
    } while (g->nvtxs > c->coarsen_to && (double)g->nvtxs < ND_COARSEN_RATIO * (double)g->finer->nvtxs &&
             g->nedges > g->nvtxs / 2);
    return g;
}

/* at most `levels` levels, stopping early like coarsen() */
static graph *coarsen_levels(nd_ctx *c, graph *g, int levels) {
    int heavy = !uniform_edges(g);
    c->maxvwgt = (Int)(1.5 * (double)g->tvwgt / (double)c->coarsen_to);
    for (int l = 0; l < levels; l++) {
        graph *cg = coarsen_level(c, g, heavy && g->nedges != 0);
        if (!cg) return NULL;
        g = cg;
        heavy = 1;
        if (g->nvtxs < c->coarsen_to || (double)g->nvtxs > ND_COARSEN_RATIO * (double)g->finer->nvtxs ||
            g->nedges < g->nvtxs / 2)
            break;
    }
    return g;
}

/* ======================================================================== */
/* edge bisection (initial partition of the coarsest graph)                  */
/* ======================================================================== */

static void edge_params(graph *g) {
    const Int n = g->nvtxs, *xadj = g->xadj, *adj = g->adjncy, *ew = g->adjwgt, *where = g->where;
    g->pwgts[0] = g->pwgts[1] = 0;
    for (Int i = 0; i < n; i++) g->pwgts[where[i]] += g->vwgt[i];
    for (Int i = 0; i < n; i++) g->bndptr[i] = -1;
    g->nbnd = 0;
    Int cut = 0;
    for (Int i = 0; i < n; i++) {
        Int in = 0, ex = 0, me = where[i];
        for (Int p = xadj[i]; p < xadj[i + 1]; p++) {
            if (where[adj[p]] == me) in += ew[p];
            else ex += ew[p];
        }
        g->id[i] = in;
        g->ed[i] = ex;
        if (ex > 0 || xadj[i] == xadj[i + 1]) {
            BND_ADD(g, i);
            cut += ex;
        }
    }
    g->mincut = cut / 2;
}

/* balance by moving boundary (or, with no boundary, any) vertices of the
 * heavy side in decreasing gain order */
static void edge_balance(nd_ctx *c, graph *g, const real *tp) {
    const Int n = g->nvtxs;
    /* load imbalance over the ub factor */
    real over = -1.0f;
    for (int s = 0; s < 2; s++) {
        real cur = (real)g->pwgts[s] * c->pijbm[s] - c->ub;
        if (cur > over) over = cur;
    }
    if (over <= 0) return;
    if (iabs64((Int)(tp[0] * (real)g->tvwgt - (real)g->pwgts[0])) < 3 * g->tvwgt / n) return;

    const int general = g->nbnd == 0;
    const Int *xadj = g->xadj, *adj = g->adjncy, *ew = g->adjwgt, *vw = g->vwgt;
    Int *where = g->where, *id = g->id, *ed = g->ed, *pw = g->pwgts;
    Int target[2];
    target[0] = (Int)((real)g->tvwgt * tp[0]);
    target[1] = g->tvwgt - target[0];
    const Int mindiff = iabs64(target[0] - pw[0]);
    const Int from = pw[0] < target[0] ? 1 : 0, to = (from + 1) % 2;
    heap q;
    Int *moved = IALLOC(c, n), *perm = IALLOC(c, n);
    if (c->oom || !heap_init(&q, n)) { c->oom = 1; free(moved); free(perm); heap_free(&q); return; }
    for (Int i = 0; i < n; i++) moved[i] = -1;
    if (!general) {
        Int nb = g->nbnd;
        shuffle(&c->rng, nb, perm, nb / 5, 1);
        for (Int t = 0; t < nb; t++) {
            Int v = g->bndind[perm[t]];
            if (where[v] == from && vw[v] <= mindiff) heap_insert(&q, v, (real)(ed[v] - id[v]));
        }
    } else {
        shuffle(&c->rng, n, perm, n / 5, 1);
        for (Int t = 0; t < n; t++) {
            Int v = perm[t];
            if (where[v] == from && vw[v] <= mindiff) heap_insert(&q, v, (real)(ed[v] - id[v]));
        }
    }
    Int cut = g->mincut;
    for (Int sw = 0; sw < n; sw++) {
        Int v = heap_pop(&q);
        if (v == -1) break;
        if (pw[to] + vw[v] > target[to]) break;
        cut -= ed[v] - id[v];
        pw[to] += vw[v];
        pw[from] -= vw[v];
        where[v] = to;
        moved[v] = sw;
        Int t = id[v]; id[v] = ed[v]; ed[v] = t;
        if (!general) {
            if (ed[v] == 0 && xadj[v] < xadj[v + 1]) BND_DEL(g, v);
        } else {
            if (ed[v] == 0 && g->bndptr[v] != -1 && xadj[v] < xadj[v + 1]) BND_DEL(g, v);
            if (ed[v] > 0 && g->bndptr[v] == -1) BND_ADD(g, v);
        }
        for (Int p = xadj[v]; p < xadj[v + 1]; p++) {
            Int k = adj[p];
            Int kw = (to == where[k]) ? ew[p] : -ew[p];
            id[k] += kw;
            ed[k] -= kw;
            const int eligible = moved[k] == -1 && where[k] == from && vw[k] <= mindiff;
            if (!general) {
                if (g->bndptr[k] != -1) {
                    if (ed[k] == 0) {
                        BND_DEL(g, k);
                        if (eligible) heap_delete(&q, k);
                    } else if (eligible) {
                        heap_update(&q, k, (real)(ed[k] - id[k]));
                    }
                } else if (ed[k] > 0) {
                    BND_ADD(g, k);
                    if (eligible) heap_insert(&q, k, (real)(ed[k] - id[k]));
                }
            } else {
                if (eligible) heap_update(&q, k, (real)(ed[k] - id[k]));
                if (ed[k] == 0 && g->bndptr[k] != -1) BND_DEL(g, k);
                else if (ed[k] > 0 && g->bndptr[k] == -1) BND_ADD(g, k);
            }
        }
    }
    g->mincut = cut;
    heap_free(&q);
    free(moved); free(perm);
}

/* Fiduccia-Mattheyses refinement of an edge cut, `passes` passes with
 * rollback to the best prefix of each pass */
static void edge_fm(nd_ctx *c, graph *g, const real *tp, Int passes) {
    const Int n = g->nvtxs, *xadj = g->xadj, *adj = g->adjncy, *ew = g->adjwgt, *vw = g->vwgt;
    Int *where = g->where, *id = g->id, *ed = g->ed, *pw = g->pwgts;
    Int *moved = IALLOC(c, n), *swaps = IALLOC(c, n), *perm = IALLOC(c, n);
    heap q[2];
    memset(q, 0, sizeof(q));
    if (c->oom || !heap_init(&q[0], n) || !heap_init(&q[1], n)) {
        c->oom = 1;
        free(moved); free(swaps); free(perm); heap_free(&q[0]); heap_free(&q[1]);
        return;
    }
    Int target[2];
    target[0] = (Int)((real)g->tvwgt * tp[0]);
    target[1] = g->tvwgt - target[0];
    double lim = 0.01 * (double)n;
    if (lim < 15) lim = 15;
    if (lim > 100) lim = 100;
    const Int limit = (Int)lim;
    Int avgvw = (pw[0] + pw[1]) / 20, a2 = 2 * (pw[0] + pw[1]) / n;
    if (a2 < avgvw) avgvw = a2;
    const Int origdiff = iabs64(target[0] - pw[0]);
    for (Int i = 0; i < n; i++) moved[i] = -1;

    for (Int pass = 0; pass < passes; pass++) {
        heap_clear(&q[0]);
        heap_clear(&q[1]);
        Int best_at = -1;
        Int cut = g->mincut, best = cut, init = cut;
        Int bestdiff = iabs64(target[0] - pw[0]);
        Int nb = g->nbnd;
        shuffle(&c->rng, nb, perm, nb, 1);
        for (Int t = 0; t < nb; t++) {
            Int v = g->bndind[perm[t]];
            heap_insert(&q[where[v]], v, (real)(ed[v] - id[v]));
        }
        Int sw;
        for (sw = 0; sw < n; sw++) {
            const Int from = (target[0] - pw[0] < target[1] - pw[1]) ? 0 : 1, to = (from + 1) % 2;
            Int v = heap_pop(&q[from]);
            if (v == -1) break;
            cut -= ed[v] - id[v];
            pw[to] += vw[v];
            pw[from] -= vw[v];
            const Int diff = iabs64(target[0] - pw[0]);
            if ((cut < best && diff <= origdiff + avgvw) || (cut == best && diff < bestdiff)) {
                best = cut;
                bestdiff = diff;
                best_at = sw;
            } else if (sw - best_at > limit) {
                cut += ed[v] - id[v];
                pw[from] += vw[v];
                pw[to] -= vw[v];
                break;
            }
            where[v] = to;
            moved[v] = sw;
            swaps[sw] = v;
            Int t = id[v]; id[v] = ed[v]; ed[v] = t;
            if (ed[v] == 0 && xadj[v] < xadj[v + 1]) BND_DEL(g, v);
            for (Int p = xadj[v]; p < xadj[v + 1]; p++) {
                Int k = adj[p];
                Int kw = (to == where[k]) ? ew[p] : -ew[p];
                id[k] += kw;
                ed[k] -= kw;
                if (g->bndptr[k] != -1) {
                    if (ed[k] == 0) {
                        BND_DEL(g, k);
                        if (moved[k] == -1) heap_delete(&q[where[k]], k);
                    } else if (moved[k] == -1) {
                        heap_update(&q[where[k]], k, (real)(ed[k] - id[k]));
                    }
                } else if (ed[k] > 0) {
                    BND_ADD(g, k);
                    if (moved[k] == -1) heap_insert(&q[where[k]], k, (real)(ed[k] - id[k]));
                }
            }
        }
        /* undo the moves after the best prefix */
        for (Int t = 0; t < sw; t++) moved[swaps[t]] = -1;
        for (sw--; sw > best_at; sw--) {
            Int v = swaps[sw];
            Int to = where[v] = (where[v] + 1) % 2;
            Int t = id[v]; id[v] = ed[v]; ed[v] = t;
            if (ed[v] == 0 && g->bndptr[v] != -1 && xadj[v] < xadj[v + 1]) BND_DEL(g, v);
            else if (ed[v] > 0 && g->bndptr[v] == -1) BND_ADD(g, v);
            pw[to] += vw[v];
            pw[(to + 1) % 2] -= vw[v];
            for (Int p = xadj[v]; p < xadj[v + 1]; p++) {
                Int k = adj[p];
                Int kw = (to == where[k]) ? ew[p] : -ew[p];
                id[k] += kw;
                ed[k] -= kw;
                if (g->bndptr[k] != -1 && ed[k] == 0) BND_DEL(g, k);
                if (g->bndptr[k] == -1 && ed[k] > 0) BND_ADD(g, k);
            }
        }
        g->mincut = best;
        if (best_at <= 0 || best == init) break;
    }
    heap_free(&q[0]); heap_free(&q[1]);
    free(moved); free(swaps); free(perm);
}

/* best of `tries` region-growing bisections, each balanced and refined */
static void grow_bisection(nd_ctx *c, graph *g, const real *tp, Int tries) {
    const Int n = g->nvtxs, *xadj = g->xadj, *adj = g->adjncy, *vw = g->vwgt;
    Int *where = g->where;
    Int *best = IALLOC(c, n), *queue = IALLOC(c, n), *seen = IALLOC(c, n);
    if (c->oom) { free(best); free(queue); free(seen); return; }
    const Int maxw1 = (Int)(c->ub * (real)g->tvwgt * tp[1]);
    const Int minw1 = (Int)((1.0 / (double)c->ub) * (double)g->tvwgt * (double)tp[1]);
    Int bestcut = 0;
    for (Int tr = 0; tr < tries; tr++) {
        for (Int i = 0; i < n; i++) { where[i] = 1; seen[i] = 0; }
        Int w0 = 0, w1 = g->tvwgt;
        queue[0] = draw_below(&c->rng, n);
        seen[queue[0]] = 1;
        Int head = 0, tail = 1, left = n - 1;
        int drain = 0;
        for (;;) {
            if (head == tail) {
                /* queue empty: restart from a random unreached vertex */
                if (left == 0 || drain) break;
                Int k = draw_below(&c->rng, left), i;
                for (i = 0; i < n; i++) {
                    if (seen[i]) continue;
                    if (k == 0) break;
                    k--;
                }
                queue[0] = i;
                seen[i] = 1;
                head = 0; tail = 1;
                left--;
            }
            Int i = queue[head++];
            if (w0 > 0 && w1 - vw[i] < minw1) { drain = 1; continue; }
            where[i] = 0;
            w0 += vw[i];
            w1 -= vw[i];
            if (w1 <= maxw1) break;
            drain = 0;
            for (Int p = xadj[i]; p < xadj[i + 1]; p++) {
                Int k = adj[p];
                if (!seen[k]) { queue[tail++] = k; seen[k] = 1; left--; }
            }
        }
        if (w1 == 0) where[draw_below(&c->rng, n)] = 1;
        if (w0 == 0) where[draw_below(&c->rng, n)] = 0;
        edge_params(g);
        edge_balance(c, g, tp);
        edge_fm(c, g, tp, 10);
        if (tr == 0 || bestcut > g->mincut) {
            bestcut = g->mincut;
            memcpy(best, where, sizeof(Int) * (size_t)n);
            if (bestcut == 0) break;
        }
    }
    g->mincut = bestcut;
    memcpy(where, best, sizeof(Int) * (size_t)n);
    free(best); free(queue); free(seen);
}

/* best of `tries` random bisections (graphs without edges) */
static void random_bisection(nd_ctx *c, graph *g, const real *tp, Int tries) {
    const Int n = g->nvtxs, *vw = g->vwgt;
    Int *where = g->where;
    Int *best = IALLOC(c, n), *perm = IALLOC(c, n);
    if (c->oom) { free(best); free(perm); return; }
    const Int maxw0 = (Int)(c->ub * (real)g->tvwgt * tp[0]);
    Int bestcut = 0;
    for (Int tr = 0; tr < tries; tr++) {
        for (Int i = 0; i < n; i++) where[i] = 1;
        if (tr > 0) {
            shuffle(&c->rng, n, perm, n / 2, 1);
            Int w0 = 0;
            for (Int t = 0; t < n; t++) {
                Int i = perm[t];
                if (w0 + vw[i] < maxw0) {
                    where[i] = 0;
                    w0 += vw[i];
                    if (w0 > maxw0) break;
                }
            }
        }
        edge_params(g);
        edge_balance(c, g, tp);
        edge_fm(c, g, tp, 4);
        if (tr == 0 || bestcut > g->mincut) {
            bestcut = g->mincut;
            memcpy(best, where, sizeof(Int) * (size_t)n);
            if (bestcut == 0) break;
        }
    }
    g->mincut = bestcut;
    memcpy(where, best, sizeof(Int) * (size_t)n);
    free(best); free(perm);
}

/* ======================================================================== */
/* vertex separators                                                         */
/* ======================================================================== */

/* side weights, separator list and each separator vertex's neighbour
 * weight on either side */
static void node_params(graph *g) {
    const Int n = g->nvtxs, *xadj = g->xadj, *adj = g->adjncy, *vw = g->vwgt, *where = g->where;
    g->pwgts[0] = g->pwgts[1] = g->pwgts[2] = 0;
    for (Int i = 0; i < n; i++) g->bndptr[i] = -1;
    g->nbnd = 0;
    for (Int i = 0; i < n; i++) {
        Int me = where[i];
        g->pwgts[me] += vw[i];
        if (me != 2) continue;
        BND_ADD(g, i);
        Int *e = g->edeg + 2 * i;
        e[0] = e[1] = 0;
        for (Int p = xadj[i]; p < xadj[i + 1]; p++) {
            Int o = where[adj[p]];
            if (o != 2) e[o] += vw[adj[p]];
        }
    }
    g->mincut = g->pwgts[2];
}

/* Pull vertex k of side `other` into the separator after a neighbour moved
 * to side `to`: recompute its side weights and fix those of its separator
 * neighbours.  `fix(kk, oldgain)` reports each separator neighbour kk. */
#define PULL_INTO_SEPARATOR(g, k, other, FIX)                                  \
    do {                                                                       \
        BND_ADD(g, k);                                                         \
        (g)->where[k] = 2;                                                     \
        (g)->pwgts[other] -= vw[k];                                            \
        Int *ek_ = (g)->edeg + 2 * (k);                                        \
        ek_[0] = ek_[1] = 0;                                                   \
        for (Int pp_ = xadj[k]; pp_ < xadj[(k) + 1]; pp_++) {                  \
            Int kk = adj[pp_];                                                 \
            if ((g)->where[kk] != 2) {                                         \
                ek_[(g)->where[kk]] += vw[kk];                                 \
            } else {                                                           \
                Int oldgain = vw[kk] - (g)->edeg[2 * kk + (other)];            \
                (g)->edeg[2 * kk + (other)] -= vw[k];                          \
                FIX;                                                           \
            }                                                                  \
        }                                                                      \
    } while (0)

/* undo the separator moves after the best prefix (moves recorded in swaps,
 * the vertices each move pulled in recorded in mind[mptr[s]..mptr[s+1])) */
static void node_rollback(graph *g, Int sw, Int best_at, const Int *swaps, const Int *mptr,
                          const Int *mind, int one_sided, Int side_to) {
    const Int *xadj = g->xadj, *adj = g->adjncy, *vw = g->vwgt;
    Int *where = g->where, *pw = g->pwgts;
    for (sw--; sw > best_at; sw--) {
        Int v = swaps[sw];
        Int to = one_sided ? side_to : where[v], other = (to + 1) % 2;
        pw[2] += vw[v];
        pw[to] -= vw[v];
        where[v] = 2;
        BND_ADD(g, v);
        Int *e = g->edeg + 2 * v;
        e[0] = e[1] = 0;
        for (Int p = xadj[v]; p < xadj[v + 1]; p++) {
            Int k = adj[p];
            if (where[k] == 2) g->edeg[2 * k + to] -= vw[v];
            else e[where[k]] += vw[k];
        }
        for (Int t = mptr[sw]; t < mptr[sw + 1]; t++) {
            Int k = mind[t];
            where[k] = other;
            pw[other] += vw[k];
            pw[2] -= vw[k];
            BND_DEL(g, k);
            for (Int p = xadj[k]; p < xadj[k + 1]; p++) {
                Int kk = adj[p];
                if (where[kk] == 2) g->edeg[2 * kk + other] += vw[k];
            }
        }
    }
}

/* FM on the separator, moves to either side (used once, right after the
 * separator is built from an edge cut) */
static void node_fm_two_sided(nd_ctx *c, graph *g, Int passes) {
    const Int n = g->nvtxs, *xadj = g->xadj, *adj = g->adjncy, *vw = g->vwgt;
    Int *where = g->where, *pw = g->pwgts;
    heap q[2];
    memset(q, 0, sizeof(q));
    Int *moved = IALLOC(c, n), *swaps = IALLOC(c, n), *mptr = IALLOC(c, n + 1), *mind = IALLOC(c, 2 * n);
    if (c->oom || !heap_init(&q[0], n) || !heap_init(&q[1], n)) {
        c->oom = 1;
        free(moved); free(swaps); free(mptr); free(mind); heap_free(&q[0]); heap_free(&q[1]);
        return;
    }
    const real mult = (real)(0.5 * (double)c->ub);
    const Int badmax = (Int)(mult * (real)(pw[0] + pw[1] + pw[2]));
    for (Int pass = 0; pass < passes; pass++) {
        for (Int i = 0; i < n; i++) moved[i] = -1;
        heap_clear(&q[0]);
        heap_clear(&q[1]);
        Int best_at = -1, init = g->mincut, best = init;
        Int nb = g->nbnd;
        shuffle(&c->rng, nb, swaps, nb, 1);
        for (Int t = 0; t < nb; t++) {
            Int v = g->bndind[swaps[t]];
            heap_insert(&q[0], v, (real)(vw[v] - g->edeg[2 * v + 1]));
            heap_insert(&q[1], v, (real)(vw[v] - g->edeg[2 * v + 0]));
        }
        const Int limit = c->compress ? (5 * nb < 400 ? 5 * nb : 400) : (2 * nb < 300 ? 2 * nb : 300);
        Int nmind = 0;
        mptr[0] = 0;
        Int bestdiff = iabs64(pw[0] - pw[1]);
        Int to = pw[0] < pw[1] ? 0 : 1;
        Int sw;
        for (sw = 0; sw < n; sw++) {
            Int u0 = heap_top(&q[0]), u1 = heap_top(&q[1]);
            if (u0 != -1 && u1 != -1) {
                Int g0 = vw[u0] - g->edeg[2 * u0 + 1], g1 = vw[u1] - g->edeg[2 * u1 + 0];
                to = g0 > g1 ? 0 : (g0 < g1 ? 1 : pass % 2);
                if (pw[to] + vw[to ? u1 : u0] > badmax) to = (to + 1) % 2;
            } else if (u0 == -1 && u1 == -1) {
                break;
            } else if (u0 != -1 && pw[0] + vw[u0] <= badmax) {
                to = 0;
            } else if (u1 != -1 && pw[1] + vw[u1] <= badmax) {
                to = 1;
            } else {
                break;
            }
            const Int other = (to + 1) % 2;
            Int v = heap_pop(&q[to]);
            if (moved[v] == -1) heap_delete(&q[other], v);
            if (nmind + xadj[v + 1] - xadj[v] >= 2 * n - 1) break;
            pw[2] -= vw[v] - g->edeg[2 * v + other];
            const Int diff = iabs64(pw[to] + vw[v] - (pw[other] - g->edeg[2 * v + other]));
            if (pw[2] < best || (pw[2] == best && diff < bestdiff)) {
                best = pw[2];
                best_at = sw;
                bestdiff = diff;
            } else if (sw - best_at > 2 * limit || (sw - best_at > limit && (double)pw[2] > 1.10 * (double)best)) {
                pw[2] += vw[v] - g->edeg[2 * v + other];
                break;
            }
            BND_DEL(g, v);
            pw[to] += vw[v];
            where[v] = to;
            moved[v] = sw;
            swaps[sw] = v;
            for (Int p = xadj[v]; p < xadj[v + 1]; p++) {
                Int k = adj[p];
                if (where[k] == 2) {
                    Int oldgain = vw[k] - g->edeg[2 * k + to];
                    g->edeg[2 * k + to] += vw[v];
                    if (moved[k] == -1 || moved[k] == -(2 + other))
                        heap_update(&q[other], k, (real)(oldgain - vw[v]));
                } else if (where[k] == other) {
                    mind[nmind++] = k;
                    PULL_INTO_SEPARATOR(g, k, other,
                        if (moved[kk] == -1 || moved[kk] == -(2 + to))
                            heap_update(&q[to], kk, (real)(oldgain + vw[k])));
                    if (moved[k] == -1) {
                        heap_insert(&q[to], k, (real)(vw[k] - g->edeg[2 * k + other]));
                        moved[k] = -(2 + to);
                    }
                }
            }
            mptr[sw + 1] = nmind;
        }
        node_rollback(g, sw, best_at, swaps, mptr, mind, 0, 0);
        g->mincut = best;
        if (best_at == -1 || best >= init) break;
    }
    heap_free(&q[0]); heap_free(&q[1]);
    free(moved); free(swaps); free(mptr); free(mind);
}

/* FM on the separator, each pass moving to one side only, sides alternating */
static void node_fm_one_sided(nd_ctx *c, graph *g, Int passes) {
    const Int n = g->nvtxs, *xadj = g->xadj, *adj = g->adjncy, *vw = g->vwgt;
    Int *where = g->where, *pw = g->pwgts;
    heap q;
    memset(&q, 0, sizeof(q));
    Int *swaps = IALLOC(c, n), *mptr = IALLOC(c, n + 1), *mind = IALLOC(c, 2 * n);
    if (c->oom || !heap_init(&q, n)) {
        c->oom = 1;
        free(swaps); free(mptr); free(mind); heap_free(&q);
        return;
    }
    const real mult = (real)(0.5 * (double)c->ub);
    const Int badmax = (Int)(mult * (real)(pw[0] + pw[1] + pw[2]));
    Int to = pw[0] < pw[1] ? 1 : 0;
    for (Int pass = 0; pass < 2 * passes; pass++) {
        const Int other = to;
        to = (to + 1) % 2;
        heap_clear(&q);
        Int best_at = -1, init = g->mincut, best = init;
        Int nb = g->nbnd;
        shuffle(&c->rng, nb, swaps, nb, 1);
        for (Int t = 0; t < nb; t++) {
            Int v = g->bndind[swaps[t]];
            heap_insert(&q, v, (real)(vw[v] - g->edeg[2 * v + other]));
        }
        const Int limit = c->compress ? (5 * nb < 500 ? 5 * nb : 500) : (3 * nb < 300 ? 3 * nb : 300);
        Int nmind = 0;
        mptr[0] = 0;
        Int bestdiff = iabs64(pw[0] - pw[1]);
        Int sw;
        for (sw = 0; sw < n; sw++) {
            Int v = heap_pop(&q);
            if (v == -1) break;
            if (nmind + xadj[v + 1] - xadj[v] >= 2 * n - 1) break;
            if (pw[to] + vw[v] > badmax) break;
            pw[2] -= vw[v] - g->edeg[2 * v + other];
            const Int diff = iabs64(pw[to] + vw[v] - (pw[other] - g->edeg[2 * v + other]));
            if (pw[2] < best || (pw[2] == best && diff < bestdiff)) {
                best = pw[2];
                best_at = sw;
                bestdiff = diff;
            } else if (sw - best_at > 3 * limit || (sw - best_at > limit && (double)pw[2] > 1.10 * (double)best)) {
                pw[2] += vw[v] - g->edeg[2 * v + other];
                break;
            }
            BND_DEL(g, v);
            pw[to] += vw[v];
            where[v] = to;
            swaps[sw] = v;
            for (Int p = xadj[v]; p < xadj[v + 1]; p++) {
                Int k = adj[p];
                if (where[k] == 2) {
                    g->edeg[2 * k + to] += vw[v];
                } else if (where[k] == other) {
                    mind[nmind++] = k;
                    PULL_INTO_SEPARATOR(g, k, other,
                        (void)oldgain;
                        heap_update(&q, kk, (real)(vw[kk] - g->edeg[2 * kk + other])));
                    heap_insert(&q, k, (real)(vw[k] - g->edeg[2 * k + other]));
                }
            }
            mptr[sw + 1] = nmind;
        }
        node_rollback(g, sw, best_at, swaps, mptr, mind, 1, to);
        g->mincut = best;
        if (pass % 2 == 1 && (best_at == -1 || best >= init)) break;
    }
    heap_free(&q);
    free(swaps); free(mptr); free(mind);
}

/* rebalance the two sides of a projected separator */
static void node_balance(nd_ctx *c, graph *g) {
    const Int n = g->nvtxs, *xadj = g->xadj, *adj = g->adjncy, *vw = g->vwgt;
    Int *where = g->where, *pw = g->pwgts;
    const real mult = (real)(0.5 * (double)c->ub);
    Int badmax = (Int)(mult * (real)(pw[0] + pw[1]));
    if ((pw[0] >= pw[1] ? pw[0] : pw[1]) < badmax) return;
    if (iabs64(pw[0] - pw[1]) < 3 * g->tvwgt / n) return;
    const Int to = pw[0] < pw[1] ? 0 : 1, other = (to + 1) % 2;
    heap q;
    memset(&q, 0, sizeof(q));
    Int *perm = IALLOC(c, n), *moved = IALLOC(c, n);
    if (c->oom || !heap_init(&q, n)) { c->oom = 1; free(perm); free(moved); heap_free(&q); return; }
    for (Int i = 0; i < n; i++) moved[i] = -1;
    Int nb = g->nbnd;
    shuffle(&c->rng, nb, perm, nb, 1);
    for (Int t = 0; t < nb; t++) {
        Int v = g->bndind[perm[t]];
        heap_insert(&q, v, (real)(vw[v] - g->edeg[2 * v + other]));
    }
    for (Int sw = 0; sw < n; sw++) {
        Int v = heap_pop(&q);
        if (v == -1) break;
        moved[v] = 1;
        Int gain = vw[v] - g->edeg[2 * v + other];
        badmax = (Int)(mult * (real)(pw[0] + pw[1]));
        if (pw[to] > pw[other]) break;
        if (gain < 0 && pw[other] < badmax) break;
        if (pw[to] + vw[v] > badmax) continue;
        pw[2] -= gain;
        BND_DEL(g, v);
        pw[to] += vw[v];
        where[v] = to;
        for (Int p = xadj[v]; p < xadj[v + 1]; p++) {
            Int k = adj[p];
            if (where[k] == 2) {
                g->edeg[2 * k + to] += vw[v];
            } else if (where[k] == other) {
                PULL_INTO_SEPARATOR(g, k, other,
                    if (moved[kk] == -1) heap_update(&q, kk, (real)(oldgain + vw[k])));
                heap_insert(&q, k, (real)(vw[k] - g->edeg[2 * k + other]));
            }
        }
    }
    g->mincut = pw[2];
    heap_free(&q);
    free(perm); free(moved);
}

/* separator of the coarsest graph: edge bisection, its boundary, then FM */
static void initial_separator(nd_ctx *c, graph *g, Int tries) {
    static const real tp[2] = {0.5f, 0.5f};
    c->pijbm[0] = g->invtvwgt / tp[0];
    c->pijbm[1] = g->invtvwgt / tp[1];
    if (!graph_partition_mem(c, g)) return;
    if (g->nedges == 0) random_bisection(c, g, tp, tries);
    else grow_bisection(c, g, tp, tries);
    if (c->oom) return;
    edge_params(g);
    /* every boundary vertex with an edge goes to the separator */
    for (Int t = 0; t < g->nbnd; t++) {
        Int v = g->bndind[t];
        if (g->xadj[v + 1] - g->xadj[v] > 0) g->where[v] = 2;
    }
    node_params(g);
    node_fm_two_sided(c, g, 1);
    node_fm_one_sided(c, g, 4);
}

/* project the separator of g->coarser onto g and refine it, level by level
 * up to `top` */
static void refine_separator(nd_ctx *c, graph *top, graph *g) {
    if (g == top) {
        node_params(g);
        return;
    }
    do {
        graph *cg = g;
        g = g->finer;
        if (!graph_partition_mem(c, g)) return;
        for (Int i = 0; i < g->nvtxs; i++) g->where[i] = cg->where[g->cmap[i]];
        graph_free(cg);
        g->coarser = NULL;
        node_params(g);
        node_balance(c, g);
        node_fm_one_sided(c, g, 10);
        if (c->oom) return;
    } while (g != top);
}

static void bisect_small(nd_ctx *c, graph *g, Int tries) {
    Int ct = g->nvtxs / 8;
    if (ct > 100) ct = 100;
    else if (ct < 40) ct = 40;
    c->coarsen_to = ct;
    graph *cg = coarsen(c, g);
    if (!cg) return;
    Int t = cg->nvtxs <= c->coarsen_to ? tries / 2 : tries;
    if (t < 1) t = 1;
    initial_separator(c, cg, t);
    if (c->oom) return;
    refine_separator(c, g, cg);
}

static void bisect(nd_ctx *c, graph *g, Int tries) {
    if (g->nvtxs < 5000) {
        bisect_small(c, g, tries);
        return;
    }
    c->coarsen_to = g->nvtxs / 30 > 100 ? g->nvtxs / 30 : 100;
    graph *cg = coarsen_levels(c, g, 4);
    if (!cg) return;
    Int *best = IALLOC(c, cg->nvtxs);
    if (!best) return;
    const int runs = 5;
    Int mincut = g->tvwgt;
    for (int r = 0; r < runs; r++) {
        bisect_small(c, cg, (Int)(0.7 * (double)tries));
        if (c->oom) { free(best); return; }
        if (r == 0 || cg->mincut < mincut) {
            mincut = cg->mincut;
            if (r < runs - 1) memcpy(best, cg->where, sizeof(Int) * (size_t)cg->nvtxs);
        }
        if (mincut == 0) break;
    }
    if (mincut != cg->mincut) memcpy(cg->where, best, sizeof(Int) * (size_t)cg->nvtxs);
    free(best);
    refine_separator(c, g, cg);
}

static void bisect_best_of(nd_ctx *c, graph *g) {
    if (c->nseps == 1 || g->nvtxs < (c->compress ? 1000 : 2000)) {
        bisect(c, g, ND_INIT_TRIES);
        return;
    }
    Int *best = IALLOC(c, g->nvtxs);
    if (!best) return;
    Int mincut = g->tvwgt;
    for (Int s = 0; s < c->nseps; s++) {
        bisect(c, g, ND_INIT_TRIES);
        if (c->oom) { free(best); return; }
        if (s == 0 || g->mincut < mincut) {
            mincut = g->mincut;
            if (s < c->nseps - 1) memcpy(best, g->where, sizeof(Int) * (size_t)g->nvtxs);
        }
        if (mincut == 0) break;
    }
    if (mincut != g->mincut) {
        memcpy(g->where, best, sizeof(Int) * (size_t)g->nvtxs);
        node_params(g);
    }
    free(best);
}

/* ======================================================================== */
/* multiple minimum degree (leaves of the dissection)                        */
/* ======================================================================== */

/* All arrays 1-based.  Quotient-graph lists end at a 0 entry or their
 * storage end; a negative entry -e continues the list in e's storage.
 * fwd/bwd hold the degree lists (fwd < 0: eliminated, = -position or
 * -(node merged into)); bwd = -deg at a list head, -INT64_MAX if merged or
 * outmatched, 0 if flagged for a degree update. */
typedef struct {
    Int n, *xadj, *adj, *head, *fwd, *bwd, *qsize, *list, *marker, tag;
} mmd_state;

#define MMD_MAX INT64_MAX

static void mmd_eliminate(mmd_state *M, Int me) {
    Int *xadj = M->xadj, *adj = M->adj, *fwd = M->fwd, *bwd = M->bwd, *qsize = M->qsize,
        *list = M->list, *marker = M->marker, *head = M->head;
    const Int tag = M->tag;
    marker[me] = tag;
    /* reachable variables of me: its uneliminated neighbours compacted in
     * place, eliminated ones chained as elements */
    Int elem = 0, rloc = xadj[me], rlmt = xadj[me + 1] - 1;
    for (Int i = xadj[me]; i <= xadj[me + 1] - 1; i++) {
        Int nb = adj[i];
        if (nb == 0) break;
        if (marker[nb] >= tag) continue;
        marker[nb] = tag;
        if (fwd[nb] < 0) { list[nb] = elem; elem = nb; }
        else adj[rloc++] = nb;
    }
    for (; elem > 0; elem = list[elem]) {
        adj[rlmt] = -elem;
        Int link = elem;
    scan_elem:
        for (Int j = xadj[link]; j <= xadj[link + 1] - 1; j++) {
            Int v = adj[j];
            if (v < 0) { link = -v; goto scan_elem; }
            if (v == 0) break;
            if (marker[v] < tag && fwd[v] >= 0) {
                marker[v] = tag;
                while (rloc >= rlmt) {
                    Int l2 = -adj[rlmt];
                    rloc = xadj[l2];
                    rlmt = xadj[l2 + 1] - 1;
                }
                adj[rloc++] = v;
            }
        }
    }
    if (rloc <= rlmt) adj[rloc] = 0;

    /* every reachable variable: out of its degree list, purge its absorbed
     * neighbours, then merge it into me or flag it for an update */
    Int link = me;
scan_reach:
    for (Int i = xadj[link]; i <= xadj[link + 1] - 1; i++) {
        Int r = adj[i];
        if (r < 0) { link = -r; goto scan_reach; }
        if (r == 0) return;
        Int pv = bwd[r];
        if (pv != 0 && pv != -MMD_MAX) {
            Int nx = fwd[r];
            if (nx > 0) bwd[nx] = pv;
            if (pv > 0) fwd[pv] = nx;
            if (pv < 0) head[-pv] = nx;
        }
        Int js = xadj[r], je = xadj[r + 1] - 1, q = js;
        for (Int j = js; j <= je; j++) {
            Int nb = adj[j];
            if (nb == 0) break;
            if (marker[nb] < tag) adj[q++] = nb;
        }
        Int kept = q - js;
        if (kept <= 0) {
            qsize[me] += qsize[r];
            qsize[r] = 0;
            marker[r] = MMD_MAX;
            fwd[r] = -me;
            bwd[r] = -MMD_MAX;
        } else {
            fwd[r] = kept + 1;
            bwd[r] = 0;
            adj[q++] = me;
            if (q <= je) adj[q] = 0;
        }
    }
}

static void mmd_link(mmd_state *M, Int v, Int deg, Int *mdeg) {
    Int f = M->head[deg];
    M->fwd[v] = f;
    M->bwd[v] = -deg;
    if (f > 0) M->bwd[f] = v;
    M->head[deg] = v;
    if (deg < *mdeg) *mdeg = deg;
}

static void mmd_update(mmd_state *M, Int ehead, Int *mdeg) {
    Int *xadj = M->xadj, *adj = M->adj, *fwd = M->fwd, *bwd = M->bwd, *qsize = M->qsize,
        *list = M->list, *marker = M->marker;
    const Int n = M->n, mdeg0 = *mdeg + 1;
    for (Int elem = ehead; elem > 0; elem = list[elem]) {
        Int mtag = M->tag + mdeg0;
        if (mtag >= MMD_MAX) {
            M->tag = 1;
            for (Int i = 1; i <= n; i++)
                if (marker[i] < MMD_MAX) marker[i] = 0;
            mtag = M->tag + mdeg0;
        }
        /* variables of the element needing an update: two-neighbour ones
         * (q2) and the rest (qx); deg0 = size of the element */
        Int q2 = 0, qx = 0, deg0 = 0, link = elem;
    scan_members:
        for (Int i = xadj[link]; i <= xadj[link + 1] - 1; i++) {
            Int v = adj[i];
            if (v < 0) { link = -v; goto scan_members; }
            if (v == 0) break;
            if (qsize[v] == 0) continue;
            deg0 += qsize[v];
            marker[v] = mtag;
            if (bwd[v] != 0) continue;
            if (fwd[v] != 2) { list[v] = qx; qx = v; }
            else { list[v] = q2; q2 = v; }
        }
        /* q2 variables: the element plus one other neighbour */
        for (Int v = q2; v > 0; v = list[v]) {
            if (bwd[v] != 0) continue;
            M->tag++;
            Int deg = deg0;
            Int nb = adj[xadj[v]];
            if (nb == elem) nb = adj[xadj[v] + 1];
            if (fwd[nb] >= 0) {
                deg += qsize[nb];
            } else {
                Int lk = nb;
            scan_other:
                for (Int i = xadj[lk]; i <= xadj[lk + 1] - 1; i++) {
                    Int u = adj[i];
                    if (u == v) continue;
                    if (u < 0) { lk = -u; goto scan_other; }
                    if (u == 0) break;
                    if (qsize[u] == 0) continue;
                    if (marker[u] < M->tag) {
                        marker[u] = M->tag;
                        deg += qsize[u];
                    } else if (bwd[u] == 0) {
                        if (fwd[u] == 2) {
                            /* indistinguishable from v: merge */
                            qsize[v] += qsize[u];
                            qsize[u] = 0;
                            marker[u] = MMD_MAX;
                            fwd[u] = -v;
                            bwd[u] = -MMD_MAX;
                        } else {
                            bwd[u] = -MMD_MAX;      /* outmatched by v */
                        }
                    }
                }
            }
            mmd_link(M, v, deg - qsize[v] + 1, mdeg);
        }
        /* qx variables: count every unmarked neighbour and element member */
        for (Int v = qx; v > 0; v = list[v]) {
            if (bwd[v] != 0) continue;
            M->tag++;
            Int deg = deg0;
            for (Int i = xadj[v]; i <= xadj[v + 1] - 1; i++) {
                Int nb = adj[i];
                if (nb == 0) break;
                if (marker[nb] >= M->tag) continue;
                marker[nb] = M->tag;
                if (fwd[nb] >= 0) {
                    deg += qsize[nb];
                    continue;
                }
                Int lk = nb;
            scan_elem2:
                for (Int j = xadj[lk]; j <= xadj[lk + 1] - 1; j++) {
                    Int u = adj[j];
                    if (u < 0) { lk = -u; goto scan_elem2; }
                    if (u == 0) break;
                    if (marker[u] < M->tag) {
                        marker[u] = M->tag;
                        deg += qsize[u];
                    }
                }
            }
            mmd_link(M, v, deg - qsize[v] + 1, mdeg);
        }
        M->tag = mtag;
    }
}

/* MMD ordering of a graph given 1-based (xadj, adj); returns pos[1..n] = the
 * 1-based position of each vertex.  adj is overwritten. */
static void mmd_order(Int n, Int *xadj, Int *adj, Int *pos, Int *work) {
    if (n <= 0) return;
    mmd_state M;
    M.n = n; M.xadj = xadj; M.adj = adj;
    M.fwd = pos;                              /* fwd doubles as the output */
    M.bwd = work;
    M.head = work + (n + 5);
    M.qsize = work + 2 * (n + 5);
    M.list = work + 3 * (n + 5);
    M.marker = work + 4 * (n + 5);
    Int *fwd = M.fwd, *bwd = M.bwd, *head = M.head, *qsize = M.qsize;
    for (Int v = 1; v <= n; v++) { head[v] = 0; qsize[v] = 1; M.marker[v] = 0; M.list[v] = 0; }
    for (Int v = 1; v <= n; v++) {
        Int d = xadj[v + 1] - xadj[v];
        if (d == 0) d = 1;
        Int f = head[d];
        fwd[v] = f;
        head[d] = v;
        if (f > 0) bwd[f] = v;
        bwd[v] = -d;
    }
    /* list 1 (degree 0 or 1) is numbered first, without elimination */
    Int num = 1;
    for (Int v = head[1]; v > 0;) {
        Int nx = fwd[v];
        M.marker[v] = MMD_MAX;
        fwd[v] = -num;
        num++;
        v = nx;
    }
    if (num <= n) {
        M.tag = 1;
        head[1] = 0;
        Int mdeg = 2;
        for (;;) {
            while (head[mdeg] <= 0) mdeg++;
            const Int mdlmt = mdeg + 1;
            Int ehead = 0;
            int done = 0;
            /* eliminate every variable of degree mdeg..mdlmt (multiple elimination) */
            for (;;) {
                Int v = head[mdeg];
                while (v <= 0) {
                    mdeg++;
                    if (mdeg > mdlmt) break;
                    v = head[mdeg];
                }
                if (v <= 0) break;
                Int nx = fwd[v];
                head[mdeg] = nx;
                if (nx > 0) bwd[nx] = -mdeg;
                fwd[v] = -num;
                if (num + qsize[v] > n) { done = 1; break; }
                M.tag++;
                if (M.tag >= MMD_MAX) {
                    M.tag = 1;
                    for (Int i = 1; i <= n; i++)
                        if (M.marker[i] < MMD_MAX) M.marker[i] = 0;
                }
                mmd_eliminate(&M, v);
                num += qsize[v];
                M.list[v] = ehead;
                ehead = v;
            }
            if (done || num > n) break;
            mmd_update(&M, ehead, &mdeg);
        }
    }
    /* number the merged variables right after their representative */
    Int *rank = bwd;
    for (Int v = 1; v <= n; v++) rank[v] = qsize[v] > 0 ? -fwd[v] : fwd[v];
    for (Int v = 1; v <= n; v++) {
        if (rank[v] > 0) continue;
        Int r = v;
        while (rank[r] <= 0) r = -rank[r];
        Int k = rank[r] + 1;
        fwd[v] = -k;
        rank[r] = k;
        Int f = v, nf = -rank[f];
        while (nf > 0) {
            rank[f] = -r;
            f = nf;
            nf = -rank[f];
        }
    }
    for (Int v = 1; v <= n; v++) pos[v] = -fwd[v];
}

/* order a leaf by MMD into positions last-n .. last-1 */
static int order_leaf(nd_ctx *c, graph *g, Int *order, Int last) {
    const Int n = g->nvtxs;
    Int *xadj = IALLOC(c, n + 2), *adj = IALLOC(c, g->nedges + 1), *pos = IALLOC(c, n + 5),
        *work = IALLOC(c, 5 * (n + 5));
    if (c->oom) { free(xadj); free(adj); free(pos); free(work); return 0; }
    for (Int i = 0; i <= n; i++) xadj[i + 1] = g->xadj[i] + 1;
    for (Int p = 0; p < g->nedges; p++) adj[p + 1] = g->adjncy[p] + 1;
    mmd_order(n, xadj, adj, pos, work);
    const Int first = last - n;
    for (Int i = 0; i < n; i++) order[g->label[i]] = first + pos[i + 1] - 1;
    free(xadj); free(adj); free(pos); free(work);
    return 1;
}

/* ======================================================================== */
/* nested dissection                                                         */
/* ======================================================================== */

/* the two sides of a separated graph as graphs of their own */
static int split_sides(nd_ctx *c, graph *g, graph **side) {
    const Int n = g->nvtxs, *xadj = g->xadj, *adj = g->adjncy, *where = g->where;
    Int *rename = IALLOC(c, n);
    if (!rename) return 0;
    Int nv[3] = {0, 0, 0}, ne[3] = {0, 0, 0};
    for (Int i = 0; i < n; i++) {
        Int k = where[i];
        rename[i] = nv[k]++;
        ne[k] += xadj[i + 1] - xadj[i];
    }
    for (int s = 0; s < 2; s++) {
        graph *h = (graph *)calloc(1, sizeof(graph));
        side[s] = h;
        if (!h) { c->oom = 1; free(rename); return 0; }
        h->xadj = IALLOC(c, nv[s] + 1);
        h->vwgt = IALLOC(c, nv[s]);
        h->label = IALLOC(c, nv[s]);
        h->adjncy = IALLOC(c, ne[s]);
        h->adjwgt = IALLOC(c, ne[s]);
        if (c->oom) { free(rename); return 0; }
        h->xadj[0] = 0;
    }
    Int cnt[2] = {0, 0}, edges[2] = {0, 0};
    for (Int i = 0; i < n; i++) {
        const Int s = where[i];
        if (s == 2) continue;
        graph *h = side[s];
        /* a valid separator leaves no edge between the sides, so the
         * neighbours on the same side are all of them except separator ones */
        for (Int p = xadj[i]; p < xadj[i + 1]; p++)
            if (where[adj[p]] == s) h->adjncy[edges[s]++] = rename[adj[p]];
        h->vwgt[cnt[s]] = g->vwgt[i];
        h->label[cnt[s]] = g->label[i];
        h->xadj[++cnt[s]] = edges[s];
    }
    for (int s = 0; s < 2; s++) {
        graph *h = side[s];
        h->nvtxs = cnt[s];
        h->nedges = edges[s];
        for (Int p = 0; p < edges[s]; p++) h->adjwgt[p] = 1;
        graph_set_total(h);
    }
    free(rename);
    return 1;
}

static void dissect(nd_ctx *c, graph *g, Int *order, Int last) {
    bisect_best_of(c, g);
    if (c->oom) { graph_free(g); return; }
    /* separator vertices take the last positions, in separator-list order */
    for (Int t = 0; t < g->nbnd; t++) order[g->label[g->bndind[t]]] = --last;
    graph *side[2] = {NULL, NULL};
    int ok = split_sides(c, g, side);
    graph_free(g);
    if (!ok) { graph_free(side[0]); graph_free(side[1]); return; }
    Int last_left = last - side[1]->nvtxs;
    if (side[0]->nvtxs > ND_MMD_LEAF && side[0]->nedges > 0) {
        dissect(c, side[0], order, last_left);
    } else {
        order_leaf(c, side[0], order, last_left);
        graph_free(side[0]);
    }
    if (c->oom) { graph_free(side[1]); return; }
    if (side[1]->nvtxs > ND_MMD_LEAF && side[1]->nedges > 0) {
        dissect(c, side[1], order, last);
    } else {
        order_leaf(c, side[1], order, last);
        graph_free(side[1]);
    }
}

/* Merge vertices with identical closed neighbourhoods.  Returns the
 * compressed graph, or NULL when it would keep more than 85% of the
 * vertices; cptr/cind list the members of each compressed vertex. */
static graph *compress_graph(nd_ctx *c, Int n, const Int *xadj, const Int *adj, Int *cptr, Int *cind) {
    Int *mark = IALLOC(c, n), *map = IALLOC(c, n);
    kv *keys = (kv *)nd_alloc(c, n, sizeof(kv));
    if (c->oom) { free(mark); free(map); free(keys); return NULL; }
    for (Int i = 0; i < n; i++) { mark[i] = -1; map[i] = -1; }
    for (Int i = 0; i < n; i++) {
        Int k = 0;
        for (Int p = xadj[i]; p < xadj[i + 1]; p++) k += adj[p];
        keys[i].key = k + i;
        keys[i].val = i;
    }
    kv_sort(keys, n);
    Int cn = 0, l = 0;
    cptr[0] = 0;
    for (Int a = 0; a < n; a++) {
        Int i = keys[a].val;
        if (map[i] != -1) continue;
        mark[i] = a;
        for (Int p = xadj[i]; p < xadj[i + 1]; p++) mark[adj[p]] = a;
        map[i] = cn;
        cind[l++] = i;
        for (Int b = a + 1; b < n; b++) {
            Int j = keys[b].val;
            if (keys[a].key != keys[b].key || xadj[i + 1] - xadj[i] != xadj[j + 1] - xadj[j]) break;
            if (map[j] != -1) continue;
            Int p = xadj[j];
            while (p < xadj[j + 1] && mark[adj[p]] == a) p++;
            if (p == xadj[j + 1]) {
                map[j] = cn;
                cind[l++] = j;
            }
        }
        cptr[++cn] = l;
    }
    graph *g = NULL;
    if ((double)cn < ND_COMPRESS_RATIO * (double)n) {
        g = (graph *)calloc(1, sizeof(graph));
        if (!g) { c->oom = 1; goto out; }
        Int ce = 0;
        for (Int k = 0; k < cn; k++) {
            Int i = cind[cptr[k]];
            ce += xadj[i + 1] - xadj[i];
        }
        g->xadj = IALLOC(c, cn + 1);
        g->vwgt = IALLOC(c, cn);
        g->adjncy = IALLOC(c, ce);
        g->adjwgt = IALLOC(c, ce);
        g->label = IALLOC(c, cn);
        if (c->oom) { graph_free(g); g = NULL; goto out; }
        for (Int i = 0; i < n; i++) mark[i] = -1;
        Int e = 0;
        g->xadj[0] = 0;
        for (Int k = 0; k < cn; k++) {
            mark[k] = k;                       /* no self edge */
            g->vwgt[k] = 0;
            for (Int t = cptr[k]; t < cptr[k + 1]; t++) {
                Int i = cind[t];
                g->vwgt[k] += 1;
                for (Int p = xadj[i]; p < xadj[i + 1]; p++) {
                    Int m = map[adj[p]];
                    if (mark[m] != k) { mark[m] = k; g->adjncy[e++] = m; }
                }
            }
            g->xadj[k + 1] = e;
        }
        for (Int p = 0; p < e; p++) g->adjwgt[p] = 1;
        g->nvtxs = cn;
        g->nedges = e;
        graph_set_total(g);
        for (Int k = 0; k < cn; k++) g->label[k] = k;
    }
out:
    free(mark); free(map); free(keys);
    return g;
}

/* METIS_NodeND with default options on the graph (xadj, adj) (no self
 * edges, symmetric).  perm[k] = vertex in position k. */
static int node_nd(Int n, const Int *xadj, const Int *adj, Int *perm) {
    nd_ctx c;
    memset(&c, 0, sizeof(c));
    mt_seed(&c.rng, 4321);
    c.ub = (real)(1.0 + 0.001 * 200);
    c.ub = (real)((double)c.ub + 0.0000499);
    c.nseps = 1;
    c.compress = 1;
    Int *iperm = IALLOC(&c, n), *cptr = IALLOC(&c, n + 1), *cind = IALLOC(&c, n);
    if (c.oom) { free(iperm); free(cptr); free(cind); return MXC_OUT_OF_MEMORY; }

    graph *g = compress_graph(&c, n, xadj, adj, cptr, cind);
    if (c.oom) { free(iperm); free(cptr); free(cind); return MXC_OUT_OF_MEMORY; }
    Int cn = 0;
    if (g) {
        cn = g->nvtxs;
        real cfactor = (real)(1.0 * (double)n / (double)cn);
        if ((double)cfactor > 1.5 && c.nseps == 1) c.nseps = 2;
    } else {
        c.compress = 0;
        g = (graph *)calloc(1, sizeof(graph));
        if (!g) { free(iperm); free(cptr); free(cind); return MXC_OUT_OF_MEMORY; }
        g->nvtxs = n;
        g->nedges = xadj[n];
        g->xadj = IALLOC(&c, n + 1);
        g->adjncy = IALLOC(&c, xadj[n]);
        g->adjwgt = IALLOC(&c, xadj[n]);
        g->vwgt = IALLOC(&c, n);
        g->label = IALLOC(&c, n);
        if (c.oom) { graph_free(g); free(iperm); free(cptr); free(cind); return MXC_OUT_OF_MEMORY; }
        memcpy(g->xadj, xadj, sizeof(Int) * (size_t)(n + 1));
        memcpy(g->adjncy, adj, sizeof(Int) * (size_t)xadj[n]);
        for (Int p = 0; p < xadj[n]; p++) g->adjwgt[p] = 1;
        for (Int i = 0; i < n; i++) { g->vwgt[i] = 1; g->label[i] = i; }
        graph_set_total(g);
    }
    const Int gn = g->nvtxs;
    dissect(&c, g, iperm, gn);
    if (c.oom) { free(iperm); free(cptr); free(cind); return MXC_OUT_OF_MEMORY; }
    if (c.compress) {
        /* positions of compressed vertices -> positions of their members */
        for (Int i = 0; i < cn; i++) perm[iperm[i]] = i;
        Int l = 0;
        for (Int k = 0; k < cn; k++) {
            Int i = perm[k];
            for (Int t = cptr[i]; t < cptr[i + 1]; t++) iperm[cind[t]] = l++;
        }
    }
    for (Int i = 0; i < n; i++) perm[iperm[i]] = i;
    free(iperm); free(cptr); free(cind);
    return MXC_OK;
}

/* ======================================================================== */
/* CHOLMOD's call                                                            */
/* ======================================================================== */

int mxc_metis_order(int64_t n, const int64_t *Ap, const int64_t *Ai, double nswitch, double dswitch,
                    int64_t *Perm) {
    if (n < 0 || !Perm || (n > 0 && (!Ap || !Ai))) return MXC_INVALID;
    if (n == 0) return MXC_OK;
    /* A + A' without the diagonal, columns laid out as cholmod_l_copy(A,0,-1) */
    Int *xadj = (Int *)calloc((size_t)n + 1, sizeof(Int));
    if (!xadj) return MXC_OUT_OF_MEMORY;
    for (Int j = 0; j < n; j++)
        for (Int p = Ap[j]; p < Ap[j + 1]; p++) {
            Int i = Ai[p];
            if (i < 0 || i >= n) { free(xadj); return MXC_INVALID; }
            if (i > j) { xadj[j + 1]++; xadj[i + 1]++; }
        }
    for (Int j = 0; j < n; j++) xadj[j + 1] += xadj[j];
    const Int nz = xadj[n];
    int identity = nz == 0;
    if (!identity && nswitch > 0) {
        double d = (double)nz / ((double)n * (double)n);
        if (n > (Int)nswitch && d > dswitch) identity = 1;
    }
    int rc = MXC_OK;
    if (identity) {
        for (Int k = 0; k < n; k++) Perm[k] = k;
    } else {
        Int *adj = (Int *)malloc(sizeof(Int) * (size_t)nz), *fill = (Int *)malloc(sizeof(Int) * (size_t)n);
        if (!adj || !fill) { free(adj); free(fill); free(xadj); return MXC_OUT_OF_MEMORY; }
        memcpy(fill, xadj, sizeof(Int) * (size_t)n);
        for (Int j = 0; j < n; j++)
            for (Int p = Ap[j]; p < Ap[j + 1]; p++) {
                Int i = Ai[p];
                if (i > j) {
                    adj[fill[j]++] = i;
                    adj[fill[i]++] = j;
                }
            }
        free(fill);
        rc = node_nd(n, xadj, adj, Perm);
        free(adj);
    }
    free(xadj);
    return rc;
}
