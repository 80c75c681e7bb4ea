/*
 * lines.c -- the algorithmic read count of the device estimator kernels,
 * restated on the CPU.  TEST INFRASTRUCTURE ONLY: the product never links or
 * calls this file; tests/ use it to check the device's own read-counting
 * launch (zg_count_reads), whose count times 128 bytes is the roofline's
 * algorithmic bytes.
 *
 * The draws follow the oracle's estimator streams (Philox keyed by (seed,
 * id), the same random choices as ZGraphInstance::calculate's estimate()
 * bodies, src/ZGraphInstance.cpp:127-441, applications/Triangle.cpp:48-73);
 * each step is charged at the device layout's cost, one 128-byte line per
 * dependent random read:
 *   - the 32-byte record of a sampled edge rank: 1 line (it carries both
 *     endpoints' list starts, degrees and the rank's position in each list,
 *     so the first conditional sample needs no search);
 *   - the sampled entry of a conditional sample: 1 line;
 *   - a search for "first rank >= min_valid" in a rank-ordered list: the
 *     distinct lines its probes touch, consecutive probes in one line
 *     counted once -- interpolation on the known rank bounds while that
 *     halves the range, bisection otherwise, then bisection below 17 entries;
 *   - a closing check of a pair {x, y}: the buckets of the edge-set hash it
 *     visits, 1 + the number of consecutive "passed over" buckets from the
 *     pair's home bucket (128-byte buckets of 8 slots, ceil(m / 4) of them,
 *     murmur3 fmix64 of (min << 32 | max), the high 32 bits scaled to the
 *     bucket count; a bucket is passed over when more keys arrive at it --
 *     its own plus those overflowing from the previous bucket -- than it has
 *     slots, which does not depend on the insertion order).
 * Plans covered: triangle, 3-chain (no entry read on a graph without
 * parallel edges: the entry can only set the parallel-copy denominator),
 * 3-motif, the three-shape 4-clique plan and (og_device_lines_house) the
 * split 5-house plan.
 */
#include "zg_oracle.h"

#include <stdlib.h>
#include <string.h>

enum { LM_SLOTS = 8, LM_EDGES_PER_BUCKET = 4 };

typedef struct {
    uint32_t nb;
    uint8_t *passed;
    int simple;   /* no vertex pair joined twice */
} lm_hash;

/* murmur3 fmix64 (Appleby, MurmurHash3) */
static uint64_t fmix64(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

static uint64_t pkey(uint32_t x, uint32_t y)
{
    return x < y ? ((uint64_t)x << 32) | y : ((uint64_t)y << 32) | x;
}

static uint32_t home_bucket(uint64_t key, uint32_t nb)
{
    return (uint32_t)(((fmix64(key) >> 32) * (uint64_t)nb) >> 32);
}

static int u64cmp(const void *a, const void *b)
{
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

static int hash_build(const og_graph *g, lm_hash *h)
{
    const uint64_t m = g->m;
    h->nb = (uint32_t)((m + LM_EDGES_PER_BUCKET - 1) / LM_EDGES_PER_BUCKET);
    h->passed = calloc(h->nb ? h->nb : 1, 1);
    uint32_t *arrive = calloc(h->nb ? h->nb : 1, sizeof(uint32_t));
    uint64_t *keys = malloc(sizeof(uint64_t) * (m ? m : 1));
    if (!h->passed || !arrive || !keys) { free(arrive); free(keys); return -1; }
    for (uint64_t r = 0; r < m; ++r) {
        keys[r] = pkey(g->src[r], g->dst[r]);
        arrive[home_bucket(keys[r], h->nb)]++;
    }
    qsort(keys, m, sizeof(uint64_t), u64cmp);
    h->simple = 1;
    for (uint64_t r = 1; r < m; ++r)
        if (keys[r] == keys[r - 1]) h->simple = 0;
    free(keys);
    /* overflow flows to the next bucket (cyclically); iterate to the fixed point */
    uint64_t carry_in = 0;
    for (int pass = 0; pass < 64 && h->nb; ++pass) {
        uint64_t carry = carry_in;
        memset(h->passed, 0, h->nb);
        for (uint32_t b = 0; b < h->nb; ++b) {
            const uint64_t a = arrive[b] + carry;
            if (a > LM_SLOTS) { h->passed[b] = 1; carry = a - LM_SLOTS; }
            else carry = 0;
        }
        if (carry == carry_in) break;
        carry_in = carry;
    }
    free(arrive);
    return 0;
}

/* buckets a closing check of {x, y} reads */
static uint64_t walk(const lm_hash *h, uint32_t x, uint32_t y)
{
    if (x == y) return 0;
    uint32_t b = home_bucket(pkey(x, y), h->nb);
    uint64_t n = 1;
    while (h->passed[b]) { b = (b + 1 == h->nb) ? 0 : b + 1; ++n; }
    return n;
}

/* some copy of {x, y} with rank >= mv; *rank = the smallest such rank */
static int pair_rank(const og_graph *g, uint32_t x, uint32_t y, uint64_t mv, uint64_t *rank)
{
    if (x == y) return 0;
    uint32_t v = x, u = y;
    if (g->off[y + 1] - g->off[y] < g->off[x + 1] - g->off[x]) { v = y; u = x; }
    uint64_t best = UINT64_MAX;
    for (uint64_t i = g->off[v]; i < g->off[v + 1]; ++i)
        if (g->adj_dst[i] == u && g->adj_rank[i] >= mv && g->adj_rank[i] < best) best = g->adj_rank[i];
    if (best == UINT64_MAX) return 0;
    if (rank) *rank = best;
    return 1;
}

typedef struct {
    const og_graph *g;
    const lm_hash *h;
    uint64_t lines;
} lm;

static int close_pair(lm *L, uint32_t x, uint32_t y, uint64_t mv, uint64_t *rank)
{
    L->lines += walk(L->h, x, y);
    return pair_rank(L->g, x, y, mv, rank);
}

typedef struct { uint32_t v, deg, off; } lm_vi;

static lm_vi vi_of(const og_graph *g, uint32_t v)
{
    lm_vi x = {v, (uint32_t)(g->off[v + 1] - g->off[v]), (uint32_t)g->off[v]};
    return x;
}

/* position of rank r in v's rank-ordered list */
static uint32_t pos_of(const og_graph *g, uint32_t v, uint64_t r)
{
    uint64_t lo = g->off[v], hi = g->off[v + 1];
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (g->adj_rank[mid] < r) lo = mid + 1; else hi = mid;
    }
    return (uint32_t)(lo - g->off[v]);
}

/* the device's first-rank->=min_valid search in [lo, hi) of x's list, with
 * known rank bounds [kl, ku); adds the distinct lines its probes touch */
static uint32_t lb_rank(lm *L, lm_vi x, uint32_t lo, uint32_t hi, uint32_t min_valid, uint64_t kl, uint64_t ku)
{
    const uint64_t *a = L->g->adj_rank + x.off;
    uint64_t probes = 0, line = UINT64_MAX;
    int bisect = 0;
    while (hi - lo > 16) {
        const uint32_t n = hi - lo;
        uint32_t p;
        if (bisect || ku <= kl) {
            p = lo + (n >> 1);
        } else {
            const uint64_t t = min_valid > kl ? (uint64_t)min_valid - kl : 0;
            p = lo + (uint32_t)(((uint64_t)n * t) / (ku - kl));
            if (p >= hi) p = hi - 1;
        }
        const uint32_t r = (uint32_t)a[p];
        if ((((uint64_t)x.off + p) >> 4) != line) { line = ((uint64_t)x.off + p) >> 4; ++probes; }
        if (r < min_valid) { lo = p + 1; kl = (uint64_t)r + 1; }
        else { hi = p; ku = r; }
        bisect = (hi - lo) > (n >> 1);
    }
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < min_valid) lo = mid + 1; else hi = mid;
        if ((((uint64_t)x.off + mid) >> 4) != line) { line = ((uint64_t)x.off + mid) >> 4; ++probes; }
    }
    L->lines += probes;
    return lo;
}

typedef struct { uint32_t rank, src, dst, pos; } lm_se;

/* the r-th valid entry over lists V[0..nv) from pos[]; 0 if none (no draw) */
static int cond_sample(lm *L, og_stream *st, const lm_vi *V, const uint32_t *pos, int nv, uint64_t *s_out,
                       lm_se *e)
{
    uint64_t sum = 0;
    for (int i = 0; i < nv; ++i) sum += V[i].deg - pos[i];
    *s_out = sum;
    if (sum == 0) return 0;
    uint64_t r = og_uniform(st, sum);
    for (int i = 0; i < nv; ++i) {
        const uint64_t cnt = V[i].deg - pos[i];
        if (r < cnt) {
            const uint64_t idx = (uint64_t)V[i].off + pos[i] + r;
            e->rank = (uint32_t)L->g->adj_rank[idx];
            e->dst = L->g->adj_dst[idx];
            e->src = V[i].v;
            e->pos = pos[i] + (uint32_t)r;
            L->lines += 1;
            return 1;
        }
        r -= cnt;
    }
    abort();
}

/* mode 0 triangle, 1 3-chain, 2 3-motif */
static void lines_wedge(lm *L, og_stream *st, int mode)
{
    const og_graph *g = L->g;
    const uint32_t r0 = (uint32_t)og_uniform(st, g->m);
    L->lines += 1;   /* e0's record */
    const uint32_t a = g->src[r0], b = g->dst[r0];
    const lm_vi A = vi_of(g, a), B = vi_of(g, b);
    const uint32_t pa = pos_of(g, a, r0) + 1, pb = pos_of(g, b, r0) + 1;
    const uint64_t ca = A.deg - pa, cb = B.deg - pb, s = ca + cb;
    if (s == 0) return;
    const uint64_t r = og_uniform(st, s);
    if (mode == 1 && L->h->simple) return;
    L->lines += 1;   /* the sampled entry */
    if (mode == 1) return;
    const int inA = r < ca;
    const uint64_t idx = inA ? (uint64_t)A.off + pa + r : (uint64_t)B.off + pb + (r - ca);
    const uint32_t r1 = (uint32_t)g->adj_rank[idx], w = g->adj_dst[idx], own = inA ? a : b;
    uint32_t tx, ty;
    if (a == own) { tx = b; ty = w; }
    else if (a == w) { tx = b; ty = own; }
    else if (b == own) { tx = a; ty = w; }
    else { tx = a; ty = own; }
    const int is_e0 = (tx == a && ty == b) || (tx == b && ty == a);
    const int is_e1 = (tx == own && ty == w) || (tx == w && ty == own);
    if (!is_e0 && !is_e1 && tx != ty) close_pair(L, tx, ty, (uint64_t)r1 + 1, NULL);
}

static void lines_k4(lm *L, og_stream *st)
{
    const og_graph *g = L->g;
    const uint64_t m = g->m;
    const uint32_t r0 = (uint32_t)og_uniform(st, m);
    L->lines += 1;
    const uint32_t a = g->src[r0], b = g->dst[r0];
    const lm_vi A = vi_of(g, a), B = vi_of(g, b);
    const uint32_t pa0 = pos_of(g, a, r0) + 1, pb0 = pos_of(g, b, r0) + 1;
    uint64_t s;
    lm_se e1;
    {
        const lm_vi V[2] = {A, B};
        const uint32_t pos[2] = {pa0, pb0};
        if (!cond_sample(L, st, V, pos, 2, &s, &e1)) return;
    }
    const uint32_t w = e1.dst, r1 = e1.rank;
    if (w == a || w == b) return;
    const int oa = e1.src == a;
    const uint32_t ob = oa ? b : a;
    uint64_t rt;
    if (!close_pair(L, ob, w, (uint64_t)r1 + 1, &rt)) return;
    const lm_vi O = oa ? A : B, OB = oa ? B : A;
    const uint32_t po1 = e1.pos;
    L->lines += 1;   /* r1's record: w's lists and r1's position in them */
    const lm_vi W = vi_of(g, w);
    const uint32_t pw1 = pos_of(g, w, r1);
    const uint32_t pob0 = oa ? pb0 : pa0;
    /* shape A */
    {
        const uint32_t pob = lb_rank(L, OB, pob0, OB.deg, r1 + 1, (uint64_t)r0 + 1, m);
        const lm_vi V[3] = {A, B, W};
        const uint32_t pos[3] = {oa ? po1 + 1 : pob, oa ? pob : po1 + 1, pw1 + 1};
        lm_se e2;
        if (cond_sample(L, st, V, pos, 3, &s, &e2)) {
            const uint32_t d = e2.dst;
            if (d != a && d != b && d != w) {
                const uint64_t mv = (uint64_t)e2.rank + 1;
                int ok = close_pair(L, ob, w, mv, NULL);
                if (ok && e2.src != a) ok = close_pair(L, a, d, mv, NULL);
                if (ok && e2.src != b) ok = close_pair(L, b, d, mv, NULL);
                if (ok && e2.src != w) ok = close_pair(L, w, d, mv, NULL);
            }
        }
    }
    /* shape B */
    {
        L->lines += 1;   /* rt's record */
        const uint32_t pob_t = pos_of(g, ob, rt) + 1, pw_t = pos_of(g, w, rt) + 1;
        const uint32_t po = lb_rank(L, O, po1 + 1, O.deg, (uint32_t)rt + 1, (uint64_t)r1 + 1, m);
        const lm_vi V[3] = {A, B, W};
        const uint32_t pos[3] = {oa ? po : pob_t, oa ? pob_t : po, pw_t};
        lm_se e3;
        if (cond_sample(L, st, V, pos, 3, &s, &e3)) {
            const uint32_t d = e3.dst;
            if (d != a && d != b && d != w) {
                const uint64_t mv = (uint64_t)e3.rank + 1;
                int ok = 1;
                if (e3.src != a) ok = close_pair(L, a, d, mv, NULL);
                if (ok && e3.src != b) ok = close_pair(L, b, d, mv, NULL);
                if (ok && e3.src != w) ok = close_pair(L, w, d, mv, NULL);
            }
        }
    }
    /* shape C */
    {
        const uint32_t pw0 = lb_rank(L, W, 0, pw1, r0 + 1, 0, r1);
        const lm_vi V[3] = {A, B, W};
        const uint32_t pos[3] = {pa0, pb0, pw0};
        lm_se e4;
        if (cond_sample(L, st, V, pos, 3, &s, &e4)) {
            const uint32_t d = e4.dst;
            if (e4.src == w && d != a && d != b && d != w && e4.rank < r1) {
                const uint64_t mv = (uint64_t)r1 + 1;
                if (close_pair(L, ob, w, mv, NULL) && close_pair(L, a, d, mv, NULL)) close_pair(L, b, d, mv, NULL);
            }
        }
    }
}

/* split 5-house (est_house_split's draws): e0's record, the roof's entry and
 * its pair's buckets; per completion of a closed roof, two entries, and the
 * pair's buckets only when x and y pass the distinctness checks */
static void lines_house(lm *L, og_stream *st, int J, uint64_t seed, uint64_t id)
{
    const og_graph *g = L->g;
    const uint64_t r0 = og_uniform(st, g->m);
    L->lines += 1;   /* e0's record */
    uint32_t a = g->src[r0], b = g->dst[r0];
    uint64_t da = g->off[a + 1] - g->off[a], db = g->off[b + 1] - g->off[b];
    if (db < da) {
        uint32_t t = a; a = b; b = t;
        uint64_t u = da; da = db; db = u;
    }
    const uint32_t w = g->adj_dst[g->off[a] + og_uniform(st, da)];
    L->lines += 1;   /* the roof's entry */
    if (w == b) return;
    L->lines += walk(L->h, w, b);
    if (!pair_rank(g, w, b, 0, NULL)) return;
    for (int j = 0; j < J; ++j) {
        og_stream c;
        og_stream_init(&c, seed, OG_TAG_COMPLETION + (uint32_t)j, id);
        const uint32_t x = g->adj_dst[g->off[a] + og_uniform(&c, da)];
        const uint32_t y = g->adj_dst[g->off[b] + og_uniform(&c, db)];
        L->lines += 2;
        if (x == b || y == a || x == y || x == w || y == w) continue;
        L->lines += walk(L->h, x, y);
    }
}

uint64_t og_device_lines_house(const og_graph *g, int completions, uint64_t seed, uint64_t first, uint64_t count)
{
    if (g->m == 0) return 0;
    lm_hash h;
    if (hash_build(g, &h)) return UINT64_MAX;
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : total)
    for (uint64_t i = 0; i < count; ++i) {
        lm L = {g, &h, 0};
        og_stream st;
        og_stream_init(&st, seed, OG_TAG_ESTIMATOR, first + i);
        lines_house(&L, &st, completions, seed, first + i);
        total += L.lines;
    }
    free(h.passed);
    return total;
}

uint64_t og_device_lines(const og_graph *g, int kind, uint64_t seed, uint64_t first, uint64_t count)
{
    if (g->m == 0) return 0;
    if (kind != OG_PLAN_TRIANGLE && kind != OG_PLAN_THREE_CHAIN && kind != OG_PLAN_THREE_MOTIF &&
        kind != OG_PLAN_K4_CLOSE)
        return UINT64_MAX;
    lm_hash h;
    if (hash_build(g, &h)) return UINT64_MAX;
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : total)
    for (uint64_t i = 0; i < count; ++i) {
        lm L = {g, &h, 0};
        og_stream st;
        og_stream_init(&st, seed, OG_TAG_ESTIMATOR, first + i);
        if (kind == OG_PLAN_K4_CLOSE) lines_k4(&L, &st);
        else lines_wedge(&L, &st, kind == OG_PLAN_TRIANGLE ? 0 : kind == OG_PLAN_THREE_CHAIN ? 1 : 2);
        total += L.lines;
    }
    free(h.passed);
    return total;
}
