/*
 * oracle.c -- CPU restatement of the ASAP estimator path of ZGraph
 * (jy-yuan/sampling-graph-mining).  TEST INFRASTRUCTURE ONLY: the product
 * never links or calls this file; tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it as the checker.
 *
 * Parity status (what pins each function; see tests/):
 *   og_philox4x32_10   Random123 known-answer vectors + rocRAND's Philox
 *                      (tests/golden/philox_vectors.json)            pinned
 *   og_uniform         exhaustive small-range rejection checks + chi^2   pinned
 *   og_graph_build     CSR invariants (rank-ascending lists, sum deg = 2m,
 *                      every rank twice) on the reference fixtures       pinned
 *   core cond. sample  brute-force enumeration of every r on tiny graphs pinned
 *   og_close           brute-force edge-set membership on tiny graphs,
 *                      parallel edges ("some copy above min_valid")     pinned
 *   estimators         exhaustive expectation E[X|graph,ranks] == exact
 *                      count (closed forms / brute force)              pinned
 *   est_k4_split       exhaustive expectation with 1, 2 and 4 completions
 *                      == exact 4-cliques; first stage == est_k4_close's
 *                      per id; variance falls with completions          pinned
 *   est_house_split    exhaustive expectation with 1 and 2 completions ==
 *                      brute-force house counts (K5: 60, C(n,5) 60 on K_n;
 *                      copy-weighted on multigraphs); the reference's
 *                      wiki-vote ground truth inside the CLT interval   pinned
 *   pattern dedupe     tests/TestPattern2.cpp values (12/24, 1/24, 5/120,
 *                      60/120 house) and TestPattern.cpp values           pinned
 *   exact counters     closed forms, brute force, SNAP's published
 *                      wiki-Vote triangle count, naive_implementation
 *                      binaries built from source into oracle/_ref       pinned
 *   og_exact_c4        brute force of the 4-cycle pattern; K_n, C_n      pinned
 *   og_exact_paths4    og_exact_paths_brute(5), itself equal to the
 *                      reference's FiveChain built from source            pinned
 *   og_exact_house     og_exact_house_brute on random graphs and
 *                      multigraphs, K_n closed forms, FiveHouse.cpp:68's
 *                      wiki-vote count                                   pinned
 *   og_regression_k    tests/TestRegression.cpp known answer k^2=609.8565 pinned
 *   og_elp_*           hand-computed closed form on 1-3 points           pinned
 *   og_device_lines    (lines.c) closed costs on CPU (3-chain = 1 line on a
 *                      simple graph, 3-motif = triangle); equals the device's
 *                      own read count on eleven graphs (GPU tests)         pinned
 *   og_device_lines_house  equals the device's read count of the split
 *                      house plan on the parity graphs (GPU tests)       pinned
 *   generators         restated from the device generators' definition
 *                      (not the reference's); the device output is
 *                      compared element by element                     pinned
 */
#include "zg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

_Static_assert(sizeof(og_outcome) == 24, "og_outcome must match the 24-byte zg_outcome record");
_Static_assert(sizeof(og_sums) == 56, "og_sums layout is mirrored by oracle/pyoracle.py");

/* ===================================================================== */
/* Philox4x32-10 (Salmon et al., SC'11), replacing include/Random.hpp     */
/* ===================================================================== */
void og_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* enumeration hook used by og_expectation (NULL in normal runs) */
typedef struct og_enum og_enum;
static __thread og_enum *tl_enum = NULL;
static uint64_t enum_draw(og_enum *e, uint64_t n);

void og_stream_init(og_stream *s, uint64_t seed, uint32_t tag, uint64_t id)
{
    s->key[0] = (uint32_t)seed;
    s->key[1] = (uint32_t)(seed >> 32);
    s->tag = tag;
    s->id = id;
    s->word = 0;
    s->buf_block = UINT64_MAX;
}

/* MT19937 (Matsumoto and Nishimura, ACM TOMACS 1998), the engine behind the
 * reference's Random (include/Random.hpp:6-16: std::mt19937 per thread).  Only
 * for the cost-faithful CPU baseline: its draws are not the Philox streams, so
 * runs in this mode are timed, never compared. */
typedef struct { uint32_t mt[624]; int idx; } og_mt;
static int og_rng_mode = 0;            /* 0 Philox streams, 1 per-thread MT19937 */
static __thread og_mt tl_mt;
static __thread int tl_mt_seeded = 0;

static void mt_seed(og_mt *s, uint32_t seed)
{
    s->mt[0] = seed;
    for (int i = 1; i < 624; ++i) s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
    s->idx = 624;
}

static uint32_t mt_next(og_mt *s)
{
    if (s->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            const uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % 624] & 0x7fffffffu);
            s->mt[i] = s->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908B0DFu : 0u);
        }
        s->idx = 0;
    }
    uint32_t y = s->mt[s->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9D2C5680u;
    y ^= (y << 15) & 0xEFC60000u;
    y ^= y >> 18;
    return y;
}

void og_set_rng_mode(int mode) { og_rng_mode = mode; }

uint32_t og_mt19937_nth(uint32_t seed, uint32_t n)
{
    og_mt s;
    mt_seed(&s, seed);
    uint32_t y = 0;
    for (uint32_t i = 0; i < n; ++i) y = mt_next(&s);
    return y;
}

uint32_t og_next_u32(og_stream *s)
{
    if (og_rng_mode) {
        if (!tl_mt_seeded) {
            mt_seed(&tl_mt, 5489u + (uint32_t)omp_get_thread_num());
            tl_mt_seeded = 1;
        }
        return mt_next(&tl_mt);
    }
    uint64_t blk = s->word >> 2;
    if (blk != s->buf_block) {
        uint32_t ctr[4] = {(uint32_t)blk, s->tag, (uint32_t)s->id, (uint32_t)(s->id >> 32)};
        og_philox4x32_10(ctr, s->key, s->buf);
        s->buf_block = blk;
    }
    return s->buf[s->word++ & 3];
}

/* Lemire, "Fast random integer generation in an interval" (TOMACS 2019). */
uint64_t og_uniform(og_stream *s, uint64_t n)
{
    if (tl_enum) return enum_draw(tl_enum, n);
    if (n <= (1ull << 32)) {
        uint64_t x = og_next_u32(s);
        uint64_t mlt = x * n;
        uint32_t l = (uint32_t)mlt;
        if ((uint64_t)l < n) {
            uint64_t t = ((1ull << 32) - n) % n;
            while ((uint64_t)l < t) {
                x = og_next_u32(s);
                mlt = x * n;
                l = (uint32_t)mlt;
            }
        }
        return mlt >> 32;
    } else {
        uint64_t x = og_next_u32(s);
        x |= (uint64_t)og_next_u32(s) << 32;
        og_u128 mlt = (og_u128)x * n;
        uint64_t l = (uint64_t)mlt;
        if (l < n) {
            uint64_t t = (0 - n) % n;
            while (l < t) {
                x = og_next_u32(s);
                x |= (uint64_t)og_next_u32(s) << 32;
                mlt = (og_u128)x * n;
                l = (uint64_t)mlt;
            }
        }
        return (uint64_t)(mlt >> 64);
    }
}

uint64_t og_key64(uint64_t seed, uint32_t tag, uint64_t id)
{
    uint32_t ctr[4] = {0, tag, (uint32_t)id, (uint32_t)(id >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    og_philox4x32_10(ctr, key, o);
    return (uint64_t)o[0] | ((uint64_t)o[1] << 32);
}

/* ===================================================================== */
/* Graph: CSRConstructor::construct_bitmap_assisted_csr at W = 1          */
/* (src/Graph.cpp:214-377)                                                */
/* ===================================================================== */
typedef struct { uint64_t key; uint64_t idx; } kv_t;

static int kv_cmp(const void *a, const void *b)
{
    const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

void og_graph_free(og_graph *g)
{
    free(g->src); free(g->dst); free(g->w); free(g->off);
    free(g->adj_rank); free(g->adj_src); free(g->adj_dst);
    memset(g, 0, sizeof(*g));
}

int og_graph_build(og_graph *g, uint32_t n, uint64_t m, const uint32_t *src,
                   const uint32_t *dst, const float *w, uint64_t shuffle_seed)
{
    memset(g, 0, sizeof(*g));
    for (uint64_t i = 0; i < m; ++i) {
        if (src[i] >= n || dst[i] >= n) return -2;
        if (src[i] == dst[i]) return -1;         /* src/Graph.cpp:454 */
    }
    g->n = n; g->m = m; g->num_edges = m;
    /* rank := position after a keyed random permutation; the reference's
     * Fisher-Yates gives rank := post-shuffle index (src/Graph.cpp:218-231,296) */
    kv_t *kv = (kv_t *)malloc(sizeof(kv_t) * (m ? m : 1));
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)m; ++i) {
        kv[i].key = og_key64(shuffle_seed, OG_TAG_SHUFFLE, (uint64_t)i);
        kv[i].idx = (uint64_t)i;
    }
    qsort(kv, m, sizeof(kv_t), kv_cmp);
    g->src = (uint32_t *)malloc(4 * (m ? m : 1));
    g->dst = (uint32_t *)malloc(4 * (m ? m : 1));
    if (w) g->w = (float *)malloc(4 * (m ? m : 1));
    for (uint64_t r = 0; r < m; ++r) {
        g->src[r] = src[kv[r].idx];
        g->dst[r] = dst[kv[r].idx];
        if (w) g->w[r] = w[kv[r].idx];
    }
    free(kv);
    /* index: degree count then prefix (src/Graph.cpp:255-272) */
    g->off = (uint64_t *)calloc((size_t)n + 1, 8);
    for (uint64_t r = 0; r < m; ++r) { g->off[g->src[r] + 1]++; g->off[g->dst[r] + 1]++; }
    uint64_t maxdeg = 0;
    for (uint32_t v = 0; v < n; ++v) {
        if (g->off[v + 1] > maxdeg) maxdeg = g->off[v + 1];
        g->off[v + 1] += g->off[v];
    }
    g->max_degree = maxdeg;                       /* src/Graph.cpp:466-474 */
    /* fill in rank order so that each list ascends by rank (:286-311) */
    g->adj_rank = (uint64_t *)malloc(8 * (m ? 2 * m : 1));
    g->adj_src = (uint32_t *)malloc(4 * (m ? 2 * m : 1));
    g->adj_dst = (uint32_t *)malloc(4 * (m ? 2 * m : 1));
    uint64_t *cur = (uint64_t *)malloc(8 * ((size_t)n + 1));
    memcpy(cur, g->off, 8 * ((size_t)n + 1));
    for (uint64_t r = 0; r < m; ++r) {
        uint32_t s = g->src[r], d = g->dst[r];
        uint64_t p = cur[s]++;
        g->adj_rank[p] = r; g->adj_src[p] = s; g->adj_dst[p] = d;
        p = cur[d]++;
        g->adj_rank[p] = r; g->adj_src[p] = d; g->adj_dst[p] = s;
    }
    free(cur);
    return 0;
}

int og_graph_from_csr(og_graph *g, uint32_t n, uint64_t m, const uint32_t *src, const uint32_t *dst,
                      const uint64_t *off, const uint32_t *adj_rank, const uint32_t *adj_nbr)
{
    memset(g, 0, sizeof(*g));
    g->n = n; g->m = m; g->num_edges = m;
    g->src = (uint32_t *)malloc(4 * (m ? m : 1));
    g->dst = (uint32_t *)malloc(4 * (m ? m : 1));
    g->off = (uint64_t *)malloc(8 * ((size_t)n + 1));
    g->adj_rank = (uint64_t *)malloc(8 * (m ? 2 * m : 1));
    g->adj_src = (uint32_t *)malloc(4 * (m ? 2 * m : 1));
    g->adj_dst = (uint32_t *)malloc(4 * (m ? 2 * m : 1));
    memcpy(g->src, src, 4 * m);
    memcpy(g->dst, dst, 4 * m);
    memcpy(g->off, off, 8 * ((size_t)n + 1));
    uint64_t maxdeg = 0;
    #pragma omp parallel for schedule(dynamic, 4096) reduction(max:maxdeg)
    for (int64_t v = 0; v < (int64_t)n; ++v) {
        uint64_t d = off[v + 1] - off[v];
        if (d > maxdeg) maxdeg = d;
        for (uint64_t x = off[v]; x < off[v + 1]; ++x) {
            g->adj_rank[x] = adj_rank[x];
            g->adj_src[x] = (uint32_t)v;
            g->adj_dst[x] = adj_nbr[x];
        }
    }
    g->max_degree = maxdeg;
    return 0;
}

/* keep-list rebuild shared by the sampler and the filter */
static int rebuild_from_kept(const og_graph *in, og_graph *out, const uint8_t *keep, uint64_t seed)
{
    uint64_t k = 0;
    for (uint64_t r = 0; r < in->m; ++r) k += keep[r];
    uint32_t *s = (uint32_t *)malloc(4 * (k ? k : 1));
    uint32_t *d = (uint32_t *)malloc(4 * (k ? k : 1));
    float *w = in->w ? (float *)malloc(4 * (k ? k : 1)) : NULL;
    uint64_t j = 0;
    for (uint64_t r = 0; r < in->m; ++r) if (keep[r]) {
        s[j] = in->src[r]; d[j] = in->dst[r]; if (w) w[j] = in->w[r]; ++j;
    }
    int rc = og_graph_build(out, in->n, k, s, d, w, seed);
    free(s); free(d); free(w);
    return rc;
}

/* GraphSampler::sample_graph (src/Graph.cpp:628-780): keep edge iff
 * rand_int(1, 2^30) <= rate * 2^30 (:652-667); at W = 1 num_edges and the
 * max degree are the sampled graph's own (:704-749, factor 1). */
int og_graph_subsample(const og_graph *in, og_graph *out, double rate, uint64_t seed)
{
    const uint64_t max_rand = 1ull << 30;
    uint64_t thr = (uint64_t)(rate * (double)max_rand);
    uint8_t *keep = (uint8_t *)malloc(in->m ? in->m : 1);
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < (int64_t)in->m; ++r) {
        og_stream st;
        og_stream_init(&st, seed, OG_TAG_SUBSAMPLE, (uint64_t)r);
        keep[r] = (1 + og_uniform(&st, max_rand)) <= thr;
    }
    int rc = rebuild_from_kept(in, out, keep, seed);
    free(keep);
    return rc;
}

/* GraphFilter::generate_graph (src/GraphFilters.cpp:16-177) */
int og_graph_filter(const og_graph *in, og_graph *out, float thr, int keep_ge, uint64_t seed)
{
    if (!in->w) return -3;
    uint8_t *keep = (uint8_t *)malloc(in->m ? in->m : 1);
    for (uint64_t r = 0; r < in->m; ++r) keep[r] = keep_ge ? (in->w[r] >= thr) : (in->w[r] < thr);
    int rc = rebuild_from_kept(in, out, keep, seed);
    free(keep);
    return rc;
}

/* ===================================================================== */
/* Seeded generators                                                       */
/* ===================================================================== */
void og_free(void *p) { free(p); }

/* open-addressing set of u64 keys */
typedef struct { uint64_t *k; uint64_t cap; } og_set;
static int set_insert(og_set *s, uint64_t key)
{
    uint64_t h = key * 0x9E3779B97F4A7C15ull;
    uint64_t i = h & (s->cap - 1);
    while (s->k[i] != ~0ull) {
        if (s->k[i] == key) return 0;
        i = (i + 1) & (s->cap - 1);
    }
    s->k[i] = key;
    return 1;
}

typedef uint64_t (*og_pair_fn)(og_stream *st, const void *ctx);

static int gen_first_distinct(uint64_t m, uint64_t seed, uint64_t maxc, og_pair_fn fn, const void *ctx,
                              uint32_t **src, uint32_t **dst)
{
    og_set set;
    set.cap = 1;
    while (set.cap < 2 * m + 16) set.cap <<= 1;
    set.k = (uint64_t *)malloc(8 * set.cap);
    memset(set.k, 0xff, 8 * set.cap);
    uint32_t *s = (uint32_t *)malloc(4 * (m ? m : 1)), *d = (uint32_t *)malloc(4 * (m ? m : 1));
    uint64_t got = 0;
    for (uint64_t i = 0; i < maxc && got < m; ++i) {
        og_stream st;
        og_stream_init(&st, seed, OG_TAG_GEN, i);
        uint64_t key = fn(&st, ctx);
        if (key == ~0ull) continue;
        if (set_insert(&set, key)) { s[got] = (uint32_t)(key >> 32); d[got] = (uint32_t)key; ++got; }
    }
    free(set.k);
    if (got < m) { free(s); free(d); return -1; }
    *src = s; *dst = d;
    return 0;
}

static uint64_t pair_key(uint32_t u, uint32_t v)
{
    if (u == v) return ~0ull;
    uint32_t a = u < v ? u : v, b = u < v ? v : u;
    return ((uint64_t)a << 32) | b;
}

static uint64_t er_pair(og_stream *st, const void *ctx)
{
    uint32_t n = *(const uint32_t *)ctx;
    uint32_t u = (uint32_t)og_uniform(st, n);
    uint32_t v = (uint32_t)og_uniform(st, n);
    return pair_key(u, v);
}

int og_gen_er(uint32_t n, uint64_t m, uint64_t seed, uint64_t maxc, uint32_t **src, uint32_t **dst)
{
    return gen_first_distinct(m, seed, maxc, er_pair, &n, src, dst);
}

typedef struct { uint32_t n; double *prob; uint32_t *alias; } cl_ctx;

static uint32_t cl_draw(og_stream *st, const cl_ctx *c)
{
    uint32_t j = (uint32_t)og_uniform(st, c->n);
    uint64_t hi = og_next_u32(st), lo = og_next_u32(st);
    double coin = (double)((hi << 21) ^ (lo >> 11)) * 0x1p-53;
    return coin < c->prob[j] ? j : c->alias[j];
}

static uint64_t cl_pair(og_stream *st, const void *ctx)
{
    const cl_ctx *c = (const cl_ctx *)ctx;
    uint32_t u = cl_draw(st, c), v = cl_draw(st, c);
    return pair_key(u, v);
}

/* Chung-Lu weights (i + i0)^(-1/(beta-1)) with i0 set by bisection so that
 * 2m w_0 / sum w = max_deg; Vose's alias table in double precision. */
int og_gen_chung_lu(uint32_t n, uint64_t m, double beta, double max_deg, uint64_t seed, uint64_t maxc,
                    uint32_t **src, uint32_t **dst)
{
    const double alpha = 1.0 / (beta - 1.0);
    const double target = max_deg / (2.0 * (double)m);
    double lo = 1e-3, hi = (double)n;
    for (int it = 0; it < 60; ++it) {
        double mid = sqrt(lo * hi), sm = 0;
        for (uint32_t i = 0; i < n; ++i) sm += pow(i + mid, -alpha);
        if (pow(mid, -alpha) / sm > target) lo = mid; else hi = mid;
    }
    const double i0 = sqrt(lo * hi);
    double *w = (double *)malloc(8 * (size_t)n), *prob = (double *)malloc(8 * (size_t)n);
    uint32_t *alias = (uint32_t *)malloc(4 * (size_t)n);
    uint32_t *small = (uint32_t *)malloc(4 * (size_t)n), *large = (uint32_t *)malloc(4 * (size_t)n);
    uint32_t ns = 0, nl = 0;
    double sm = 0;
    for (uint32_t i = 0; i < n; ++i) { w[i] = pow(i + i0, -alpha); sm += w[i]; }
    for (uint32_t i = 0; i < n; ++i) {
        w[i] = w[i] * n / sm;
        if (w[i] < 1.0) small[ns++] = i; else large[nl++] = i;
    }
    while (ns && nl) {
        uint32_t l = small[--ns];
        uint32_t g = large[nl - 1];
        prob[l] = w[l]; alias[l] = g;
        w[g] = (w[g] + w[l]) - 1.0;
        if (w[g] < 1.0) { --nl; small[ns++] = g; }
    }
    for (uint32_t i = 0; i < nl; ++i) { prob[large[i]] = 1.0; alias[large[i]] = large[i]; }
    for (uint32_t i = 0; i < ns; ++i) { prob[small[i]] = 1.0; alias[small[i]] = small[i]; }
    cl_ctx c = {n, prob, alias};
    int rc = gen_first_distinct(m, seed, maxc, cl_pair, &c, src, dst);
    free(w); free(prob); free(alias); free(small); free(large);
    return rc;
}

typedef struct { int scale; double a, ab, abc; } rmat_ctx;

static uint64_t rmat_pair(og_stream *st, const void *ctx)
{
    const rmat_ctx *c = (const rmat_ctx *)ctx;
    uint32_t u = 0, v = 0;
    for (int l = 0; l < c->scale; ++l) {
        double t = (double)og_next_u32(st) * 0x1p-32;
        uint32_t bu = t >= c->ab, bv = (t >= c->a && t < c->ab) || t >= c->abc;
        u = (u << 1) | bu;
        v = (v << 1) | bv;
    }
    return pair_key(u, v);
}

int og_gen_rmat(int scale, uint64_t m, double a, double b, double c, uint64_t seed, uint64_t maxc,
                uint32_t **src, uint32_t **dst)
{
    rmat_ctx x = {scale, a, a + b, a + b + c};
    return gen_first_distinct(m, seed, maxc, rmat_pair, &x, src, dst);
}

/* ===================================================================== */
/* Per-estimator state: Edge / MyList / SubGraph                          */
/* (include/ZGraphInstance.hpp:16-105, src/ZGraphInstance.cpp:39-113)     */
/* ===================================================================== */
#define OG_MAX_LIST 16
typedef struct { uint64_t rank; uint32_t src, dst; } og_edge;
typedef struct { int ne, nv; og_edge e[OG_MAX_LIST]; uint32_t v[OG_MAX_LIST]; uint64_t min_valid; } og_sub;

static void sub_init(og_sub *s) { s->ne = s->nv = 0; s->min_valid = 0; }
static int sub_has_vertex(const og_sub *s, uint32_t v)
{
    for (int i = 0; i < s->nv; ++i) if (s->v[i] == v) return 1;
    return 0;
}
static void sub_add_vertex(og_sub *s, uint32_t v)
{
    if (!sub_has_vertex(s, v)) { if (s->nv < OG_MAX_LIST) s->v[s->nv++] = v; else abort(); }
}
/* SubGraph::add_edge (:56-73): append, bump min_valid, add src then dst */
static void sub_add_edge(og_sub *s, og_edge e)
{
    if (s->ne >= OG_MAX_LIST) abort();
    s->e[s->ne++] = e;
    if (e.rank >= s->min_valid) s->min_valid = e.rank + 1;
    sub_add_vertex(s, e.src);
    sub_add_vertex(s, e.dst);
}
/* SubGraph::operator- (:96-113): undirected endpoint difference */
static void sub_minus(const og_sub *a, const og_sub *b, og_sub *out)
{
    sub_init(out);
    for (int i = 0; i < a->ne; ++i) {
        const og_edge *x = &a->e[i];
        int del = 0;
        for (int j = 0; j < b->ne; ++j) {
            const og_edge *y = &b->e[j];
            if ((x->src == y->src && x->dst == y->dst) || (x->src == y->dst && x->dst == y->src)) { del = 1; break; }
        }
        if (!del) sub_add_edge(out, *x);
    }
}

/* ===================================================================== */
/* Primitives (src/ZGraphInstance.cpp)                                    */
/* ===================================================================== */
typedef struct {
    const og_graph *g;
    og_stream st;
    og_u128 num;        /* m * prod s_i (* method factor) */
    int log2den;        /* number of c_i == 2 */
    int ncond;
    int overflow;
    uint32_t e0;
} og_ctx;

static void ctx_mul(og_ctx *c, uint64_t f)
{
    og_u128 r;
    if (__builtin_mul_overflow(c->num, (og_u128)f, &r)) c->overflow = 1;
    c->num = r;
}

/* sample_edge (:314-334) */
static int prim_sample_edge(og_ctx *c, og_edge *e)
{
    const og_graph *g = c->g;
    if (g->m == 0) return 0;
    uint64_t r = og_uniform(&c->st, g->m);
    e->rank = r; e->src = g->src[r]; e->dst = g->dst[r];
    c->e0 = (uint32_t)r;
    c->num = g->m;
    c->log2den = 0;
    return 1;
}

#define OG_BINARY_SEARCH_THRESHOLD 32   /* src/ZGraphInstance.cpp:11 */

/* sizing pass of core_conditional_sample_edge (:127-184) */
uint64_t og_cond_valid_count(const og_graph *g, uint64_t min_valid, const uint32_t *verts,
                             int nv, uint64_t *pos_out)
{
    uint64_t sum = 0;
    for (int i = 0; i < nv; ++i) {
        uint32_t v = verts[i];
        uint64_t b = g->off[v], e = g->off[v + 1];
        if (b == e) { pos_out[i] = e; continue; }      /* bitmap bit unset */
        if (min_valid <= g->adj_rank[b]) { sum += e - b; pos_out[i] = b; }
        else if (min_valid > g->adj_rank[e - 1]) { pos_out[i] = e; }
        else if (e - b <= OG_BINARY_SEARCH_THRESHOLD) {
            for (uint64_t x = b; x < e; ++x)
                if (min_valid <= g->adj_rank[x]) { sum += e - x; pos_out[i] = x; break; }
        } else {
            uint64_t l = b, r = e - 1;                  /* rank[l] < min_valid <= rank[r] */
            while (r - l > 1) {
                uint64_t mid = (l + r) >> 1;
                if (g->adj_rank[mid] < min_valid) l = mid; else r = mid;
            }
            sum += e - r; pos_out[i] = r;
        }
    }
    return sum;
}

/* selection pass (:186-208) */
int og_cond_pick(const og_graph *g, uint64_t min_valid, const uint32_t *verts, int nv,
                 uint64_t r, uint64_t *rank, uint32_t *src, uint32_t *dst)
{
    uint64_t pos[OG_MAX_LIST];
    og_cond_valid_count(g, min_valid, verts, nv, pos);
    for (int i = 0; i < nv; ++i) {
        uint64_t cur = pos[i], e = g->off[verts[i] + 1];
        if (g->off[verts[i]] == e) continue;
        if (cur + r < e) {
            cur += r;
            *rank = g->adj_rank[cur]; *src = g->adj_src[cur]; *dst = g->adj_dst[cur];
            return 1;
        }
        r -= e - cur;
    }
    return 0;
}

/* core_conditional_sample_edge (:127-222); returns 0 when sum == 0 (p = -1) */
static int prim_cond(og_ctx *c, const og_edge *edges, int ne, const uint32_t *verts, int nv, og_edge *out,
                     uint64_t *s_out, int *c_out)
{
    const og_graph *g = c->g;
    c->ncond++;
    uint64_t min_valid = 0;
    for (int i = 0; i < ne; ++i) if (edges[i].rank + 1 > min_valid) min_valid = edges[i].rank + 1;
    uint64_t pos[OG_MAX_LIST];
    uint64_t sum = og_cond_valid_count(g, min_valid, verts, nv, pos);
    if (sum == 0) return 0;
    uint64_t r = og_uniform(&c->st, sum);
    int found = 0;
    for (int i = 0; i < nv; ++i) {
        uint64_t cur = pos[i], e = g->off[verts[i] + 1];
        if (g->off[verts[i]] == e) continue;
        if (cur + r < e) {
            cur += r;
            out->rank = g->adj_rank[cur]; out->src = g->adj_src[cur]; out->dst = g->adj_dst[cur];
            found = 1;
            break;
        }
        r -= e - cur;
    }
    if (!found) abort();
    int cnt = 0;
    for (int i = 0; i < nv; ++i) { cnt += verts[i] == out->src; cnt += verts[i] == out->dst; }
    if (cnt != 1 && cnt != 2) abort();
    *s_out = sum; *c_out = cnt;
    ctx_mul(c, sum);
    if (cnt == 2) c->log2den++;
    return 1;
}

/* conditional_close (:371-407) and the rank-reporting MyList variant
 * (:409-441): linear scans of every listed vertex's whole adjacency */
static int close_scan(const og_graph *g, uint64_t min_valid, const uint32_t *verts, int nv,
                      uint32_t a, uint32_t b, uint64_t *rank_out)
{
    for (int j = 0; j < nv; ++j) {
        uint32_t v = verts[j];
        for (uint64_t x = g->off[v]; x < g->off[v + 1]; ++x) {
            if (g->adj_rank[x] < min_valid) continue;
            if ((g->adj_src[x] == a && g->adj_dst[x] == b) || (g->adj_dst[x] == a && g->adj_src[x] == b)) {
                if (rank_out) *rank_out = g->adj_rank[x];
                return 1;
            }
        }
    }
    return 0;
}

int og_close(const og_graph *g, uint64_t min_valid, const uint32_t *verts, int nv,
             const uint32_t *wsrc, const uint32_t *wdst, int nw)
{
    for (int i = 0; i < nw; ++i)
        if (!close_scan(g, min_valid, verts, nv, wsrc[i], wdst[i], NULL)) return 0;
    return 1;
}

static int prim_close_sub(og_ctx *c, const og_sub *sampled, const og_sub *wait)
{
    for (int i = 0; i < wait->ne; ++i)
        if (!close_scan(c->g, sampled->min_valid, sampled->v, sampled->nv, wait->e[i].src, wait->e[i].dst, NULL))
            return 0;
    return 1;
}

/* ===================================================================== */
/* estimate() bodies                                                      */
/* ===================================================================== */
static void set_ok(og_outcome *o, const og_ctx *c)
{
    o->status = c->overflow ? 3 : 1;
    o->num_lo = (uint64_t)c->num;
    o->num_hi = (uint64_t)(c->num >> 64);
    o->log2den = (uint8_t)c->log2den;
}

/* Triangle(e0, e1) (applications/Triangle.cpp:17-40) */
static int triangle_third(const og_edge *e0, const og_edge *e1, og_edge *t)
{
    t->rank = 0;
    if (e0->src == e1->src) { t->src = e0->dst; t->dst = e1->dst; }
    else if (e0->src == e1->dst) { t->src = e0->dst; t->dst = e1->src; }
    else if (e0->dst == e1->src) { t->src = e0->src; t->dst = e1->dst; }
    else if (e0->dst == e1->dst) { t->src = e0->src; t->dst = e1->src; }
    else return 0;
    return 1;
}

/* wedge prefix shared by Triangle / ThreeChain / ThreeMotif:
 * sample_edge, then conditional_sample_edge(subgraph{e0}) */
static int wedge(og_ctx *c, og_sub *sub, og_edge *e0, og_edge *e1)
{
    sub_init(sub);
    if (!prim_sample_edge(c, e0)) return 0;
    sub_add_edge(sub, *e0);
    uint64_t s; int cc;
    if (!prim_cond(c, sub->e, sub->ne, sub->v, sub->nv, e1, &s, &cc)) return 0;
    sub_add_edge(sub, *e1);
    return 1;
}

/* TriangleCoutingInstance::estimate (applications/Triangle.cpp:48-73) */
static int est_triangle_after_wedge(og_ctx *c, const og_sub *sub, const og_edge *e0, const og_edge *e1)
{
    og_sub tri, wait;
    og_edge t;
    sub_init(&tri);
    sub_add_edge(&tri, *e0);
    sub_add_edge(&tri, *e1);
    if (!triangle_third(e0, e1, &t)) abort();
    sub_add_edge(&tri, t);
    sub_minus(&tri, sub, &wait);
    return prim_close_sub(c, sub, &wait);
}

/* canonicality_check (src/SamplerGenerator.cpp:286-303,
 * applications/ChainMining.cpp:23-40) */
static int canonical(const og_edge *edges, int ne, const og_edge *nw)
{
    if (ne <= 1) return 1;
    int found = 0;
    for (int i = 0; i < ne; ++i) {
        if (!found) {
            if (nw->src == edges[i].src || nw->src == edges[i].dst || nw->dst == edges[i].src || nw->dst == edges[i].dst)
                found = 1;
        } else if (edges[i].rank >= nw->rank) return 0;
    }
    return 1;
}

/* ChainMiningInstance::estimate (applications/ChainMining.cpp:45-84) */
static int est_chain(og_ctx *c, int order)
{
    og_edge edges[OG_MAX_LIST], e0, e;
    og_sub sub;
    uint32_t ends[2];
    int ne = 0;
    sub_init(&sub);
    if (!prim_sample_edge(c, &e0)) return 0;
    edges[ne++] = e0;
    sub_add_edge(&sub, e0);
    ends[0] = e0.src; ends[1] = e0.dst;
    for (int i = 0; i < order - 2; ++i) {
        uint64_t s; int cc;
        if (!prim_cond(c, &e0, 1, ends, 2, &e, &s, &cc)) return 0;
        if (!canonical(edges, ne, &e)) return 0;
        edges[ne++] = e;
        sub_add_edge(&sub, e);
        if (sub.nv != i + 3) return 0;
        if (ends[0] == e.src) ends[0] = e.dst;
        else if (ends[0] == e.dst) ends[0] = e.src;
        else if (ends[1] == e.src) ends[1] = e.dst;
        else if (ends[1] == e.dst) ends[1] = e.src;
        else abort();
    }
    return 1;
}

/* the vertex of e1 that the next FourChain step extends from
 * (applications/FourChain.cpp:42-53) */
static uint32_t fc_far(const og_edge *e0, const og_edge *e1)
{
    if (e0->src == e1->src) return e1->dst;
    if (e0->src == e1->dst) return e1->src;
    if (e0->dst == e1->src) return e1->dst;
    if (e0->dst == e1->dst) return e1->src;
    abort();
}

/* FourChainCountingInstance::estimate (applications/FourChain.cpp:24-124) */
static int est_four_chain(og_ctx *c)
{
    uint64_t method = 1 + og_uniform(&c->st, 3);     /* sample_interger(1, 3) */
    og_edge e0, e1, e2, edges[2];
    uint32_t v[2];
    uint64_t s; int cc;
    if (!prim_sample_edge(c, &e0)) return 0;
    ctx_mul(c, 3);                                    /* p *= 1/3 */
    if (method == 1) {
        edges[0] = e0; v[0] = e0.src; v[1] = e0.dst;
        if (!prim_cond(c, edges, 1, v, 2, &e1, &s, &cc)) return 0;
        edges[1] = e1;
        v[0] = fc_far(&e0, &e1);
        if (!prim_cond(c, edges, 2, v, 1, &e2, &s, &cc)) return 0;
        if (e2.src == e0.src || e2.src == e0.dst) return 0;
        if (e2.dst == e0.src || e2.dst == e0.dst) return 0;
        return 1;
    } else if (method == 2) {
        edges[0] = e0; v[0] = e0.src; v[1] = e0.dst;
        if (!prim_cond(c, edges, 1, v, 2, &e1, &s, &cc)) return 0;
        v[0] = fc_far(&e0, &e1);
        if (!prim_cond(c, edges, 1, v, 1, &e2, &s, &cc)) return 0;
        if (e2.src == e0.src || e2.src == e0.dst) return 0;
        if (e2.dst == e0.src || e2.dst == e0.dst) return 0;
        if (e2.rank >= e1.rank) return 0;
        return 1;
    } else {
        edges[0] = e0; v[0] = e0.src;
        if (!prim_cond(c, edges, 1, v, 1, &e1, &s, &cc)) return 0;
        v[0] = e0.dst;
        if (!prim_cond(c, edges, 1, v, 1, &e2, &s, &cc)) return 0;
        if (e2.src == e1.src || e2.src == e1.dst) return 0;
        if (e2.dst == e1.src || e2.dst == e1.dst) return 0;
        return 1;
    }
}

/* ---- SamplerGenerator (src/SamplerGenerator.cpp) ---------------------- */
typedef struct { int k, e; uint32_t s[OG_MAX_PATTERN_E], d[OG_MAX_PATTERN_E]; } og_pat;

static int pat_has_edge(const og_pat *p, uint32_t a, uint32_t b)
{
    for (int i = 0; i < p->e; ++i)
        if ((p->s[i] == a && p->d[i] == b) || (p->s[i] == b && p->d[i] == a)) return 1;
    return 0;
}
/* Pattern::is_automorphic (:225-242): same sizes, every listed edge present */
static int pat_is_automorphic(const og_pat *p, int k, int e, const uint32_t *s, const uint32_t *d)
{
    if (k != p->k || e != p->e) return 0;
    for (int i = 0; i < e; ++i) if (!pat_has_edge(p, s[i], d[i])) return 0;
    return 1;
}

typedef struct { int nuniq, total; og_pat *uniq; } og_patset;

/* generate_mapping (:262-284) enumerates permutations by DFS in index order */
static void gen_perms(int k, int idx, uint32_t *map, int *used, const og_pat *base, og_pat *out, int *pos)
{
    if (idx == k) {
        og_pat *p = &out[(*pos)++];
        *p = *base;
        for (int i = 0; i < base->e; ++i) { p->s[i] = map[base->s[i]]; p->d[i] = map[base->d[i]]; }
        return;
    }
    for (int v = 0; v < k; ++v) if (!used[v]) {
        used[v] = 1; map[idx] = (uint32_t)v;
        gen_perms(k, idx + 1, map, used, base, out, pos);
        used[v] = 0;
    }
}

/* SamplerGenerator ctor dedupe (:311-363) */
static void patset_build(og_patset *ps, int k, int e, const uint32_t *s, const uint32_t *d)
{
    int total = 1;
    for (int i = 2; i <= k; ++i) total *= i;
    og_pat base; base.k = k; base.e = e;
    for (int i = 0; i < e; ++i) { base.s[i] = s[i]; base.d[i] = d[i]; }
    og_pat *all = (og_pat *)malloc(sizeof(og_pat) * total);
    uint32_t map[OG_MAX_PATTERN_V]; int used[OG_MAX_PATTERN_V] = {0}; int pos = 0;
    gen_perms(k, 0, map, used, &base, all, &pos);
    int *dup = (int *)calloc(total, sizeof(int));
    int nu = 0;
    for (int i = 0; i < total; ++i) {
        for (int j = 0; j < i; ++j)
            if (!dup[j] && pat_is_automorphic(&all[j], all[i].k, all[i].e, all[i].s, all[i].d)) { dup[i] = 1; break; }
        if (!dup[i]) ++nu;
    }
    ps->uniq = (og_pat *)malloc(sizeof(og_pat) * (nu ? nu : 1));
    int t = 0;
    for (int i = 0; i < total; ++i) if (!dup[i]) ps->uniq[t++] = all[i];
    ps->nuniq = nu; ps->total = total;
    free(all); free(dup);
}

int og_pattern_unique_count(int k, int e, const uint32_t *src, const uint32_t *dst, int *total)
{
    og_patset ps;
    patset_build(&ps, k, e, src, dst);
    if (total) *total = ps.total;
    free(ps.uniq);
    return ps.nuniq;
}

/* Pattern::has_same_edge_ordering (:101-204) restated: BFS vertex matching
 * seeded by the rank-0 edge in both orientations. Edge ids of the pattern are
 * given by the rank permutation. */
typedef struct { int k, e; uint32_t s[OG_MAX_PATTERN_E], d[OG_MAX_PATTERN_E], id[OG_MAX_PATTERN_E]; } og_rpat;

static int rp_idx(const og_rpat *p, uint32_t id)
{
    for (int i = 0; i < p->e; ++i) if (p->id[i] == id) return i;
    abort();
}
static int rp_core(const og_rpat *a, uint32_t src, uint32_t dst, const og_rpat *b)
{
    int k = a->k;
    uint32_t map[OG_MAX_PATTERN_V]; int ma[OG_MAX_PATTERN_V] = {0}, mb[OG_MAX_PATTERN_V] = {0};
    uint32_t q[OG_MAX_PATTERN_V]; int qh = 0, qt = 0;
    int ip = rp_idx(b, 0);
    map[src] = b->s[ip]; map[dst] = b->d[ip];
    ma[src] = ma[dst] = 1; mb[b->s[ip]] = mb[b->d[ip]] = 1;
    q[qt++] = src; q[qt++] = dst;
    int ok = 1;
    /* adjacency of a in construct_csr order (:11-48): per vertex, edges in list order */
    while (qh < qt && ok) {
        uint32_t v = q[qh++];
        for (int i = 0; i < a->e && ok; ++i) {
            uint32_t u;
            if (a->s[i] == v) u = a->d[i]; else if (a->d[i] == v) u = a->s[i]; else continue;
            if (ma[u]) continue;
            int jp = rp_idx(b, a->id[i]);
            if (b->s[jp] == map[v]) {
                if (!mb[b->d[jp]]) { map[u] = b->d[jp]; ma[u] = 1; mb[b->d[jp]] = 1; q[qt++] = u; }
                else ok = 0;
            } else if (b->d[jp] == map[v]) {
                if (!mb[b->s[jp]]) { map[u] = b->s[jp]; ma[u] = 1; mb[b->s[jp]] = 1; q[qt++] = u; }
                else ok = 0;
            } else ok = 0;
        }
    }
    for (int v = 0; v < k; ++v) if (!ma[v] || !mb[v]) ok = 0;
    if (!ok) return 0;
    /* is_vertex_mapping_correct (:78-99) */
    for (int i = 0; i < a->e; ++i) {
        int jp = rp_idx(b, a->id[i]);
        uint32_t s = a->s[i], d = a->d[i];
        if (!((map[s] == b->s[jp] && map[d] == b->d[jp]) || (map[s] == b->d[jp] && map[d] == b->s[jp]))) return 0;
    }
    return 1;
}
static int rp_same(const og_rpat *a, const og_rpat *b)
{
    int i0 = rp_idx(a, 0);
    return rp_core(a, a->s[i0], a->d[i0], b) || rp_core(a, a->d[i0], a->s[i0], b);
}

/* number of edge orderings (of e! rank assignments) that are distinct up to
 * automorphism, as printed by tests/TestPattern.cpp:87-125 */
int og_pattern_edge_order_classes(int k, int e, const uint32_t *src, const uint32_t *dst, int *total)
{
    int tot = 1;
    for (int i = 2; i <= e; ++i) tot *= i;
    og_rpat *all = (og_rpat *)malloc(sizeof(og_rpat) * tot);
    uint32_t perm[OG_MAX_PATTERN_E];
    for (int i = 0; i < e; ++i) perm[i] = (uint32_t)i;
    for (int t = 0; t < tot; ++t) {
        og_rpat *p = &all[t];
        p->k = k; p->e = e;
        for (int i = 0; i < e; ++i) { p->s[i] = src[i]; p->d[i] = dst[i]; p->id[i] = perm[i]; }
        /* next permutation (lexicographic) */
        int i = e - 2;
        while (i >= 0 && perm[i] >= perm[i + 1]) --i;
        if (i >= 0) {
            int j = e - 1;
            while (perm[j] <= perm[i]) --j;
            uint32_t x = perm[i]; perm[i] = perm[j]; perm[j] = x;
            for (int a = i + 1, b = e - 1; a < b; ++a, --b) { x = perm[a]; perm[a] = perm[b]; perm[b] = x; }
        }
    }
    int *dup = (int *)calloc(tot, sizeof(int));
    int nu = 0;
    for (int i = 0; i < tot; ++i) {
        for (int j = 0; j < i; ++j) if (!dup[j] && rp_same(&all[i], &all[j])) { dup[i] = 1; break; }
        if (!dup[i]) ++nu;
    }
    free(all); free(dup);
    if (total) *total = tot;
    return nu;
}

/* SamplerGenerator::estimate (src/SamplerGenerator.cpp:378-464) */
static int est_pattern(og_ctx *c, const og_patset *ps, int k, int pe)
{
    og_edge sampled[OG_MAX_LIST], e0, e;
    og_sub sub;
    sub_init(&sub);
    if (!prim_sample_edge(c, &e0)) return 0;
    int ns = 0;
    sampled[ns++] = e0;
    sub_add_edge(&sub, e0);
    for (int i = 1; i < pe; ++i) {
        uint64_t s; int cc;
        if (!prim_cond(c, &e0, 1, sub.v, sub.nv, &e, &s, &cc)) return 0;
        if (!canonical(sampled, ns, &e)) return 0;
        sampled[ns++] = e;
        sub_add_edge(&sub, e);
        if (sub.nv > k) return 0;
    }
    if (sub.nv != k) return 0;
    uint32_t vs[OG_MAX_LIST];
    for (int i = 0; i < k; ++i) vs[i] = sub.v[i];
    for (int i = 0; i < k; ++i)            /* selection sort (:415-423) */
        for (int j = i + 1; j < k; ++j)
            if (vs[j] < vs[i]) { uint32_t t = vs[j]; vs[j] = vs[i]; vs[i] = t; }
    uint32_t rs[OG_MAX_LIST], rd[OG_MAX_LIST];
    for (int i = 0; i < ns; ++i) {          /* relabel (:425-445) */
        int fs = 0, fd = 0;
        for (int j = 0; j < k; ++j) if (sampled[i].src == vs[j]) { rs[i] = (uint32_t)j; fs = 1; break; }
        for (int j = 0; j < k; ++j) if (sampled[i].dst == vs[j]) { rd[i] = (uint32_t)j; fd = 1; break; }
        if (!fs || !fd) abort();
    }
    for (int u = 0; u < ps->nuniq; ++u)
        if (pat_is_automorphic(&ps->uniq[u], k, pe, rs, rd)) return 1;
    return 0;
}

/* ---- 4-clique by three spanning shapes plus closing checks ------------- *
 * No reference counterpart is unbiased: applications/FourClique.cpp:41-76
 * captures a clique only when its three lowest-rank edges form a spanning
 * tree whose second edge touches the first (E = 3/5 of the count on K4).
 * This plan adds the two missing rank shapes.  For a clique with ranks
 * r1 < ... < r6: shape A (r1, r2 adjacent; r3 reaches the 4th vertex),
 * shape B (r1, r2, r3 form a triangle), shape C (r1, r2 disjoint).  Each
 * shape has exactly one draw sequence that reaches it, with probability
 * 1/(m s1 s2), so X = X_A + X_B + X_C is unbiased for every rank order. */
static int close1(og_ctx *c, uint64_t min_valid, const uint32_t *v, int nv, uint32_t a, uint32_t b, uint64_t *rank)
{
    return close_scan(c->g, min_valid, v, nv, a, b, rank);
}

/* first stage: e0, e1 from {a, b} above rank(e0), and the triangle edge
 * {ob, w} above rank(e1) (every shape needs it); fills the stage state and
 * leaves c->num = m s1 */
typedef struct { og_edge e0, e1; uint32_t a, b, o, w, ob; uint64_t rt; } k4_stage;

static int k4_first(og_ctx *c, k4_stage *k)
{
    og_edge edges[1];
    uint64_t s1; int c1;
    if (!prim_sample_edge(c, &k->e0)) return 0;
    k->a = k->e0.src; k->b = k->e0.dst;
    uint32_t v2[2] = {k->a, k->b};
    edges[0] = k->e0;
    if (!prim_cond(c, edges, 1, v2, 2, &k->e1, &s1, &c1)) return 0;
    if (c1 == 2) return 0;
    k->o = k->e1.src; k->w = k->e1.dst; k->ob = (k->o == k->a) ? k->b : k->a;
    uint32_t v3[3] = {k->a, k->b, k->w};
    return close1(c, k->e1.rank + 1, v3, 3, k->ob, k->w, &k->rt);
}

/* the three shapes after a closed first stage, each with one conditional
 * sample from c->st; returns the sum of the shapes' values m s1 s2 */
static og_u128 k4_shapes(og_ctx *c, const k4_stage *k, og_u128 base, int base_ov, int *ov)
{
    const uint32_t a = k->a, b = k->b, o = k->o, w = k->w, ob = k->ob;
    const og_edge e0 = k->e0, e1 = k->e1;
    og_edge e2, edges[3];
    uint64_t s2; int c2;
    uint32_t v3[3] = {a, b, w};
    og_u128 total = 0;
    /* shape A: e2 from V3 above rank(e1), new vertex d */
    edges[0] = e0; edges[1] = e1;
    c->num = base; c->overflow = base_ov;
    if (prim_cond(c, edges, 2, v3, 3, &e2, &s2, &c2) && c2 == 1) {
        uint32_t d = e2.dst;
        uint32_t v4[4] = {a, b, w, d};
        uint32_t P[6][2] = {{a, b}, {a, w}, {a, d}, {b, w}, {b, d}, {w, d}};
        int okc = 1;
        for (int i = 0; i < 6 && okc; ++i) {
            uint32_t x = P[i][0], y = P[i][1];
            int have = ((x == a && y == b) || (x == b && y == a)) ||
                       ((x == o && y == w) || (x == w && y == o)) ||
                       ((x == e2.src && y == d) || (x == d && y == e2.src));
            if (!have && !close1(c, e2.rank + 1, v4, 4, x, y, NULL)) okc = 0;
        }
        if (okc) { total += c->num; *ov |= c->overflow; }
    }
    /* shape B: triangle {e0, e1, t}, then e2 above rank(t) to new vertex d */
    {
        og_edge t; t.rank = k->rt; t.src = ob; t.dst = w;
        edges[1] = e1; edges[2] = t;
        c->num = base; c->overflow = base_ov;
        if (prim_cond(c, edges, 3, v3, 3, &e2, &s2, &c2) && c2 == 1) {
            uint32_t d = e2.dst;
            uint32_t v4[4] = {a, b, w, d};
            int okc = 1;
            for (int i = 0; i < 3 && okc; ++i) {
                uint32_t x = v3[i];
                if (x == e2.src) continue;
                if (!close1(c, e2.rank + 1, v4, 4, x, d, NULL)) okc = 0;
            }
            if (okc) { total += c->num; *ov |= c->overflow; }
        }
    }
    /* shape C: e2 from V3 above rank(e0), disjoint from e0, below rank(e1) */
    c->num = base; c->overflow = base_ov;
    if (prim_cond(c, edges, 1, v3, 3, &e2, &s2, &c2) && c2 == 1 && e2.src == w && e2.rank < e1.rank) {
        uint32_t d = e2.dst;
        uint32_t v4[4] = {a, b, w, d};
        if (close1(c, e1.rank + 1, v4, 4, ob, w, NULL) && close1(c, e1.rank + 1, v4, 4, a, d, NULL) &&
            close1(c, e1.rank + 1, v4, 4, b, d, NULL)) {
            total += c->num; *ov |= c->overflow;
        }
    }
    return total;
}

static int est_k4_close(og_ctx *c)
{
    k4_stage k;
    if (!k4_first(c, &k)) return 0;
    int ov = 0;
    og_u128 total = k4_shapes(c, &k, c->num, c->overflow, &ov);
    c->num = total; c->overflow = ov; c->log2den = 0;
    return total != 0;
}

/* Split 4-clique plan: est_k4_close's first stage and, when its triangle
 * closes, J independent completions -- the three shapes, completion j drawing
 * from the stream (seed, OG_TAG_COMPLETION + j, id) -- averaged: value =
 * (sum over completions of the shape values) / J, J a power of two.  Each
 * completion has the distribution of est_k4_close's shapes given the first
 * stage, so the mean is unchanged; the rare completions that give the value
 * its variance are drawn J times per closed triangle.  nwords counts the
 * first stage's stream, ncond every conditional sample (saturating). */
static int est_k4_split(og_ctx *c, int J, uint64_t seed, uint64_t id)
{
    k4_stage k;
    if (!k4_first(c, &k)) return 0;
    const og_u128 base = c->num;
    const int base_ov = c->overflow;
    og_stream main = c->st;
    og_u128 total = 0;
    int ov = 0;
    for (int j = 0; j < J; ++j) {
        og_stream_init(&c->st, seed, OG_TAG_COMPLETION + (uint32_t)j, id);
        const og_u128 t = k4_shapes(c, &k, base, base_ov, &ov);
        if (total + t < total) ov = 1;
        total += t;
    }
    c->st = main;
    int lg = 0;
    while ((1 << lg) < J) ++lg;
    c->num = total; c->overflow = ov; c->log2den = lg;
    return total != 0;
}

/* Multiplicity of the pair {x, y}: the stored edges joining them (parallel
 * copies count one each), by a scan of the shorter of the two lists. */
static uint32_t pair_mult(const og_graph *g, uint32_t x, uint32_t y)
{
    if (g->off[y + 1] - g->off[y] < g->off[x + 1] - g->off[x]) { uint32_t t = x; x = y; y = t; }
    uint32_t n = 0;
    for (uint64_t i = g->off[x]; i < g->off[x + 1]; ++i) n += g->adj_dst[i] == y;
    return n;
}

/* 5-house, split (no reference counterpart: the replacement of
 * SamplerGenerator's house estimator for time to a target error, as
 * est_k4_split is for 4-cliques).  The house of applications/FiveHouse.cpp:18-27
 * is the roof triangle {0, 1, 2} on the square 1-3-4-2; exactly one of its
 * edges, the base {1, 2}, lies in both its triangle and its 4-cycle, so a
 * house is one tuple (base edge {a, b}, apex w ~ a, b; square vertices x ~ a
 * and y ~ b with x ~ y; five distinct vertices), and the copy-weighted count
 * weighs it by the multiplicities of its six pairs.
 * First stage: e0 of rank r0 ~ U[m]; a = its endpoint of smaller degree (the
 * source on ties), b the other; an entry ~ U[deg a] of a's list gives the apex
 * w; the roof closes iff w != b and mult(w, b) > 0.  Completion j (stream
 * (seed, OG_TAG_COMPLETION + j, id)): an entry ~ U[deg a] of a's list gives x,
 * then an entry ~ U[deg b] of b's list gives y; it counts mult(x, y) when
 * x != b, y != a, x != y and w is neither.  value = m deg(a) mult(w, b)
 * deg(a) deg(b) (sum_j count_j) / J, so E[value | e0] = m times the
 * copy-weighted houses on base e0.  ncond: 1 (the roof) + 2 J once it closes. */
static int est_house_split(og_ctx *c, int J, uint64_t seed, uint64_t id)
{
    const og_graph *g = c->g;
    if (g->m == 0) return 0;
    const uint64_t r0 = og_uniform(&c->st, g->m);
    c->e0 = (uint32_t)r0;
    uint32_t a = g->src[r0], b = g->dst[r0];
    uint64_t da = g->off[a + 1] - g->off[a], db = g->off[b + 1] - g->off[b];
    if (db < da) {
        uint32_t t = a; a = b; b = t;
        uint64_t u = da; da = db; db = u;
    }
    c->ncond = 1;
    const uint32_t w = g->adj_dst[g->off[a] + og_uniform(&c->st, da)];
    if (w == b) return 0;
    const uint32_t mwb = pair_mult(g, w, b);
    if (!mwb) return 0;
    og_stream main = c->st;
    uint64_t S = 0;
    for (int j = 0; j < J; ++j) {
        og_stream_init(&c->st, seed, OG_TAG_COMPLETION + (uint32_t)j, id);
        const uint32_t x = g->adj_dst[g->off[a] + og_uniform(&c->st, da)];
        const uint32_t y = g->adj_dst[g->off[b] + og_uniform(&c->st, db)];
        if (x == b || y == a || x == y || x == w || y == w) continue;
        S += pair_mult(g, x, y);
    }
    c->st = main;
    c->ncond = 1 + 2 * J;
    if (!S) return 0;
    c->num = g->m;
    c->overflow = 0;
    ctx_mul(c, da); ctx_mul(c, mwb); ctx_mul(c, da); ctx_mul(c, db); ctx_mul(c, S);
    int lg = 0;
    while ((1 << lg) < J) ++lg;
    c->log2den = lg;
    return 1;
}

/* ---- plan dispatch ----------------------------------------------------- */
int og_plan_channels(const og_plan *p) { return p->kind == OG_PLAN_THREE_MOTIF ? 2 : 1; }

int og_plan_max_log2den(const og_plan *p)
{
    switch (p->kind) {
    case OG_PLAN_TRIANGLE: case OG_PLAN_THREE_CHAIN: case OG_PLAN_THREE_MOTIF: return 1;
    case OG_PLAN_CHAIN: return p->order - 2;
    case OG_PLAN_FOUR_CHAIN: return 2;
    case OG_PLAN_PATTERN: return p->pe - 1;
    case OG_PLAN_K4_CLOSE: return 0;
    case OG_PLAN_K4_SPLIT: case OG_PLAN_HOUSE_SPLIT: { int lg = 0; while ((1 << lg) < p->order) ++lg; return lg; }
    default: return 0;
    }
}

double og_plan_cond_calls(const og_plan *p)
{
    switch (p->kind) {
    case OG_PLAN_TRIANGLE: case OG_PLAN_THREE_CHAIN: case OG_PLAN_THREE_MOTIF: return 1;
    case OG_PLAN_CHAIN: return p->order - 2;
    case OG_PLAN_FOUR_CHAIN: return 2;
    case OG_PLAN_PATTERN: return p->pe - 1;   /* SamplerGenerator.cpp:471-474 */
    case OG_PLAN_K4_CLOSE: case OG_PLAN_K4_SPLIT: return 2;
    case OG_PLAN_HOUSE_SPLIT: return 3;   /* the roof, x and y */
    default: return 0;
    }
}

/* pattern sets are built once per plan and cached (not thread-safe to build) */
static og_patset *patset_for(const og_plan *p)
{
    static og_plan cached;
    static og_patset ps;
    static int have = 0;
    if (!have || memcmp(&cached, p, sizeof(og_plan)) != 0) {
        if (have) free(ps.uniq);
        patset_build(&ps, p->pk, p->pe, p->pat_src, p->pat_dst);
        cached = *p; have = 1;
    }
    return &ps;
}

static void estimate_with(const og_graph *g, const og_plan *p, const og_patset *ps, uint64_t seed, uint64_t id, og_outcome *out)
{
    int nch = og_plan_channels(p);
    og_ctx c;
    c.g = g; c.num = 0; c.log2den = 0; c.ncond = 0; c.overflow = 0; c.e0 = 0;
    og_stream_init(&c.st, seed, OG_TAG_ESTIMATOR, id);
    for (int i = 0; i < nch; ++i) memset(&out[i], 0, sizeof(og_outcome));
    int ok = 0;
    og_sub sub; og_edge e0, e1;
    switch (p->kind) {
    case OG_PLAN_TRIANGLE:
        ok = wedge(&c, &sub, &e0, &e1) && est_triangle_after_wedge(&c, &sub, &e0, &e1);
        break;
    case OG_PLAN_THREE_CHAIN:
        ok = wedge(&c, &sub, &e0, &e1);
        break;
    case OG_PLAN_THREE_MOTIF:
        ok = wedge(&c, &sub, &e0, &e1);
        if (ok) set_ok(&out[0], &c);
        if (ok && est_triangle_after_wedge(&c, &sub, &e0, &e1)) set_ok(&out[1], &c);
        ok = 0;
        break;
    case OG_PLAN_CHAIN: ok = est_chain(&c, p->order); break;
    case OG_PLAN_FOUR_CHAIN: ok = est_four_chain(&c); break;
    case OG_PLAN_PATTERN: ok = est_pattern(&c, ps, p->pk, p->pe); break;
    case OG_PLAN_K4_CLOSE: ok = est_k4_close(&c); break;
    case OG_PLAN_K4_SPLIT: ok = est_k4_split(&c, p->order, seed, id); break;
    case OG_PLAN_HOUSE_SPLIT: ok = est_house_split(&c, p->order, seed, id); break;
    default: abort();
    }
    if (ok) set_ok(&out[0], &c);
    uint8_t nw = c.st.word > 255 ? 255 : (uint8_t)c.st.word;
    for (int i = 0; i < nch; ++i) {
        out[i].e0 = c.e0;
        out[i].ncond = c.ncond > 255 ? 255 : (uint8_t)c.ncond;
        out[i].nwords = nw;
        if (!out[i].status) { out[i].num_lo = out[i].num_hi = 0; out[i].log2den = 0; }
    }
}

void og_estimate(const og_graph *g, const og_plan *p, uint64_t seed, uint64_t id, og_outcome *out)
{
    const og_patset *ps = p->kind == OG_PLAN_PATTERN ? patset_for(p) : NULL;
    estimate_with(g, p, ps, seed, id, out);
}

/* ---- calculate (src/ZGraphInstance.cpp:256-297) ------------------------ */
static void acc_add(og_sums *s, const og_outcome *o)
{
    s->count++;
    if (o->status != 1) return;
    s->successes++;
    int sh = s->K - o->log2den;
    og_u128 v = ((og_u128)o->num_hi << 64) | o->num_lo;
    og_u128 lo = v << sh;
    uint64_t hi = sh ? (uint64_t)(v >> (128 - sh)) : 0;
    og_u128 cur = ((og_u128)s->sum[1] << 64) | s->sum[0];
    og_u128 nxt = cur + lo;
    s->sum[0] = (uint64_t)nxt; s->sum[1] = (uint64_t)(nxt >> 64);
    s->sum[2] += hi + (nxt < cur);
    long double x = (long double)v / ldexpl(1.0L, o->log2den);
    s->sumsq += (double)(x * x);
}

static void acc_merge(og_sums *a, const og_sums *b)
{
    og_u128 x = ((og_u128)a->sum[1] << 64) | a->sum[0];
    og_u128 y = ((og_u128)b->sum[1] << 64) | b->sum[0];
    og_u128 z = x + y;
    a->sum[0] = (uint64_t)z; a->sum[1] = (uint64_t)(z >> 64);
    a->sum[2] += b->sum[2] + (z < x);
    a->sumsq += b->sumsq;
    a->successes += b->successes;
    a->count += b->count;
}

void og_run(const og_graph *g, const og_plan *p, uint64_t seed, uint64_t first,
            uint64_t count, int threads, og_sums *sums)
{
    int nch = og_plan_channels(p);
    int K = og_plan_max_log2den(p);
    const og_patset *ps = p->kind == OG_PLAN_PATTERN ? patset_for(p) : NULL;
    for (int ch = 0; ch < nch; ++ch) { memset(&sums[ch], 0, sizeof(og_sums)); sums[ch].K = K; }
    if (threads <= 0) threads = omp_get_max_threads();
    #pragma omp parallel num_threads(threads)
    {
        og_sums loc[2];
        memset(loc, 0, sizeof(loc));
        loc[0].K = loc[1].K = K;
        og_outcome o[2];
        #pragma omp for schedule(dynamic, 4096)
        for (int64_t i = 0; i < (int64_t)count; ++i) {
            estimate_with(g, p, ps, seed, first + (uint64_t)i, o);
            for (int ch = 0; ch < nch; ++ch) acc_add(&loc[ch], &o[ch]);
        }
        #pragma omp critical
        for (int ch = 0; ch < nch; ++ch) acc_merge(&sums[ch], &loc[ch]);
    }
}

void og_run_outcomes(const og_graph *g, const og_plan *p, uint64_t seed, uint64_t first,
                     uint64_t count, int threads, og_outcome *out)
{
    int nch = og_plan_channels(p);
    const og_patset *ps = p->kind == OG_PLAN_PATTERN ? patset_for(p) : NULL;
    if (threads <= 0) threads = omp_get_max_threads();
    #pragma omp parallel for num_threads(threads) schedule(dynamic, 4096)
    for (int64_t i = 0; i < (int64_t)count; ++i)
        estimate_with(g, p, ps, seed, first + (uint64_t)i, &out[(uint64_t)i * nch]);
}

double og_sums_mean(const og_sums *s)
{
    if (!s->count) return 0;
    long double v = (long double)s->sum[2] * ldexpl(1.0L, 128) + (long double)s->sum[1] * ldexpl(1.0L, 64) +
                    (long double)s->sum[0];
    return (double)(v / ldexpl(1.0L, s->K) / (long double)s->count);
}

/* ===================================================================== */
/* Exhaustive expectation by replaying every choice sequence              */
/* ===================================================================== */
struct og_enum {
    int depth, len;
    uint64_t choice[64], n[64];
};

static uint64_t enum_draw(og_enum *e, uint64_t n)
{
    if (e->depth == e->len) {
        if (e->len >= 64) abort();
        e->choice[e->len] = 0; e->n[e->len] = n; e->len++;
    }
    if (e->n[e->depth] != n) abort();          /* replay must be deterministic */
    return e->choice[e->depth++];
}

double og_expectation(const og_graph *g, const og_plan *p, int channel, uint64_t max_paths)
{
    og_enum en; memset(&en, 0, sizeof(en));
    const og_patset *ps = p->kind == OG_PLAN_PATTERN ? patset_for(p) : NULL;
    int nch = og_plan_channels(p);
    long double E = 0;
    uint64_t paths = 0;
    tl_enum = &en;
    for (;;) {
        en.depth = 0;
        og_outcome o[2];
        estimate_with(g, p, ps, 0, 0, o);
        if (en.depth != en.len) abort();
        if (++paths > max_paths) { tl_enum = NULL; return -1; }
        if (channel < nch && o[channel].status == 1) {
            long double wgt = 1.0L;
            for (int i = 0; i < en.len; ++i) wgt /= (long double)en.n[i];
            og_u128 v = ((og_u128)o[channel].num_hi << 64) | o[channel].num_lo;
            E += (long double)v / ldexpl(1.0L, o[channel].log2den) * wgt;
        }
        /* odometer */
        int j = en.len - 1;
        while (j >= 0 && en.choice[j] + 1 >= en.n[j]) --j;
        if (j < 0) break;
        en.choice[j]++;
        en.len = j + 1;
    }
    tl_enum = NULL;
    return (double)E;
}

/* ===================================================================== */
/* Exact counts (naive_implementation sources, non-induced counts)      */
/* ===================================================================== */
typedef struct { uint32_t n; uint64_t *off; uint32_t *nb; } og_simple;

static int u32cmp(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

/* simple undirected graph: sorted, de-duplicated neighbour lists */
static void simple_build(const og_graph *g, og_simple *s)
{
    s->n = g->n;
    uint64_t *off = (uint64_t *)calloc((size_t)g->n + 1, 8);
    uint32_t *nb = (uint32_t *)malloc(4 * (g->m ? 2 * g->m : 1));
    for (uint32_t v = 0; v < g->n; ++v) {
        uint64_t b = g->off[v], e = g->off[v + 1];
        for (uint64_t x = b; x < e; ++x) nb[x] = g->adj_dst[x];
        qsort(nb + b, e - b, 4, u32cmp);
    }
    uint64_t w = 0;
    for (uint32_t v = 0; v < g->n; ++v) {
        uint64_t b = g->off[v], e = g->off[v + 1];
        off[v] = w;
        for (uint64_t x = b; x < e; ++x)
            if (x == b || nb[x] != nb[x - 1]) nb[w++] = nb[x];
    }
    off[g->n] = w;
    s->off = off; s->nb = nb;
}
static void simple_free(og_simple *s) { free(s->off); free(s->nb); }
static uint64_t sdeg(const og_simple *s, uint32_t v) { return s->off[v + 1] - s->off[v]; }
static int s_less(const og_simple *s, uint32_t a, uint32_t b)
{   /* degree order, ties by id */
    uint64_t da = sdeg(s, a), db = sdeg(s, b);
    return da < db || (da == db && a < b);
}
static int s_has(const og_simple *s, uint32_t a, uint32_t b)
{
    uint64_t lo = s->off[a], hi = s->off[a + 1];
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (s->nb[mid] < b) lo = mid + 1; else hi = mid;
    }
    return lo < s->off[a + 1] && s->nb[lo] == b;
}

/* oriented out-lists (to higher degree order), sorted by id */
static void orient(const og_simple *s, uint64_t **ooff, uint32_t **onb)
{
    uint64_t *off = (uint64_t *)calloc((size_t)s->n + 1, 8);
    for (uint32_t v = 0; v < s->n; ++v)
        for (uint64_t x = s->off[v]; x < s->off[v + 1]; ++x)
            if (s_less(s, v, s->nb[x])) off[v + 1]++;
    for (uint32_t v = 0; v < s->n; ++v) off[v + 1] += off[v];
    uint32_t *nb = (uint32_t *)malloc(4 * (off[s->n] ? off[s->n] : 1));
    for (uint32_t v = 0; v < s->n; ++v) {
        uint64_t w = off[v];
        for (uint64_t x = s->off[v]; x < s->off[v + 1]; ++x)
            if (s_less(s, v, s->nb[x])) nb[w++] = s->nb[x];
    }
    *ooff = off; *onb = nb;
}

uint64_t og_exact_triangles(const og_graph *g, int threads)
{
    og_simple s; simple_build(g, &s);
    uint64_t *off; uint32_t *nb; orient(&s, &off, &nb);
    uint64_t T = 0;
    if (threads <= 0) threads = omp_get_max_threads();
    #pragma omp parallel for num_threads(threads) reduction(+:T) schedule(dynamic, 256)
    for (int64_t u = 0; u < (int64_t)s.n; ++u)
        for (uint64_t x = off[u]; x < off[u + 1]; ++x) {
            uint32_t v = nb[x];
            uint64_t i = off[u], j = off[v];
            while (i < off[u + 1] && j < off[v + 1]) {
                if (nb[i] < nb[j]) ++i; else if (nb[i] > nb[j]) ++j; else { ++T; ++i; ++j; }
            }
        }
    free(off); free(nb); simple_free(&s);
    return T;
}

uint64_t og_exact_k4(const og_graph *g, int threads)
{
    og_simple s; simple_build(g, &s);
    uint64_t *off; uint32_t *nb; orient(&s, &off, &nb);
    uint64_t K = 0;
    if (threads <= 0) threads = omp_get_max_threads();
    #pragma omp parallel num_threads(threads) reduction(+:K)
    {
        uint32_t *W = (uint32_t *)malloc(4 * ((size_t)g->max_degree + 1));
        #pragma omp for schedule(dynamic, 64)
        for (int64_t u = 0; u < (int64_t)s.n; ++u)
            for (uint64_t x = off[u]; x < off[u + 1]; ++x) {
                uint32_t v = nb[x];
                uint64_t nw = 0, i = off[u], j = off[v];
                while (i < off[u + 1] && j < off[v + 1]) {
                    if (nb[i] < nb[j]) ++i; else if (nb[i] > nb[j]) ++j; else { W[nw++] = nb[i]; ++i; ++j; }
                }
                for (uint64_t a = 0; a < nw; ++a) {   /* edges w1 -> w2 inside W */
                    uint32_t w1 = W[a];
                    uint64_t p = off[w1], q = 0;
                    while (p < off[w1 + 1] && q < nw) {
                        if (nb[p] < W[q]) ++p; else if (nb[p] > W[q]) ++q; else { ++K; ++p; ++q; }
                    }
                }
            }
        free(W);
    }
    free(off); free(nb); simple_free(&s);
    return K;
}

uint64_t og_exact_wedges(const og_graph *g)
{
    og_simple s; simple_build(g, &s);
    uint64_t W = 0;
    for (uint32_t v = 0; v < s.n; ++v) { uint64_t d = sdeg(&s, v); W += d * (d - (d > 0)) / 2; }
    simple_free(&s);
    return W;
}

og_u128 og_exact_stars(const og_graph *g, int leaves)
{
    og_simple s; simple_build(g, &s);
    og_u128 S = 0;
    for (uint32_t v = 0; v < s.n; ++v) {
        uint64_t d = sdeg(&s, v);
        if (d < (uint64_t)leaves) continue;
        og_u128 c = 1;
        for (int i = 0; i < leaves; ++i) c = c * (og_u128)(d - i) / (og_u128)(i + 1);
        S += c;
    }
    simple_free(&s);
    return S;
}

void og_exact_stars64(const og_graph *g, int leaves, uint64_t *lo, uint64_t *hi)
{
    const og_u128 s = og_exact_stars(g, leaves);
    *lo = (uint64_t)s;
    *hi = (uint64_t)(s >> 64);
}

/* 3-edge simple paths: sum over edges (du-1)(dv-1) - 3 T */
uint64_t og_exact_paths3(const og_graph *g, int threads)
{
    og_simple s; simple_build(g, &s);
    uint64_t P = 0;
    for (uint32_t u = 0; u < s.n; ++u)
        for (uint64_t x = s.off[u]; x < s.off[u + 1]; ++x) {
            uint32_t v = s.nb[x];
            if (u < v) P += (sdeg(&s, u) - 1) * (sdeg(&s, v) - 1);
        }
    simple_free(&s);
    return P - 3 * og_exact_triangles(g, threads);
}

/* 4-cycles by their definition: every 4-cycle has two diagonals, and a pair
 * of vertices {x, y} with c common neighbours is the diagonal of C(c, 2)
 * 4-cycles, so C4 = (1/2) sum over pairs x < y of C(c_xy, 2).  For each x the
 * common-neighbour counts of all y > x come from walking x's 2-hop paths into
 * a dense counter: O(sum of d^2). */
uint64_t og_exact_c4(const og_graph *g)
{
    og_simple s; simple_build(g, &s);
    uint32_t *cnt = (uint32_t *)calloc((size_t)s.n + 1, 4);
    uint32_t *seen = (uint32_t *)malloc(4 * ((size_t)s.n + 1));
    uint64_t pairs = 0;
    for (uint32_t x = 0; x < s.n; ++x) {
        uint64_t ns = 0;
        for (uint64_t i = s.off[x]; i < s.off[x + 1]; ++i) {
            uint32_t v = s.nb[i];
            for (uint64_t j = s.off[v]; j < s.off[v + 1]; ++j) {
                uint32_t y = s.nb[j];
                if (y <= x) continue;
                if (cnt[y]++ == 0) seen[ns++] = y;
            }
        }
        for (uint64_t k = 0; k < ns; ++k) {
            uint64_t c = cnt[seen[k]];
            pairs += c * (c - 1) / 2;
            cnt[seen[k]] = 0;
        }
    }
    free(cnt); free(seen); simple_free(&s);
    return pairs / 2;
}

/* 5-houses by a closed form: with t_e the triangles and c4_e the 4-cycles
 * through edge e of the simple graph, a house (roof triangle on a square) is
 * an edge e in both, an apex from e's t_e common neighbours and a 4-cycle
 * through e, minus the choices whose apex lies on the cycle; over the
 * house's two automorphisms,
 *   houses = sum_e t_e c4_e - 2 sum_e t_e (t_e - 1)
 * (K5: 10 * 3 * 6 - 2 * 10 * 6 = 60).  sum_e t_e c4_e = the sum over 4-cycles
 * of their edges' t, counted here from both diagonals {x, y} of each cycle:
 * k common neighbours v_i of x and y close C(k, 2) cycles whose weights sum
 * to (k - 1) sum_i (t(x, v_i) + t(v_i, y)).  Pinned against
 * og_exact_house_brute and FiveHouse.cpp:68's wiki-vote count. */
uint64_t og_exact_house(const og_graph *g)
{
    og_simple s; simple_build(g, &s);
    const uint64_t ne = s.off[s.n];
    uint32_t *te = (uint32_t *)malloc(4 * (ne ? ne : 1));
    unsigned __int128 tt = 0;   /* sum over entries (each edge twice) of t (t - 1) */
    for (uint32_t v = 0; v < s.n; ++v)
        for (uint64_t i = s.off[v]; i < s.off[v + 1]; ++i) {
            const uint32_t u = s.nb[i];
            uint64_t a = s.off[v], b = s.off[u], c = 0;
            while (a < s.off[v + 1] && b < s.off[u + 1]) {
                if (s.nb[a] < s.nb[b]) ++a;
                else if (s.nb[a] > s.nb[b]) ++b;
                else { ++c; ++a; ++b; }
            }
            te[i] = (uint32_t)c;
            tt += (unsigned __int128)c * (c ? c - 1 : 0);
        }
    uint64_t *cnt = (uint64_t *)calloc((size_t)s.n + 1, 8);
    unsigned __int128 *wsum = (unsigned __int128 *)calloc((size_t)s.n + 1, sizeof(unsigned __int128));
    uint32_t *seen = (uint32_t *)malloc(4 * ((size_t)s.n + 1));
    unsigned __int128 w2 = 0;   /* every 4-cycle's weight, once per diagonal */
    for (uint32_t x = 0; x < s.n; ++x) {
        uint64_t ns = 0;
        for (uint64_t i = s.off[x]; i < s.off[x + 1]; ++i) {
            const uint32_t v = s.nb[i];
            for (uint64_t j = s.off[v]; j < s.off[v + 1]; ++j) {
                const uint32_t y = s.nb[j];
                if (y <= x) continue;
                if (cnt[y]++ == 0) seen[ns++] = y;
                wsum[y] += (unsigned __int128)te[i] + te[j];
            }
        }
        for (uint64_t k = 0; k < ns; ++k) {
            const uint32_t y = seen[k];
            w2 += (unsigned __int128)(cnt[y] - 1) * wsum[y];
            cnt[y] = 0;
            wsum[y] = 0;
        }
    }
    free(cnt); free(wsum); free(seen); free(te); simple_free(&s);
    return (uint64_t)(w2 / 2 - tt);
}

/* 5-vertex chains (simple paths of four edges, FiveChain.cpp's count) by a
 * closed form over the path's middle vertex v and its neighbours x, y on the
 * path: the ordered choices (x, v, y) with an end beyond x and one beyond y
 * number (d_x - 1 - a)(d_y - 1 - a) - (c_xy - 1), a = [x ~ y], c_xy the
 * common neighbours of x and y.  Summed over ordered wedges:
 *   2P = sum_v [(sum_{x~v} (d_x - 1))^2 - sum_{x~v} (d_x - 1)^2]
 *        - 4 sum_v (d_v - 1) t_v + 6 T - 8 C4,
 * t_v the triangles at v.  Pinned against og_exact_paths_brute(5). */
uint64_t og_exact_paths4(const og_graph *g)
{
    og_simple s; simple_build(g, &s);
    __int128 two_p = 0, T = 0, TW = 0;
    for (uint32_t v = 0; v < s.n; ++v) {
        unsigned __int128 s1 = 0, s2 = 0;
        uint64_t tv = 0;
        for (uint64_t i = s.off[v]; i < s.off[v + 1]; ++i) {
            uint32_t x = s.nb[i];
            uint64_t a = sdeg(&s, x) - 1;
            s1 += a; s2 += (unsigned __int128)a * a;
            for (uint64_t j = i + 1; j < s.off[v + 1]; ++j) tv += s_has(&s, x, s.nb[j]);
        }
        two_p += (__int128)(s1 * s1 - s2);
        T += tv;
        TW += (__int128)(sdeg(&s, v) - (sdeg(&s, v) > 0)) * tv;
    }
    simple_free(&s);
    two_p += -4 * TW + 6 * (T / 3) - 8 * (__int128)og_exact_c4(g);
    return (uint64_t)(two_p / 2);
}

/* brute force: injective maps of pattern vertices with every pattern edge
 * present, divided by the automorphism count */
static uint64_t bf_rec(const og_simple *s, int k, int e, const uint32_t *ps, const uint32_t *pd, int idx, uint32_t *map)
{
    if (idx == k) return 1;
    uint64_t c = 0;
    for (uint32_t v = 0; v < s->n; ++v) {
        int ok = 1;
        for (int j = 0; j < idx && ok; ++j) if (map[j] == v) ok = 0;
        for (int i = 0; i < e && ok; ++i) {
            uint32_t a = ps[i], b = pd[i];
            if (a == (uint32_t)idx && b < (uint32_t)idx && !s_has(s, v, map[b])) ok = 0;
            if (b == (uint32_t)idx && a < (uint32_t)idx && !s_has(s, v, map[a])) ok = 0;
        }
        if (!ok) continue;
        map[idx] = v;
        c += bf_rec(s, k, e, ps, pd, idx + 1, map);
    }
    return c;
}

uint64_t og_exact_pattern_brute(const og_graph *g, int k, int e, const uint32_t *ps, const uint32_t *pd)
{
    og_simple s; simple_build(g, &s);
    uint32_t map[OG_MAX_PATTERN_V];
    uint64_t emb = bf_rec(&s, k, e, ps, pd, 0, map);
    simple_free(&s);
    /* automorphisms: permutations mapping the edge set onto itself */
    int tot; int uniq = og_pattern_unique_count(k, e, ps, pd, &tot);
    uint64_t aut = (uint64_t)(tot / uniq);
    return emb / aut;
}

uint64_t og_exact_paths_brute(const og_graph *g, int nverts)
{
    uint32_t ps[OG_MAX_PATTERN_E], pd[OG_MAX_PATTERN_E];
    for (int i = 0; i + 1 < nverts; ++i) { ps[i] = (uint32_t)i; pd[i] = (uint32_t)i + 1; }
    return og_exact_pattern_brute(g, nverts, nverts - 1, ps, pd);
}

uint64_t og_exact_house_brute(const og_graph *g)
{   /* applications/FiveHouse.cpp:18-27 pattern */
    uint32_t ps[6] = {0, 0, 1, 1, 2, 3}, pd[6] = {1, 2, 2, 3, 4, 4};
    return og_exact_pattern_brute(g, 5, 6, ps, pd);
}

/* ===================================================================== */
/* Regression ELP (src/Regression.cpp, src/ErrorLatencyProfile.cpp)       */
/* ===================================================================== */
long double og_regression_k(const long double *x, const long double *y, unsigned len)
{   /* UnbiasedLinearRegression::solve (src/Regression.cpp:17-28) */
    long double t0 = 0, t1 = 0;
    for (unsigned i = 0; i < len; ++i) t0 += x[i] * y[i];
    for (unsigned i = 0; i < len; ++i) t1 += x[i] * x[i];
    return t0 / t1;
}

double og_elp_bound(double eps, double delta, double a, double b, double est,
                    uint64_t max_degree, uint64_t num_edges)
{   /* ErrorProfile::calculate_bound (:51-59) */
    double f0 = 1. / pow(eps, 2.);
    double f1 = pow((double)max_degree, b) * pow((double)num_edges, a) / est;
    double f2 = log(2. / delta);
    return f0 * f1 * f2;
}

double og_elp_num_estimators(const long double *est, uint64_t npts, long double F,
                             double a, double b, uint64_t samp_maxdeg, uint64_t samp_m,
                             uint64_t maxdeg, uint64_t m, double eps, double conf,
                             long double *K_out)
{   /* ErrorProfile::get_num_estimators (:127-200) */
    const unsigned long long per = 16;
    long double *X = (long double *)malloc(sizeof(long double) * (npts ? npts : 1));
    long double *Y = (long double *)malloc(sizeof(long double) * (npts ? npts : 1));
    long double ema = 100;
    double error_times = 1, alpha = 0.04;
    double Fd = (double)F;
    for (uint64_t i = 0; i < npts; ++i) {
        long double err = (est[i] - Fd) / Fd * 5;
        if (err < 0) err = -err;
        ema = ema * (1 - alpha) + err * alpha;
        double xt = (double)((long double)og_elp_bound((double)(ema * error_times), 1 - conf, a, b, Fd, samp_maxdeg, samp_m)
                             * powl(ema * error_times, 2));
        X[i] = sqrt(xt / (double)((i + 1) * per));
        Y[i] = ema;
    }
    double K = (double)og_regression_k(X, Y, (unsigned)npts);
    long double rough = (long double)Fd / powl(0.1L, (long double)(a + b));
    double N = og_elp_bound(eps, 1. - conf, a, b, (double)rough, maxdeg, m);
    N *= (K * K);
    if (K_out) *K_out = K;
    free(X); free(Y);
    return N;
}
