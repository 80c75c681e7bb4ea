/*
 * zg_oracle.h -- CPU restatement of ZGraph's ASAP neighbour-sampling estimator
 * path, used ONLY as test infrastructure (the checker in tests/, smoke() and
 * bench.py's cpu_baseline leg).  Nothing in the product links or calls it.
 *
 * Every function restates a reference routine and cites it (paths relative to
 * the jy-yuan/sampling-graph-mining tree).  Two deliberate, documented changes
 * from the reference, shared with the product so that per-estimator outcomes
 * are comparable bit for bit:
 *   - RNG: std::mt19937 + uniform_int_distribution (include/Random.hpp:6-16,
 *     src/Random.cpp:11-14) is replaced by Philox4x32-10 keyed by
 *     (seed, estimator id) with Lemire's nearly-divisionless bounded draw.
 *   - Edge ranks: the per-socket Fisher-Yates shuffle (src/Graph.cpp:218-231)
 *     is replaced by a keyed random permutation (sort by a Philox key per
 *     input edge), which is uniform and reproducible.
 * Values are kept exact: an estimator value 1/p = m * prod(s_i) / prod(c_i)
 * (src/ZGraphInstance.cpp:209-220) is returned as an unsigned 128-bit
 * numerator and a power-of-two denominator exponent.
 *
 * Parity status per function is listed at the top of oracle.c.
 */
#ifndef ZG_ORACLE_H
#define ZG_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned __int128 og_u128;

/* ---- Philox4x32-10 and the estimator streams -------------------------- */
void og_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* stream tags (counter word 1) */
enum { OG_TAG_ESTIMATOR = 0, OG_TAG_SHUFFLE = 1, OG_TAG_SUBSAMPLE = 2, OG_TAG_GEN = 3,
       OG_TAG_COMPLETION = 0x1000 /* + j: completion j of a split estimator */ };

typedef struct {
    uint32_t key[2];
    uint32_t tag;
    uint64_t id;
    uint64_t word;     /* next word index */
    uint32_t buf[4];
    uint64_t buf_block;/* block held in buf, or UINT64_MAX */
} og_stream;

void og_stream_init(og_stream *s, uint64_t seed, uint32_t tag, uint64_t id);
uint32_t og_next_u32(og_stream *s);
/* 0: Philox streams (parity); 1: per-thread MT19937 draws (cost-faithful timing only) */
void og_set_rng_mode(int mode);
uint32_t og_mt19937_nth(uint32_t seed, uint32_t n);   /* n-th output of MT19937 seeded `seed` */
/* U[0, n-1], 1 <= n <= 2^64-1 */
uint64_t og_uniform(og_stream *s, uint64_t n);
/* first 64 bits of block 0 of (seed, tag, id): the per-edge permutation key */
uint64_t og_key64(uint64_t seed, uint32_t tag, uint64_t id);

/* ---- graph (one socket, one MPI rank: W = 1) -------------------------- */
typedef struct {
    uint32_t n;
    uint64_t m;               /* stored edges ("num_local_edges") */
    uint64_t num_edges;       /* graph.num_edges used by the ELP bound */
    uint64_t max_degree;
    uint32_t *src, *dst;      /* edges by rank ("local_edges") */
    float *w;                 /* per-edge weight by rank, or NULL */
    uint64_t *off;            /* n+1 ("outgoing_adj_index") */
    uint64_t *adj_rank;       /* 2m ("AdjUnit.edge_id"), ascending per vertex */
    uint32_t *adj_src;        /* 2m (owner) */
    uint32_t *adj_dst;        /* 2m (neighbour) */
} og_graph;

/* Build from an input edge list (in file order). Returns 0 on success,
 * -1 on a self-loop (src/Graph.cpp:454 asserts), -2 on an id >= n. */
int og_graph_build(og_graph *g, uint32_t n, uint64_t m, const uint32_t *src,
                   const uint32_t *dst, const float *w, uint64_t shuffle_seed);
void og_graph_free(og_graph *g);
/* Adopts a CSR built elsewhere (copies it): src/dst by rank [m], off [n+1],
 * per-list ranks/neighbours [2m].  Used to time the CPU path on a graph the
 * device built; tests pin og_graph_build against the device build separately. */
int og_graph_from_csr(og_graph *g, uint32_t n, uint64_t m, const uint32_t *src, const uint32_t *dst,
                      const uint64_t *off, const uint32_t *adj_rank, const uint32_t *adj_nbr);
/* GraphSampler::sample_graph (src/Graph.cpp:628-780) at W=1 */
int og_graph_subsample(const og_graph *in, og_graph *out, double rate, uint64_t seed);
/* GraphFilter::generate_graph (src/GraphFilters.cpp:16-177); keeps edges with
 * (w >= thr) when keep_ge != 0, else (w < thr). */
int og_graph_filter(const og_graph *in, og_graph *out, float thr, int keep_ge, uint64_t shuffle_seed);

/* ---- seeded generators (restating the library's device generators) ---- *
 * Candidate pair i draws from stream (seed, OG_TAG_GEN, i); the output is the
 * first m distinct non-loop pairs (min, max) in candidate order.  Caller frees
 * *src and *dst.  Returns 0, or -1 if max_candidates were not enough. */
int og_gen_er(uint32_t n, uint64_t m, uint64_t seed, uint64_t max_candidates, uint32_t **src, uint32_t **dst);
int og_gen_chung_lu(uint32_t n, uint64_t m, double beta, double max_deg, uint64_t seed,
                    uint64_t max_candidates, uint32_t **src, uint32_t **dst);
int og_gen_rmat(int scale, uint64_t m, double a, double b, double c, uint64_t seed,
                uint64_t max_candidates, uint32_t **src, uint32_t **dst);
void og_free(void *p);

/* ---- plans (estimate() bodies) ---------------------------------------- */
enum {
    OG_PLAN_TRIANGLE = 1,      /* applications/Triangle.cpp:48-73 */
    OG_PLAN_THREE_CHAIN = 2,   /* applications/ThreeChain.cpp:22-37 */
    OG_PLAN_THREE_MOTIF = 3,   /* applications/ThreeMotif.cpp:48-121 (2 channels: 3-chain, triangle) */
    OG_PLAN_CHAIN = 4,         /* applications/ChainMining.cpp:45-84, param = order */
    OG_PLAN_FOUR_CHAIN = 5,    /* applications/FourChain.cpp:24-124 */
    OG_PLAN_PATTERN = 6,       /* src/SamplerGenerator.cpp:378-464 */
    OG_PLAN_K4_CLOSE = 7,      /* 4-clique: three spanning shapes + closing checks */
    OG_PLAN_FOUR_CLIQUE_REF = 8,/* applications/FourClique.cpp:41-76 (biased, kept for the record) */
    OG_PLAN_K4_SPLIT = 9,      /* OG_PLAN_K4_CLOSE's first stage + `order` (a power of two) completions */
    OG_PLAN_HOUSE_SPLIT = 11   /* 5-house: base edge + roof, then `order` (a power of two) square completions */
};

#define OG_MAX_PATTERN_V 8
#define OG_MAX_PATTERN_E 16

typedef struct {
    int kind;
    int order;                 /* OG_PLAN_CHAIN */
    int pk, pe;                /* OG_PLAN_PATTERN: vertices, edges */
    uint32_t pat_src[OG_MAX_PATTERN_E], pat_dst[OG_MAX_PATTERN_E];
} og_plan;

typedef struct {
    uint64_t num_lo, num_hi;   /* value = num / 2^log2den (num as two 64-bit limbs: 24-byte record) */
    uint32_t e0;               /* rank of the first sampled edge (diagnostic) */
    uint8_t status;            /* 1 = success */
    uint8_t log2den;
    uint8_t ncond;             /* conditional_sample_edge calls made */
    uint8_t nwords;            /* RNG words consumed (saturating) */
} og_outcome;

int og_plan_channels(const og_plan *p);
/* max denominator exponent of the plan (number of conditional samples) */
int og_plan_max_log2den(const og_plan *p);
/* conditional calls per success ("get_num_conditional_sample_edge_calls") */
double og_plan_cond_calls(const og_plan *p);

/* run one estimator id; out has og_plan_channels() entries */
void og_estimate(const og_graph *g, const og_plan *p, uint64_t seed, uint64_t id, og_outcome *out);

typedef struct {
    uint64_t sum[3];           /* sum of num * 2^(K - log2den), 192 bits little-endian */
    double sumsq;              /* sum of value^2 (value = num/2^log2den) */
    uint64_t successes;
    uint64_t count;
    int K;                     /* og_plan_max_log2den */
} og_sums;

/* ZGraphInstance::calculate's loop (src/ZGraphInstance.cpp:256-297) over ids
 * [first, first+count) with `threads` OpenMP threads (0 = default). */
void og_run(const og_graph *g, const og_plan *p, uint64_t seed, uint64_t first,
            uint64_t count, int threads, og_sums *sums /* [channels] */);
void og_run_outcomes(const og_graph *g, const og_plan *p, uint64_t seed, uint64_t first,
                     uint64_t count, int threads, og_outcome *out /* [count*channels] */);
double og_sums_mean(const og_sums *s);

/* Exhaustive expectation: E[X | graph, ranks] over every random choice, as
 * a double (sum of dyadic rationals).  Returns -1 if more than max_paths
 * execution paths would be needed. channel selects the output. */
double og_expectation(const og_graph *g, const og_plan *p, int channel, uint64_t max_paths);

/* ---- SamplerGenerator pattern machinery ------------------------------- */
/* unique relabelings (src/SamplerGenerator.cpp:311-363) */
int og_pattern_unique_count(int k, int e, const uint32_t *src, const uint32_t *dst, int *total);
/* edge-order classes (has_same_edge_ordering, src/SamplerGenerator.cpp:101-204) */
int og_pattern_edge_order_classes(int k, int e, const uint32_t *src, const uint32_t *dst, int *total);

/* ---- reference primitives exposed for brute-force tests --------------- */
/* core_conditional_sample_edge sizing: number of valid entries and the
 * lower_bound positions (src/ZGraphInstance.cpp:127-184). */
uint64_t og_cond_valid_count(const og_graph *g, uint64_t min_valid, const uint32_t *verts,
                             int nv, uint64_t *pos_out);
/* the r-th valid entry in list order (src/ZGraphInstance.cpp:186-208) */
int og_cond_pick(const og_graph *g, uint64_t min_valid, const uint32_t *verts, int nv,
                 uint64_t r, uint64_t *rank, uint32_t *src, uint32_t *dst);
/* conditional_close (src/ZGraphInstance.cpp:371-407), linear scans */
int og_close(const og_graph *g, uint64_t min_valid, const uint32_t *verts, int nv,
             const uint32_t *wsrc, const uint32_t *wdst, int nw);

/* ---- algorithmic line count of the device kernels (lines.c) ------------ */
/* 128-byte lines that the device estimator kernel of `kind` (triangle, 3-chain,
 * 3-motif, 4-clique) reads for estimators [first, first+count) under the
 * device layout; UINT64_MAX for other kinds. */
uint64_t og_device_lines(const og_graph *g, int kind, uint64_t seed, uint64_t first, uint64_t count);
/* the same for OG_PLAN_HOUSE_SPLIT with `completions` per closed roof */
uint64_t og_device_lines_house(const og_graph *g, int completions, uint64_t seed, uint64_t first, uint64_t count);

/* ---- exact counts (naive_implementation semantics) --------------------- */
uint64_t og_exact_triangles(const og_graph *g, int threads);
uint64_t og_exact_k4(const og_graph *g, int threads);
uint64_t og_exact_wedges(const og_graph *g);              /* ThreeChain.cpp:44-51 */
uint64_t og_exact_paths3(const og_graph *g, int threads);  /* 3-edge simple paths */
uint64_t og_exact_c4(const og_graph *g);                  /* 4-cycles, by pair counting */
uint64_t og_exact_paths4(const og_graph *g);              /* 4-edge simple paths (FiveChain.cpp), closed form */
og_u128 og_exact_stars(const og_graph *g, int leaves);     /* FiveStar.cpp:46-60 for leaves=4 */
void og_exact_stars64(const og_graph *g, int leaves, uint64_t *lo, uint64_t *hi);
uint64_t og_exact_paths_brute(const og_graph *g, int nverts); /* simple paths with nverts vertices */
uint64_t og_exact_house_brute(const og_graph *g);          /* 5-house, non-induced */
uint64_t og_exact_house(const og_graph *g);                /* 5-house by per-edge triangles and 4-cycles */
uint64_t og_exact_pattern_brute(const og_graph *g, int k, int e, const uint32_t *ps, const uint32_t *pd);

/* ---- regression error-latency profile (src/ErrorLatencyProfile.cpp) --- */
long double og_regression_k(const long double *x, const long double *y, unsigned len);
double og_elp_bound(double eps, double delta, double a, double b, double est,
                    uint64_t max_degree, uint64_t num_edges);
/* get_num_estimators' arithmetic (:127-200) given the pilot running means
 * est[i] (after 16*(i+1) estimators), the final mean F, call exponents a, b,
 * the sampled graph's (Delta', m'), and the original graph's (Delta, m). */
double og_elp_num_estimators(const long double *est, uint64_t npts, long double F,
                             double a, double b, uint64_t samp_maxdeg, uint64_t samp_m,
                             uint64_t maxdeg, uint64_t m, double eps, double conf,
                             long double *K_out);

#ifdef __cplusplus
}
#endif
#endif
