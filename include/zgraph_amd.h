/*
 * zgraph_amd.h -- C ABI of the MI355X-native ASAP neighbour-sampling
 * estimator library (libzgraph_amd.so).
 *
 * Drop-in boundary for ZGraph's estimator path (jy-yuan/sampling-graph-mining):
 * the reference runs `ZGraphInstance<E>::calculate` (src/ZGraphInstance.cpp:256-297)
 * over user `estimate()` callbacks on a per-socket CSR (include/Graph.hpp:148-166).
 * Here the CSR lives in HBM on one GPU, each estimate() body is a device
 * "plan", and calculate() is one kernel launch over estimator ids
 * [first_id, first_id + count), each drawing from its own Philox4x32-10
 * stream keyed by (seed, estimator id).  Plain pointers and sizes only; every
 * function returns 0 on success or a negative ZG_E* code, with a message
 * available from zg_last_error().  Calls on one graph handle are serialised
 * on that handle's stream.
 */
#ifndef ZGRAPH_AMD_H
#define ZGRAPH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define ZG_OK 0
#define ZG_EINVAL (-1)      /* bad argument (self-loop, id >= n, size limits) */
#define ZG_EHIP (-2)        /* HIP runtime error (message has the HIP string) */
#define ZG_ENOMEM (-3)      /* device allocation failed */
#define ZG_EOVERFLOW (-4)   /* an estimator value exceeded 128 bits */
#define ZG_ESTATE (-5)      /* no GPU / library not initialised / empty result */

typedef struct zg_graph zg_graph;
typedef struct zg_plan zg_plan;

/* Copies the last error message of the calling thread (NUL-terminated). */
int zg_last_error(char *buf, size_t len);

/* Library/device initialisation (device 0).  Called implicitly by the first
 * graph upload; fails with ZG_ESTATE when no HIP device is present. */
int zg_init(void);
/* Device properties used by the launch heuristics: CU count, LDS per CU. */
int zg_device_info(int *num_cus, int *lds_per_cu, int *warp_size);
/* Free and total device memory of the library's device, in bytes (no
 * reference counterpart: callers size graphs to HBM with it). */
int zg_device_memory(uint64_t *free_bytes, uint64_t *total_bytes);

/* ---- graphs ------------------------------------------------------------ *
 * Replaces DistGraphLoader<E>::load_graph + CSRConstructor<E>::
 * construct_bitmap_assisted_csr (src/Graph.cpp:214-377, :386-615) at one
 * worker (W = 1: no partitioning, no cross-partition edge dropping).
 * Input: m edges in file order (src[i], dst[i]) over vertices [0, n), an
 * optional float weight per edge, and a shuffle seed.  Edge ranks ("edge
 * ids") are a keyed random permutation of the input order; every vertex's
 * adjacency is ascending by rank, as the reference asserts
 * (src/Graph.cpp:324-334).  Self-loops are rejected (src/Graph.cpp:454).
 * Requires m < 2^31.  Host pointers. */
int zg_graph_create(const uint32_t *src, const uint32_t *dst, const float *weights,
                    uint64_t m, uint32_t n, uint64_t shuffle_seed, zg_graph **out);
/* Same, with the edge list already in device memory (not consumed). */
int zg_graph_create_device(const uint32_t *d_src, const uint32_t *d_dst, const float *d_weights,
                           uint64_t m, uint32_t n, uint64_t shuffle_seed, zg_graph **out);
int zg_graph_destroy(zg_graph *g);

typedef struct {
    uint32_t n;
    uint64_t m;            /* stored edges ("num_local_edges") */
    uint64_t num_edges;    /* graph.num_edges (ELP bound input) */
    uint64_t max_degree;   /* graph.max_degree */
    uint64_t device_bytes; /* HBM held by the handle */
    int has_weights;
    int has_parallel_edges; /* some vertex pair is joined by more than one edge */
} zg_graph_info;
int zg_graph_info_get(const zg_graph *g, zg_graph_info *out);

/* Copies the device CSR back for inspection.  Any pointer may be NULL.
 * src/dst/w: [m] by rank; off: [n+1]; adj_rank/adj_nbr: [2m], each list
 * ascending by rank (the reference's AdjUnit edge_id / dst). */
int zg_graph_export(const zg_graph *g, uint32_t *src, uint32_t *dst, float *w, uint64_t *off,
                    uint32_t *adj_rank, uint32_t *adj_nbr);

/* GraphSampler<E>::sample_graph (src/Graph.cpp:628-780): keeps each edge iff
 * 1 + U[0, 2^30) <= floor(rate * 2^30) drawn from Philox (seed, edge rank),
 * rebuilds the CSR with a fresh rank permutation (same seed). */
int zg_graph_subsample(const zg_graph *in, double rate, uint64_t seed, zg_graph **out);
/* GraphFilter<E>::generate_graph (src/GraphFilters.cpp:16-177) with the
 * predicate `w >= threshold` (keep_ge != 0) or `w < threshold`. */
int zg_graph_filter(const zg_graph *in, float threshold, int keep_ge, uint64_t shuffle_seed,
                    zg_graph **out);

/* ---- seeded synthetic graphs (device generators) ----------------------- *
 * No reference counterpart (the reference reads SNAP files); these build
 * the benchmark inputs in HBM.  Outputs are device arrays owned by the
 * library, freed with zg_edges_free.  All generators are deterministic in
 * (parameters, seed), de-duplicate undirected pairs and drop self-loops. */
typedef struct {
    uint32_t *d_src, *d_dst;   /* device pointers, m entries, src < dst */
    float *d_w;                /* device pointer or NULL */
    uint64_t m;
    uint32_t n;
} zg_edges;
/* Erdos-Renyi G(n, m): m distinct uniform pairs. */
int zg_gen_erdos_renyi(uint32_t n, uint64_t m, uint64_t seed, zg_edges *out);
/* Chung-Lu: weights w_i = (i + i0)^(-1/(beta-1)); pairs drawn from w x w
 * until m distinct pairs; i0 chosen so the expected top degree is max_deg. */
int zg_gen_chung_lu(uint32_t n, uint64_t m, double beta, double max_deg, uint64_t seed, zg_edges *out);
/* R-MAT (a, b, c, 1-a-b-c) at 2^scale vertices until m distinct pairs. */
int zg_gen_rmat(int scale, uint64_t m, double a, double b, double c, uint64_t seed, zg_edges *out);
/* Attach float weights U[0,1) (Philox keyed by (seed, edge index)). */
int zg_edges_add_uniform_weights(zg_edges *e, uint64_t seed);
/* Copy device edges to host buffers (any may be NULL). */
int zg_edges_download(const zg_edges *e, uint32_t *src, uint32_t *dst, float *w);
int zg_edges_free(zg_edges *e);

/* ---- plans: the estimate() bodies -------------------------------------- */
enum {
    ZG_PLAN_TRIANGLE = 1,     /* TriangleCoutingInstance (applications/Triangle.cpp:48-73) */
    ZG_PLAN_THREE_CHAIN = 2,  /* ThreeChainCountingInstance (applications/ThreeChain.cpp:22-37) */
    ZG_PLAN_THREE_MOTIF = 3,  /* ThreeMotif.cpp:48-121 fused: channel 0 = 3-chain, 1 = triangle */
    ZG_PLAN_CHAIN = 4,        /* ChainMiningInstance(order) (applications/ChainMining.cpp:45-84) */
    ZG_PLAN_FOUR_CHAIN = 5,   /* FourChainCountingInstance (applications/FourChain.cpp:24-124) */
    ZG_PLAN_PATTERN = 6,      /* SamplerGenerator<Empty>(Pattern) (src/SamplerGenerator.cpp:378-464) */
    ZG_PLAN_K4_CLOSE = 7,     /* 4-clique: three spanning shapes + closing checks (unbiased;
                                 replaces the biased applications/FourClique.cpp:41-76) */
    ZG_PLAN_TRIANGLE_K4 = 8,  /* fused: channel 0 = triangle (ZG_PLAN_TRIANGLE), channel 1 =
                                 4-clique (ZG_PLAN_K4_CLOSE) from one first edge, conditional
                                 sample and closing check per id -- the cached sub-pattern
                                 sharing of ThreeMotif (src/ZGraphInstance.cpp:596-833) applied
                                 to triangle -> 4-clique; each channel equals its own plan's run */
    ZG_PLAN_K4_SPLIT = 9,     /* 4-clique, split: ZG_PLAN_K4_CLOSE's first stage (e0, e1, the
                                 triangle edge's closing check), then, when the triangle closes,
                                 `completions` independent draws of the three shapes averaged:
                                 the same mean with far less variance per estimator (no
                                 reference counterpart; each completion is FourClique.cpp's
                                 remaining conditional samples and closing checks) */
    ZG_PLAN_TRIANGLE_K4_SPLIT = 10 /* fused: channel 0 = triangle (ZG_PLAN_TRIANGLE), channel 1 =
                                 the split 4-clique (ZG_PLAN_K4_SPLIT, `completions`) from one
                                 first stage per id; each channel equals its own plan's run */,
    ZG_PLAN_HOUSE_SPLIT = 11  /* 5-house (the pattern of applications/FiveHouse.cpp:18-27), split:
                                 base edge e0 and its roof apex from the smaller-degree endpoint's
                                 list, closed by one pair lookup, then `completions` draws of the
                                 square's two other vertices per closed roof, averaged.  Counts
                                 each house once per set of edge copies (the simple graph's count
                                 when there are no parallel edges).  No reference counterpart:
                                 SamplerGenerator's house estimator rarely succeeds on large
                                 graphs (none in 2^26 on a 69 M-edge Chung-Lu graph) */
};

#define ZG_MAX_PATTERN_EDGES 15
typedef struct {
    int kind;
    int order;                               /* ZG_PLAN_CHAIN: vertices in the path (>= 3) */
    int pattern_vertices;                    /* ZG_PLAN_PATTERN: k (2..6) */
    int pattern_edges;                       /* ZG_PLAN_PATTERN: E (1..15) */
    uint32_t pattern_src[ZG_MAX_PATTERN_EDGES];
    uint32_t pattern_dst[ZG_MAX_PATTERN_EDGES];
    int completions;                         /* ZG_PLAN_K4_SPLIT, ZG_PLAN_TRIANGLE_K4_SPLIT: completions
                                                per closed triangle; ZG_PLAN_HOUSE_SPLIT: per closed
                                                roof; a power of two in [1, 4096] */
} zg_plan_desc;

int zg_plan_create(const zg_plan_desc *desc, zg_plan **out);
int zg_plan_destroy(zg_plan *p);
/* outputs per estimator (2 for ZG_PLAN_THREE_MOTIF, ZG_PLAN_TRIANGLE_K4 and
 * ZG_PLAN_TRIANGLE_K4_SPLIT, else 1) */
int zg_plan_channels(const zg_plan *p);
/* conditional_sample_edge calls per successful estimate
 * (get_num_conditional_sample_edge_calls, src/ZGraphInstance.cpp:470-488;
 * SamplerGenerator override src/SamplerGenerator.cpp:471-474) */
double zg_plan_cond_calls(const zg_plan *p);
/* SamplerGenerator unique relabelings (src/SamplerGenerator.cpp:336-359). */
int zg_plan_unique_patterns(const zg_plan *p, int *unique, int *total);
/* The same count for a pattern description, on the host (no device needed):
 * the k! relabelings and those left after Pattern::is_automorphic
 * deduplication (src/SamplerGenerator.cpp:225-242, 311-363). */
int zg_pattern_unique(const zg_plan_desc *desc, int *unique, int *total);

/* ---- runs (calculate) -------------------------------------------------- */
typedef struct {
    uint64_t sum[3];      /* exact sum of values * 2^log2_den, 192-bit little-endian */
    int32_t log2_den;     /* the estimate is sum / 2^log2_den / count */
    int32_t status;       /* 0 ok, ZG_EOVERFLOW if some value exceeded 128 bits */
    uint64_t successes;   /* estimators with a nonzero value */
    uint64_t count;       /* estimators run */
    double sumsq;         /* sum of squared values (for the CLT interval) */
    double mean;          /* sum / 2^log2_den / count */
} zg_sums;

/* Per-estimator record (parity dumps). */
typedef struct {
    uint64_t num_lo, num_hi; /* value = num / 2^log2den */
    uint32_t e0;             /* rank of the first sampled edge */
    uint8_t status;          /* 0 fail, 1 success, 3 overflow */
    uint8_t log2den;
    uint8_t ncond;           /* conditional_sample_edge calls made */
    uint8_t nwords;          /* Philox words consumed (saturating at 255) */
} zg_outcome;

typedef struct {
    int block_size;     /* threads per workgroup (multiple of 64; 0 = default 256) */
    int blocks_per_cu;  /* workgroups per CU in the grid (0 = default) */
    int order;          /* 0 or 1: estimator-id order; b in 2..16: estimators sorted
                           by the first edge's rank into 2^b buckets before the launch
                           (cache locality of the edge records; same outcomes and exact
                           sums, sumsq up to rounding).  Opt-in: measured slower on
                           MI355X for the 69 M-edge benchmark graph, sort included. */
} zg_launch;

/* The calculate() loop: runs ids [first_id, first_id+count) and reduces
 * exactly (first_id + count must not exceed 2^64 - 1, else ZG_EINVAL; the
 * same holds for zg_run_outcomes, zg_count_reads and batched queries).
 * sums has zg_plan_channels() entries.  kernel_ms (may be NULL)
 * receives the estimator kernel's duration from HIP events on the handle's
 * stream.  Blocks until done. */
int zg_run(zg_graph *g, const zg_plan *p, uint64_t seed, uint64_t first_id, uint64_t count,
           const zg_launch *launch, zg_sums *sums, float *kernel_ms);
/* zg_run that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and sums as zg_run (estimator-id order only: a launch
 * order of 2..16 fails with ZG_EINVAL).  The block partials and the result
 * record live in a fixed device array of the library (16384 partial records
 * of 64 bytes plus the result, about 1 MiB, held by each GPU until the
 * process ends); a launch whose grid needs more partial records than that
 * (workgroups x channels > 16384, e.g. blocks_per_cu > 32 for two-channel
 * plans on 256 CUs) fails with ZG_EINVAL before anything runs.  Every call on
 * one GPU shares that array: calls that overlap in time on one GPU (from two
 * threads, or on two graph handles' streams) race and must be serialised by
 * the caller.  Nothing needs clearing between calls (each launch writes every
 * record it reads). */
int zg_run_noalloc(zg_graph *g, const zg_plan *p, uint64_t seed, uint64_t first_id, uint64_t count,
                   const zg_launch *launch, zg_sums *sums, float *kernel_ms);
/* Same, and copies every estimator's outcome(s) to host memory
 * (count * channels records, channel-minor). */
int zg_run_outcomes(zg_graph *g, const zg_plan *p, uint64_t seed, uint64_t first_id,
                    uint64_t count, zg_outcome *host_out, zg_sums *sums);

/* ---- batched queries (one replayable HIP graph) ------------------------ *
 * No single reference counterpart: many ZGraphInstance::calculate calls
 * (src/ZGraphInstance.cpp:256-297), each a (graph, plan, seed, id range),
 * served together.  One persistent work-queue launch runs the chunks of every
 * query, whatever its plan kind: wavefronts claim chunks of estimator ids
 * from one counter, queries laid out by decreasing cost per estimator; the
 * whole batch is captured once in a HIP graph and replayed by zg_batch_run.  Each query's sums equal zg_run's for the same
 * (graph, plan, seed, first_id, count) exactly (sumsq up to rounding).  The
 * graphs and plans must outlive the batch; all queries run on device 0.
 * Output: one zg_sums per query and channel, in query order (a
 * two-channel query takes two records); zg_batch_records() gives the
 * total.  count <= 2^31 per query. */
typedef struct {
    zg_graph *graph;
    const zg_plan *plan;
    uint64_t seed, first_id, count;
} zg_query;
typedef struct {
    int block_size;             /* threads per workgroup (64..256, multiple of 64; 0 = 256) */
    int blocks_per_cu;          /* persistent workgroups per CU (0 = 8) */
    uint32_t chunk_estimators;  /* estimator ids one wavefront claims at a time (a multiple of
                                   64 in 64..2^24; 0 = 1024); a query with more than 16384 such
                                   chunks takes proportionally longer ones */
} zg_batch_config;
typedef struct zg_batch zg_batch;
int zg_batch_create(const zg_query *queries, int nq, const zg_batch_config *config, zg_batch **out);
int zg_batch_records(const zg_batch *b);
/* replays the graph; ms (may be NULL) = device time of the replay */
int zg_batch_run(zg_batch *b, zg_sums *out, float *ms);
/* new seeds and/or first ids (arrays of nq, either may be NULL) for later runs */
int zg_batch_set_ids(zg_batch *b, const uint64_t *seeds, const uint64_t *first_ids);
/* times replays over block size x workgroups per CU x estimators per
 * wavefront claim {256, 1024, 4096, 16384} (coordinate descent, median of
 * `trials`), keeps the fastest configuration */
int zg_batch_autotune(zg_batch *b, int trials, zg_batch_config *chosen, float *best_ms);
int zg_batch_config_get(const zg_batch *b, zg_batch_config *out);
int zg_batch_destroy(zg_batch *b);
/* create + run + destroy */
int zg_run_batch(const zg_query *queries, int nq, const zg_batch_config *config, zg_sums *out, float *ms);

/* Algorithmic read count of a run: the number of dependent global reads the
 * estimators of ids [first_id, first_id+count) issue, counted in distinct
 * 128-byte lines (MI355X's DRAM read granule), for the roofline's
 * algorithmic bytes.  Same outcomes as zg_run;
 * slower (it counts with atomics).  No reference counterpart. */
int zg_count_reads(zg_graph *g, const zg_plan *p, uint64_t seed, uint64_t first_id, uint64_t count,
                   uint64_t *reads);

/* ---- exact counts (semantics of the naive_implementation programs) ----- *
 * Non-induced counts on the simple graph underlying the edge list: what = 0
 * triangles (TriangleCounting.cpp:44-72), 1 4-cliques (FourClique.cpp:44-78),
 * 2 wedges (ThreeChain.cpp:44-51), 3 three-edge paths (FourChain.cpp:46-66),
 * 4 5-stars (FiveStar.cpp:51-59), 5 5-houses (FiveHouse.cpp's count, as
 * sum_e t_e c4_e - 2 sum_e t_e (t_e - 1) with t_e and c4_e the triangles and
 * 4-cycles through edge e; needs 2m < 2^32 like the 4-cycles), 6 5-vertex
 * chains (FiveChain.cpp:45-72, simple paths of four edges), 7 4-cycles (no
 * reference program; the 5-chain and house counts need them).  Counts
 * that exceed 64 bits fail with ZG_EOVERFLOW (5-stars, 5-chains; 4-cycles are
 * summed in 128 bits and checked too, though at most C(m, 2) < 2^61 of them
 * exist for m < 2^31).  Scratch: 4-cliques work in vertex ranges
 * whose bit rows fit min(32 GiB, max(1 GiB, free / 3)) (the environment
 * variable ZGRAPH_K4_SCRATCH_WORDS lowers it, in 8-byte words); 4-cycles and
 * 5-chains sort the hubs' wedges in batches of up to min(16 GiB, max(1 GiB, free / 4)) / 24 wedges
 * (houses: / 88, their wedges carrying weights)
 * (ZGRAPH_C4_SORT_WEDGES lowers it; a hub with more wedges is counted with
 * a dense counter of n entries instead; ZGRAPH_C4_WIDE_KEYS=1 and
 * ZGRAPH_C4_BATCH_VERTICES force the 64-bit sort keys and cap the vertices of
 * a batch, and ZGRAPH_HOUSE_WIDE_VALUES=1 the houses' unpacked 16-byte
 * {count, weight} values, for tests). */
int zg_exact_counts(zg_graph *g, int what, uint64_t *out);
/* Degree moments sum d^2 and sum d^3 over all vertices (ELP features). */
int zg_graph_degree_moments(zg_graph *g, double *sum_d2, double *sum_d3);

/* ---- measurement helper (no reference counterpart) ---------------------- *
 * Achievable random-gather bandwidth: `loads` random 8-byte reads from a
 * fresh device buffer of `buffer_bytes`, each lane running `chains` (1, 2, 4)
 * dependent-load chains, launched like zg_run (block_size, blocks_per_cu; 0 =
 * defaults).  *gbs counts one 128-byte line per read. */
int zg_random_gather_bw(uint64_t buffer_bytes, int chains, uint64_t loads, int block_size, int blocks_per_cu,
                        double *gbs, double *ms);
/* Achievable rate of the triangle estimator's access chain: `estimators`
 * chains of three dependent random reads -- a 32-byte record (two 16-byte
 * loads) of an m-record array, an 8-byte entry of a 2m-entry array at an
 * index taken from the record, a 16-byte header of one of m/4 128-byte
 * buckets at an index taken from the entry -- over fresh buffers of the
 * graph's array sizes, no estimator logic.  *gbs counts 3 lines of 128 bytes
 * per chain: the denominator matched to an estimator kernel's line traffic. */
int zg_random_chain_bw(uint64_t m, uint64_t estimators, int block_size, int blocks_per_cu, double *gbs,
                       double *ms);
/* Random-line read rate by load shape: `lines` random 128-byte lines of a
 * fresh buffer of `buffer_bytes`, one dependent chain per reader.  With
 * cooperative = 0 every lane reads `loads_per_line` (1..8) consecutive 16-byte
 * pieces of its own line, one load instruction each; with cooperative = 1
 * groups of 8 lanes read one line, one 16-byte piece per lane, in a single
 * instruction (loads_per_line is then ignored).  *gbs counts 128 bytes per
 * line.  Shows what each extra load instruction per random line costs (each
 * pays its own address translation) -- the measurement behind the edge-set
 * hash's one-load header. */
int zg_random_line_bw(uint64_t buffer_bytes, int loads_per_line, int cooperative, uint64_t lines, int block_size,
                      int blocks_per_cu, double *gbs, double *ms);
/* Random lines confined to regions: `lines` random 16-byte reads, one per
 * 128-byte line, each address taken from the previous line, inside regions of
 * region_bytes of a buffer_bytes buffer; workgroup b reads region
 * (b mod 8 + 8 * phase) when xcd_matched (the workgroups of one XCD share a
 * region: L2-resident when it fits), ((b / 8) mod 8 + 8 * phase) otherwise,
 * with phase = round / rounds_per_phase.  *gbs counts 128 bytes per line.
 * Measurement helper (the reuse each cache level would give the estimator
 * reads); no reference counterpart. */
int zg_random_region_bw(uint64_t buffer_bytes, uint64_t region_bytes, int xcd_matched, uint32_t rounds_per_phase,
                        uint64_t lines, int block_size, int blocks_per_cu, double *gbs, double *ms);

/* ---- error-latency profile (src/ErrorLatencyProfile.cpp) --------------- */
/* ErrorProfile<E>::get_num_estimators (:71-207) on the device: subsample at
 * rate 0.1, 2^14 pilot batches of 16 estimators plus 2^23 more, EMA error
 * curve, through-origin regression K, N = bound(eps, 1-conf, a, b, f, Delta, m) * K^2.
 * a = 1 (:449-452), b = zg_plan_cond_calls.  Writes N (rounded down), K and the
 * rough estimate f.  `pilot_scale` divides the pilot sizes by 2^pilot_scale (0 = reference sizes). */
typedef struct {
    double num_estimators;   /* N before truncation */
    double K;                /* regression slope */
    double rough_estimate;   /* F / 0.1^(a+b) */
    double sampled_estimate; /* F on the subsampled graph */
    uint64_t sampled_m, sampled_max_degree;
    double elapsed_ms;
} zg_elp_result;
int zg_elp_regression(zg_graph *g, const zg_plan *p, double eps, double conf, uint64_t seed,
                      int pilot_scale, zg_elp_result *out);
/* The bound alone (ErrorProfile::calculate_bound, :51-59). */
double zg_elp_bound(double eps, double delta, double a, double b, double est, uint64_t max_degree,
                    uint64_t num_edges);
/* UnbiasedLinearRegression::solve (src/Regression.cpp:17-28) on host arrays. */
long double zg_regression_k(const long double *x, const long double *y, unsigned len);
/* The arithmetic of get_num_estimators (:127-200) given pilot running means. */
double zg_elp_from_pilots(const long double *est, uint64_t npts, long double F, double a, double b,
                          uint64_t samp_maxdeg, uint64_t samp_m, uint64_t maxdeg, uint64_t m,
                          double eps, double conf, long double *K_out);

/* ---- learned error-latency profile network ----------------------------- *
 * No reference counterpart: a 16-64-64-64-1 ReLU MLP whose prediction (the
 * log-ratio of the estimator's coefficient of variation to a pilot's) sizes
 * the estimator budget in place of the regression fit of
 * src/ErrorLatencyProfile.cpp:127-200.  BF16 operands on MFMA, FP32
 * accumulation, FP32 master weights, Adam.  Host pointers; X is [B][16]
 * row-major fp32, targets [B].  Parameter order: W1[16][64], b1[64],
 * W2[64][64], b2, W3[64][64], b3, w4[64], b4 (row-major, input-major). */
typedef struct zg_mlp zg_mlp;
int zg_mlp_num_params(void);
int zg_mlp_input_dim(void);
int zg_mlp_create(uint64_t seed, zg_mlp **out);
int zg_mlp_destroy(zg_mlp *m);
int zg_mlp_get_params(const zg_mlp *m, float *params);
int zg_mlp_set_params(zg_mlp *m, const float *params);
int zg_mlp_forward(zg_mlp *m, const float *X, int B, float *y);
/* gradient of the mean squared error over the B rows (and the loss) */
int zg_mlp_gradient(zg_mlp *m, const float *X, const float *T, int B, float *grad, float *loss);
/* `steps` Adam steps on minibatches of `batch` rows drawn from the N rows */
int zg_mlp_train(zg_mlp *m, const float *X, const float *T, int N, int steps, int batch, float lr, uint64_t seed,
                 float *final_loss);

#ifdef __cplusplus
}
#endif
#endif
