// zg_kernels.hip -- the estimator loop of ZGraphInstance::calculate
// (src/ZGraphInstance.cpp:256-297) as one HIP kernel per plan: one estimator
// per lane, grid-stride over estimator ids, exact integer reduction of the
// values (32-bit digit sums), wave -> workgroup -> grid, all deterministic.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "zg_internal.h"

namespace zg {

// one estimator's outputs (up to two channels)
struct Est {
    U128 v[2];
    uint8_t st[2];     // 0 fail, 1 success, 3 value overflowed 128 bits
    uint8_t den[2];    // log2 of the denominator (number of c_i == 2)
    uint32_t ncond;
    uint32_t e0;
};

__device__ __forceinline__ void set_out(Est &o, int ch, const U128 &v, bool ov, int den)
{
    o.v[ch] = v;
    o.st[ch] = ov ? 3 : 1;
    o.den[ch] = (uint8_t)den;
}

// pick the r-th valid entry over lists V[0..nv) starting at pos[] (the
// selection walk of core_conditional_sample_edge, src/ZGraphInstance.cpp:186-208)
template <int MAXV>
__device__ __forceinline__ bool cond_sample(const GView &g, Rng &rng, const VI (&V)[MAXV],
                                            const uint32_t (&pos)[MAXV], int nv, uint64_t &s, SE &e)
{
    uint64_t sum = 0;
#pragma unroll
    for (int i = 0; i < MAXV; ++i)
        if (i < nv) sum += V[i].deg - pos[i];
    s = sum;
    if (sum == 0) return false;
    uint64_t r = rng.uniform(sum);
    uint64_t idx = 0;
    uint32_t own = 0, p = 0;
    bool done = false;
#pragma unroll
    for (int i = 0; i < MAXV; ++i) {
        if (i < nv && !done) {
            const uint64_t cnt = V[i].deg - pos[i];
            if (r < cnt) {
                idx = (uint64_t)V[i].off + pos[i] + r;
                own = V[i].v;
                p = pos[i] + (uint32_t)r;
                done = true;
            } else {
                r -= cnt;
            }
        }
    }
    const uint64_t ent = g.adj[idx];
    count_reads(g, 1);
    e.rank = (uint32_t)(ent >> 32);
    e.dst = (uint32_t)ent;
    e.src = own;
    e.pos = p;
    return true;
}

// canonicality_check (src/SamplerGenerator.cpp:286-303,
// applications/ChainMining.cpp:23-40) over the sampled list
template <int N>
__device__ __forceinline__ bool canonical(const uint32_t (&sr)[N], const uint32_t (&ss)[N],
                                          const uint32_t (&sd)[N], int ns, const SE &e)
{
    if (ns <= 1) return true;
    bool found = false;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (i < ns) {
            if (!found) {
                if (e.src == ss[i] || e.src == sd[i] || e.dst == ss[i] || e.dst == sd[i]) found = true;
            } else if (sr[i] >= e.rank) {
                return false;
            }
        }
    }
    return true;
}

// ---- Triangle / ThreeChain / ThreeMotif ---------------------------------
// MODE 0: triangle (applications/Triangle.cpp:48-73)
// MODE 1: 3-chain (applications/ThreeChain.cpp:22-37)
// MODE 2: 3-motif, channel 0 = 3-chain, channel 1 = triangle on the same
//         wedge (the cached sub-pattern reuse of applications/ThreeMotif.cpp:48-86)
template <int MODE>
__device__ __forceinline__ void plan_wedge(const GView &g, Rng &rng, Est &o)
{
    if (g.m == 0) return;
    const uint32_t r0 = (uint32_t)rng.uniform(g.m);
    o.e0 = r0;
    if (MODE == 1 && g.simple) {
        // 3-chain on a graph without parallel edges: the sampled entry's neighbour is
        // neither endpoint of e0 (that would be a second copy of e0 or a self-loop), so
        // the denominator is 1 and only the sizes s = ca + cb are needed -- the record's
        // first half (one 16-byte load); the index draw keeps the stream's words
        const uint4 dp = load_edge_degpos(g, r0);   // s_deg, s_pos, d_deg, d_pos
        const uint64_t s = (uint64_t)(dp.x - dp.y - 1) + (dp.z - dp.w - 1);
        o.ncond = 1;
        if (s == 0) return;
        (void)rng.uniform(s);
        U128 val{g.m, 0};
        const bool ov = mul_u128(val, s);
        set_out(o, 0, val, ov, 0);
        return;
    }
    const ERec E = load_edge(g, r0);
    const VI A = src_vi(E), B = dst_vi(E);
    const uint32_t pa = E.s_pos + 1, pb = E.d_pos + 1;   // lower_bound(r0 + 1) in both endpoint lists
    const uint32_t ca = A.deg - pa, cb = B.deg - pb;
    const uint64_t s = (uint64_t)ca + cb;
    o.ncond = 1;
    if (s == 0) return;
    const uint64_t r = rng.uniform(s);
    U128 val{g.m, 0};
    const bool ov = mul_u128(val, s);
    const bool inA = r < ca;
    const uint64_t idx = inA ? (uint64_t)A.off + pa + r : (uint64_t)B.off + pb + (r - ca);
    const uint64_t ent = g.adj[idx];
    count_reads(g, 1);
    const uint32_t r1 = (uint32_t)(ent >> 32), w = (uint32_t)ent, own = inA ? E.src : E.dst;
    const int den = (w == E.src || w == E.dst) ? 1 : 0;
    if (MODE != 0) set_out(o, 0, val, ov, den);
    if (MODE == 1) return;
    // third edge of Triangle(e0, e1) (applications/Triangle.cpp:17-40)
    uint32_t tx, ty;
    if (E.src == own) { tx = E.dst; ty = w; }
    else if (E.src == w) { tx = E.dst; ty = own; }
    else if (E.dst == own) { tx = E.src; ty = w; }
    else { tx = E.src; ty = own; }
    const bool is_e0 = (tx == E.src && ty == E.dst) || (tx == E.dst && ty == E.src);
    const bool is_e1 = (tx == own && ty == w) || (tx == w && ty == own);
    bool closed;
    if (is_e0 || is_e1) {
        closed = true;     // nothing left to wait for
    } else if (tx == ty) {
        closed = false;
    } else {
        closed = close_pair(g, tx, ty, r1 + 1, nullptr);
    }
    if (closed) set_out(o, MODE == 2 ? 1 : 0, val, ov, den);
}

// ---- ChainMining(order) (applications/ChainMining.cpp:45-84) -------------
__device__ __forceinline__ void plan_chain(const GView &g, const PlanView &P, Rng &rng, Est &o)
{
    constexpr int MAXE = 8;
    if (g.m == 0) return;
    const uint32_t r0 = (uint32_t)rng.uniform(g.m);
    o.e0 = r0;
    const ERec E = load_edge(g, r0);
    VI ends[2] = {src_vi(E), dst_vi(E)};
    uint32_t pos[2] = {E.s_pos + 1, E.d_pos + 1};
    uint32_t sr[MAXE], ss[MAXE], sd[MAXE], vs[MAXE + 1];
    sr[0] = r0; ss[0] = E.src; sd[0] = E.dst;
    vs[0] = E.src; vs[1] = E.dst;
    int ns = 1, nv = 2;
    U128 val{g.m, 0};
    bool ov = false;
    int den = 0;
    for (int i = 0; i < P.order - 2; ++i) {
        o.ncond++;
        uint64_t s;
        SE e;
        if (!cond_sample<2>(g, rng, ends, pos, 2, s, e)) return;
        den += (e.dst == ends[0].v || e.dst == ends[1].v);
        ov |= mul_u128(val, s);
        if (!canonical<MAXE>(sr, ss, sd, ns, e)) return;
        if (ns < MAXE) { sr[ns] = e.rank; ss[ns] = e.src; sd[ns] = e.dst; }
        ++ns;
        bool in = false;
#pragma unroll
        for (int j = 0; j < MAXE + 1; ++j) in |= (j < nv && vs[j] == e.dst);
        if (!in) { if (nv < MAXE + 1) vs[nv] = e.dst; ++nv; }
        if (nv != i + 3) return;
        // the new end's lists and e's position in them; rank(e) > r0 bounds the search
        uint32_t pe;
        const VI nvi = end_vi(load_edge(g, e.rank), e.dst, pe);
        const uint32_t np = lb_rank(g, nvi, 0, pe, r0 + 1, 0, e.rank);
        if (ends[0].v == e.src) { ends[0] = nvi; pos[0] = np; } else { ends[1] = nvi; pos[1] = np; }
    }
    set_out(o, 0, val, ov, den);
}

// ---- FourChain (applications/FourChain.cpp:24-124) ------------------------
__device__ __forceinline__ uint32_t fc_far(const ERec &E, const SE &e1)
{
    if (E.src == e1.src) return e1.dst;
    if (E.src == e1.dst) return e1.src;
    if (E.dst == e1.src) return e1.dst;
    return e1.src;
}

__device__ __forceinline__ void plan_four_chain(const GView &g, Rng &rng, Est &o)
{
    const uint32_t method = 1 + (uint32_t)rng.uniform(3);   // sample_interger(1, 3)
    if (g.m == 0) return;
    const uint32_t r0 = (uint32_t)rng.uniform(g.m);
    o.e0 = r0;
    const ERec E = load_edge(g, r0);
    const VI A = src_vi(E), B = dst_vi(E);
    U128 val{g.m, 0};
    bool ov = mul_u128(val, 3);
    int den = 0;
    uint64_t s;
    SE e1, e2;
    if (method != 3) {
        VI V[2] = {A, B};
        uint32_t pos[2] = {E.s_pos + 1, E.d_pos + 1};
        o.ncond = 1;
        if (!cond_sample<2>(g, rng, V, pos, 2, s, e1)) return;
        den += (e1.dst == E.src || e1.dst == E.dst);
        ov |= mul_u128(val, s);
        const uint32_t far = fc_far(E, e1);
        VI F[1];
        uint32_t pf;   // position of rank(e1) in far's list
        if (far == e1.src) { F[0] = e1.src == E.src ? A : B; pf = e1.pos; }
        else F[0] = end_vi(load_edge(g, e1.rank), far, pf);
        uint32_t fp[1];
        if (method == 1)   // edges {e0, e1}: lower_bound(r1 + 1) is right after e1
            fp[0] = pf + 1;
        else               // edges {e0}: lower_bound(r0 + 1), at most pf
            fp[0] = lb_rank(g, F[0], 0, pf, r0 + 1, 0, e1.rank);
        o.ncond = 2;
        if (!cond_sample<1>(g, rng, F, fp, 1, s, e2)) return;
        den += (e2.dst == far);
        ov |= mul_u128(val, s);
        if (e2.src == E.src || e2.src == E.dst || e2.dst == E.src || e2.dst == E.dst) return;
        if (method == 2 && e2.rank >= e1.rank) return;
    } else {
        VI V1[1] = {A};
        uint32_t p1[1] = {E.s_pos + 1};
        o.ncond = 1;
        if (!cond_sample<1>(g, rng, V1, p1, 1, s, e1)) return;
        den += (e1.dst == E.src);
        ov |= mul_u128(val, s);
        VI V2[1] = {B};
        uint32_t p2[1] = {E.d_pos + 1};
        o.ncond = 2;
        if (!cond_sample<1>(g, rng, V2, p2, 1, s, e2)) return;
        den += (e2.dst == E.dst);
        ov |= mul_u128(val, s);
        if (e2.src == e1.src || e2.src == e1.dst || e2.dst == e1.src || e2.dst == e1.dst) return;
    }
    set_out(o, 0, val, ov, den);
}

// ---- SamplerGenerator<Empty>::estimate (src/SamplerGenerator.cpp:378-464) --
// The isomorphism loop over unique relabelings is a lookup in a bit table
// indexed by the mask of relabeled vertex pairs: bit set iff every sampled
// pair lies in some relabeling of the pattern (Pattern::is_automorphic).
template <int MAXV>
__device__ __forceinline__ void plan_pattern(const GView &g, const PlanView &P, Rng &rng, Est &o)
{
    constexpr int MAXE = MAXV * (MAXV - 1) / 2;   // a pattern on MAXV vertices has at most this many edges
    constexpr int NW = (MAXE + 4) / 5;            // 5 edges per word, 6 bits each
    if (g.m == 0) return;
    const uint32_t r0 = (uint32_t)rng.uniform(g.m);
    o.e0 = r0;
    const ERec E = load_edge(g, r0);
    VI V[MAXV];
    uint32_t pos[MAXV];
#pragma unroll
    for (int i = 0; i < MAXV; ++i) { V[i] = VI{0, 0, 0}; pos[i] = 0; }
    V[0] = src_vi(E); V[1] = dst_vi(E);
    pos[0] = E.s_pos + 1; pos[1] = E.d_pos + 1;
    int nv = 2;
    // sampled edges in sampling order: ranks, and endpoints as indices into V
    // (3 bits each) -- vertex ids are distinct, so index equality is id equality
    uint32_t sr[MAXE], ends[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) ends[w] = 0;
    sr[0] = r0;
    ends[0] = 0u | (1u << 3);
    int ns = 1;
    U128 val{g.m, 0};
    bool ov = false;
    int den = 0;
    for (int i = 1; i < P.pe; ++i) {
        o.ncond++;
        uint64_t s;
        SE e;
        if (!cond_sample<MAXV>(g, rng, V, pos, nv, s, e)) return;
        uint32_t si = 0, di = (uint32_t)nv;
#pragma unroll
        for (int j = 0; j < MAXV; ++j) {
            if (j < nv && V[j].v == e.src) si = (uint32_t)j;
            if (j < nv && V[j].v == e.dst) di = (uint32_t)j;
        }
        const bool in = di < (uint32_t)nv;
        den += in;
        ov |= mul_u128(val, s);
        // canonicality_check (src/SamplerGenerator.cpp:286-303): from the first earlier
        // edge sharing a vertex with e on, every edge must rank below e
        if (ns > 1) {
            bool found = false, ok = true;
#pragma unroll
            for (int q = 0; q < MAXE; ++q) {
                if (q < ns) {
                    const uint32_t ab = (ends[q / 5] >> (6 * (q % 5))) & 63u;
                    const uint32_t a = ab & 7u, b = ab >> 3;
                    if (!found) found = a == si || a == di || b == si || b == di;
                    else if (sr[q] >= e.rank) ok = false;
                }
            }
            if (!ok) return;
        }
#pragma unroll
        for (int q = 0; q < MAXE; ++q)
            if (q == ns) {
                sr[q] = e.rank;
                ends[q / 5] |= (si | (di << 3)) << (6 * (q % 5));
            }
        ++ns;
        if (!in) {
            if (nv >= P.pk || nv >= MAXV) return;   // |V| would exceed k
            uint32_t pe;
            const VI nvi = end_vi(load_edge(g, e.rank), e.dst, pe);
            const uint32_t np = lb_rank(g, nvi, 0, pe, r0 + 1, 0, e.rank);   // rank(e) > r0
#pragma unroll
            for (int j = 0; j < MAXV; ++j)
                if (j == nv) { V[j] = nvi; pos[j] = np; }
            ++nv;
        }
    }
    if (nv != P.pk) return;
    // relabel: a vertex's new label is its position among the sorted ids
    uint32_t rk[MAXV];
#pragma unroll
    for (int j = 0; j < MAXV; ++j) {
        uint32_t c = 0;
#pragma unroll
        for (int l = 0; l < MAXV; ++l) c += V[l].v < V[j].v;
        rk[j] = c;
    }
    uint32_t mask = 0;
#pragma unroll
    for (int q = 0; q < MAXE; ++q) {
        if (q < ns) {
            const uint32_t ab = (ends[q / 5] >> (6 * (q % 5))) & 63u;
            uint32_t a = 0, b = 0;
#pragma unroll
            for (int j = 0; j < MAXV; ++j) {
                if ((ab & 7u) == (uint32_t)j) a = rk[j];
                if ((ab >> 3) == (uint32_t)j) b = rk[j];
            }
            const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
            mask |= 1u << (hi * (hi - 1) / 2 + lo);
        }
    }
    count_reads(g, 1);
    if ((P.accept[mask >> 5] >> (mask & 31)) & 1u) set_out(o, 0, val, ov, den);
}

// ---- 4-clique: three spanning shapes + closing checks --------------------
// For a clique with ranks r1 < ... < r6: shape A (r1, r2 adjacent, r3 reaches
// the 4th vertex), shape B (r1, r2, r3 a triangle), shape C (r1, r2 disjoint).
// Each shape has one draw sequence of probability 1/(m s1 s2), so the sum of
// the three shape values is unbiased for every rank order.  (The reference's
// applications/FourClique.cpp:41-76 covers shape A only.)

// The first stage's state: e0 = {a, b} of rank r0 (a, b's lists from position
// pa0, pb0 on hold the ranks above r0), e1 = {o, w} of rank r1 drawn from
// o's list at position po1 (o = a when oa), s1 the valid entries it was drawn
// from, {ob, w} closed at rank rt > r1, and W with rank r1 at position pw1.
struct K4S {
    VI A, B, W;
    uint32_t r0, r1, rt, pa0, pb0, po1, pw1, oa;
    uint64_t s1;
};

// e0, e1 and the closing check of the triangle edge {ob, w} above r1, which
// every shape needs; true when it closes.  TRI: channel 0 = the triangle
// plan's outcome for the same id.  Both plans draw the first edge and the
// conditional sample identically, and for w outside {a, b} the triangle
// plan's closing check is this one, with value m * s1 (for w in {a, b}
// neither succeeds).
template <bool TRI>
__device__ __forceinline__ bool k4_first(const GView &g, Rng &rng, Est &o, K4S &k)
{
    if (g.m == 0) return false;
    const uint32_t r0 = (uint32_t)rng.uniform(g.m);
    o.e0 = r0;
    const ERec E = load_edge(g, r0);
    const uint32_t a = E.src, b = E.dst;
    k.A = src_vi(E);
    k.B = dst_vi(E);
    k.r0 = r0;
    k.pa0 = E.s_pos + 1;
    k.pb0 = E.d_pos + 1;
    SE e1;
    {
        VI V[2] = {k.A, k.B};
        uint32_t pos[2] = {k.pa0, k.pb0};
        o.ncond = 1;
        if (!cond_sample<2>(g, rng, V, pos, 2, k.s1, e1)) return false;
    }
    const uint32_t w = e1.dst;
    if (w == a || w == b) return false;
    k.oa = e1.src == a ? 1u : 0u;
    k.r1 = e1.rank;
    if (!close_pair(g, k.oa ? b : a, w, k.r1 + 1, &k.rt)) return false;
    if constexpr (TRI) {
        U128 tv{g.m, 0};
        const bool tov = mul_u128(tv, k.s1);
        set_out(o, 0, tv, tov, 0);
    }
    k.po1 = e1.pos;                              // rank r1 in o's list
    k.W = end_vi(load_edge(g, k.r1), w, k.pw1);  // and in w's
    return true;
}

// What every completion of a closed first stage shares: the three shapes'
// starting positions in the lists of a, b, w (shape A: ranks above r1; B:
// above the triangle edge's rank rt; C: above r0), found once per estimator
// by the searches of the shapes (rank r1 in ob's list, rt in o's, r0 in w's)
// and the triangle edge's record.
struct K4Q {
    VI A, B, W;
    uint32_t r0, r1, oa;
    uint32_t pA[3], pB[3], pC[3];
    uint64_t s1;
};

__device__ __forceinline__ void k4_prepare(const GView &g, const K4S &k, K4Q &q)
{
    const bool oa = k.oa != 0;
    const uint32_t ob = oa ? k.B.v : k.A.v, w = k.W.v;
    const VI O = oa ? k.A : k.B, OB = oa ? k.B : k.A;
    const uint32_t pob0 = oa ? k.pb0 : k.pa0;   // lower_bound(r0 + 1) in ob's list
    q.A = k.A;
    q.B = k.B;
    q.W = k.W;
    q.r0 = k.r0;
    q.r1 = k.r1;
    q.oa = k.oa;
    q.s1 = k.s1;
    const uint32_t pob = lb_rank(g, OB, pob0, OB.deg, k.r1 + 1, (uint64_t)k.r0 + 1, g.m);
    q.pA[0] = oa ? k.po1 + 1 : pob;
    q.pA[1] = oa ? pob : k.po1 + 1;
    q.pA[2] = k.pw1 + 1;
    const ERec ET = load_edge(g, k.rt);
    const uint32_t pob_t = (ET.src == ob ? ET.s_pos : ET.d_pos) + 1;
    const uint32_t pw_t = (ET.src == w ? ET.s_pos : ET.d_pos) + 1;
    const uint32_t po = lb_rank(g, O, k.po1 + 1, O.deg, k.rt + 1, (uint64_t)k.r1 + 1, g.m);
    q.pB[0] = oa ? po : pob_t;
    q.pB[1] = oa ? pob_t : po;
    q.pB[2] = pw_t;
    q.pC[0] = k.pa0;
    q.pC[1] = k.pb0;
    q.pC[2] = lb_rank(g, k.W, 0, k.pw1, k.r0 + 1, 0, k.r1);
}

// One completion: the three shapes from their prepared positions, one
// conditional sample each from rng, and their closing checks; returns the sum
// of their values m s1 s2, and sets ov when one of them passed 128 bits.
__device__ __forceinline__ U128 k4_complete(const GView &g, Rng &rng, const K4Q &q, bool &ov)
{
    const VI A = q.A, B = q.B, W = q.W;
    const uint32_t a = A.v, b = B.v, w = W.v;
    const uint32_t ob = q.oa ? b : a;
    U128 base{g.m, 0};
    const bool base_ov = mul_u128(base, q.s1);
    U128 total{0, 0};
    VI V[3] = {A, B, W};
    // shape A: e2 above r1 from {a, b, w} to a new vertex d; close the rest above r2
    {
        uint32_t pos[3] = {q.pA[0], q.pA[1], q.pA[2]};
        uint64_t s2;
        SE e2;
        if (cond_sample<3>(g, rng, V, pos, 3, s2, e2)) {
            const uint32_t d = e2.dst;
            if (d != a && d != b && d != w) {
                const uint32_t mv = e2.rank + 1;
                bool ok = close_pair(g, ob, w, mv, nullptr);
                if (ok && e2.src != a) ok = close_pair(g, a, d, mv, nullptr);
                if (ok && e2.src != b) ok = close_pair(g, b, d, mv, nullptr);
                if (ok && e2.src != w) ok = close_pair(g, w, d, mv, nullptr);
                if (ok) {
                    U128 v = base;
                    const bool o2 = mul_u128(v, s2);
                    add_u128(total, v);
                    ov |= base_ov | o2;
                }
            }
        }
    }
    // shape B: triangle {e0, e1, t}; e2 above rank(t) to a new vertex d; close the rest above r2
    {
        uint32_t pos[3] = {q.pB[0], q.pB[1], q.pB[2]};
        uint64_t s3;
        SE e3;
        if (cond_sample<3>(g, rng, V, pos, 3, s3, e3)) {
            const uint32_t d = e3.dst;
            if (d != a && d != b && d != w) {
                const uint32_t mv = e3.rank + 1;
                bool ok = true;
                if (e3.src != a) ok = close_pair(g, a, d, mv, nullptr);
                if (ok && e3.src != b) ok = close_pair(g, b, d, mv, nullptr);
                if (ok && e3.src != w) ok = close_pair(g, w, d, mv, nullptr);
                if (ok) {
                    U128 v = base;
                    const bool o3 = mul_u128(v, s3);
                    add_u128(total, v);
                    ov |= base_ov | o3;
                }
            }
        }
    }
    // shape C: e2 above r0 from w to a new vertex d, below r1; close the rest above r1
    {
        uint32_t pos[3] = {q.pC[0], q.pC[1], q.pC[2]};
        uint64_t s4;
        SE e4;
        if (cond_sample<3>(g, rng, V, pos, 3, s4, e4)) {
            const uint32_t d = e4.dst;
            if (e4.src == w && d != a && d != b && d != w && e4.rank < q.r1) {
                const uint32_t mv = q.r1 + 1;
                if (close_pair(g, ob, w, mv, nullptr) && close_pair(g, a, d, mv, nullptr) &&
                    close_pair(g, b, d, mv, nullptr)) {
                    U128 v = base;
                    const bool o4 = mul_u128(v, s4);
                    add_u128(total, v);
                    ov |= base_ov | o4;
                }
            }
        }
    }
    return total;
}

// The 4-clique plan (TRI: and the triangle channel) as one body: the steps of
// k4_first, k4_prepare and k4_complete on the estimator's own stream.  Written
// out rather than composed from them: the composed form compiled to 85 VGPRs
// (5 waves/SIMD) and ran 4 % slower on the 69 M-edge graph (16.9 against 16.3
// ms per 2^28 estimators, alternating runs of both libraries).
template <bool TRI>
__device__ __forceinline__ void plan_k4_close(const GView &g, Rng &rng, Est &o)
{
    if (g.m == 0) return;
    const uint32_t r0 = (uint32_t)rng.uniform(g.m);
    o.e0 = r0;
    const ERec E = load_edge(g, r0);
    const uint32_t a = E.src, b = E.dst;
    const VI A = src_vi(E), B = dst_vi(E);
    const uint32_t pa0 = E.s_pos + 1, pb0 = E.d_pos + 1;
    uint64_t s1;
    SE e1;
    {
        VI V[2] = {A, B};
        uint32_t pos[2] = {pa0, pb0};
        o.ncond = 1;
        if (!cond_sample<2>(g, rng, V, pos, 2, s1, e1)) return;
    }
    const uint32_t w = e1.dst, r1 = e1.rank;
    if (w == a || w == b) return;
    const bool oa = e1.src == a;
    const uint32_t ob = oa ? b : a;
    // every shape needs the triangle edge {ob, w} above r1: test it first
    uint32_t rt;
    if (!close_pair(g, ob, w, r1 + 1, &rt)) return;
    if constexpr (TRI) {
        U128 tv{g.m, 0};
        const bool tov = mul_u128(tv, s1);
        set_out(o, 0, tv, tov, 0);
    }
    const VI O = oa ? A : B, OB = oa ? B : A;
    const uint32_t po1 = e1.pos;                 // rank r1 in o's list
    uint32_t pw1;                                // rank r1 in w's list
    const VI W = end_vi(load_edge(g, r1), w, pw1);
    const uint32_t pob0 = oa ? pb0 : pa0;        // lower_bound(r0 + 1) in ob's list
    U128 base{g.m, 0};
    const bool base_ov = mul_u128(base, s1);
    U128 total{0, 0};
    bool ov = false;

    // shape A: e2 above r1 from {a, b, w} to a new vertex d; close the rest above r2
    {
        const uint32_t pob = lb_rank(g, OB, pob0, OB.deg, r1 + 1, (uint64_t)r0 + 1, g.m);
        VI V[3] = {A, B, W};
        uint32_t pos[3] = {oa ? po1 + 1 : pob, oa ? pob : po1 + 1, pw1 + 1};
        uint64_t s2;
        SE e2;
        o.ncond++;
        if (cond_sample<3>(g, rng, V, pos, 3, s2, e2)) {
            const uint32_t d = e2.dst;
            if (d != a && d != b && d != w) {
                const uint32_t mv = e2.rank + 1;
                bool ok = close_pair(g, ob, w, mv, nullptr);
                if (ok && e2.src != a) ok = close_pair(g, a, d, mv, nullptr);
                if (ok && e2.src != b) ok = close_pair(g, b, d, mv, nullptr);
                if (ok && e2.src != w) ok = close_pair(g, w, d, mv, nullptr);
                if (ok) {
                    U128 v = base;
                    const bool o2 = mul_u128(v, s2);
                    add_u128(total, v);
                    ov |= base_ov | o2;
                }
            }
        }
    }
    // shape B: triangle {e0, e1, t}; e2 above rank(t) to a new vertex d; close the rest above r2
    {
        {
            const ERec ET = load_edge(g, rt);
            const uint32_t pob_t = (ET.src == ob ? ET.s_pos : ET.d_pos) + 1;
            const uint32_t pw_t = (ET.src == w ? ET.s_pos : ET.d_pos) + 1;
            const uint32_t po = lb_rank(g, O, po1 + 1, O.deg, rt + 1, (uint64_t)r1 + 1, g.m);
            VI V[3] = {A, B, W};
            uint32_t pos[3] = {oa ? po : pob_t, oa ? pob_t : po, pw_t};
            uint64_t s3;
            SE e3;
            o.ncond++;
            if (cond_sample<3>(g, rng, V, pos, 3, s3, e3)) {
                const uint32_t d = e3.dst;
                if (d != a && d != b && d != w) {
                    const uint32_t mv = e3.rank + 1;
                    bool ok = true;
                    if (e3.src != a) ok = close_pair(g, a, d, mv, nullptr);
                    if (ok && e3.src != b) ok = close_pair(g, b, d, mv, nullptr);
                    if (ok && e3.src != w) ok = close_pair(g, w, d, mv, nullptr);
                    if (ok) {
                        U128 v = base;
                        const bool o3 = mul_u128(v, s3);
                        add_u128(total, v);
                        ov |= base_ov | o3;
                    }
                }
            }
        }
    }
    // shape C: e2 above r0 from w to a new vertex d, below r1; close the rest above r1
    {
        const uint32_t pw0 = lb_rank(g, W, 0, pw1, r0 + 1, 0, r1);
        VI V[3] = {A, B, W};
        uint32_t pos[3] = {pa0, pb0, pw0};
        uint64_t s4;
        SE e4;
        o.ncond++;
        if (cond_sample<3>(g, rng, V, pos, 3, s4, e4)) {
            const uint32_t d = e4.dst;
            if (e4.src == w && d != a && d != b && d != w && e4.rank < r1) {
                const uint32_t mv = r1 + 1;
                if (close_pair(g, ob, w, mv, nullptr) && close_pair(g, a, d, mv, nullptr) &&
                    close_pair(g, b, d, mv, nullptr)) {
                    U128 v = base;
                    const bool o4 = mul_u128(v, s4);
                    add_u128(total, v);
                    ov |= base_ov | o4;
                }
            }
        }
    }
    if (total.lo | total.hi) set_out(o, TRI ? 1 : 0, total, ov, 0);
}

// ---- split 4-clique plan (ZG_PLAN_K4_SPLIT) ---------------------------------
// The 4-clique plan's first stage per lane; for every lane whose triangle
// closes, the whole wavefront runs that estimator's J completions -- the
// three shapes, completion j drawing from the stream (seed, TAG_COMPLETION +
// j, id) -- lane l taking completions l, l + 64, ...  The value is the sum of
// the completions' values over J (a power of two: the channel's denominator
// exponent, P.K, or P.K1 beside the triangle channel): each completion has the
// distribution of the shapes given the first stage, so the mean is the
// 4-clique plan's, while the rare completions that make its variance are
// drawn J times per closed triangle.  One lane running its own J completions
// would hold its wavefront for J dependent chains; spread over the lanes they
// cost about one.  The searches that fix the shapes' starting positions depend
// only on the first stage, so the owning lane runs them once (k4_prepare) and
// the completions share the result.  Every lane of the wavefront calls this
// (active = the lane has an estimator); outcomes match the oracle's sequential
// completions.
__device__ __forceinline__ uint32_t lane_u32(uint32_t x, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)x, l); }
__device__ __forceinline__ VI lane_vi(const VI &x, int l) { return VI{lane_u32(x.v, l), lane_u32(x.deg, l), lane_u32(x.off, l)}; }

// TRI: channel 0 = the triangle plan's outcome (k4_first<true>), channel 1 the split 4-clique's.
template <bool TRI>
__device__ __forceinline__ void k4_split(const GView &g, const PlanView &P, uint64_t seed, uint64_t id, bool active,
                                         Est &o, uint32_t &words)
{
    o.st[0] = o.st[1] = 0;
    o.den[0] = o.den[1] = 0;
    o.v[0] = o.v[1] = U128{0, 0};
    o.ncond = 0;
    o.e0 = 0;
    Rng rng;
    rng.init(seed, TAG_ESTIMATOR, id);
    K4S k;
    const bool closed = active && k4_first<TRI>(g, rng, o, k);
    words = rng.word;
    K4Q q;
    if (closed) k4_prepare(g, k, q);   // the completions' shared searches, once per estimator
    const uint32_t lane = threadIdx.x & 63;
    const int lg = TRI ? P.K1 : P.K;
    const uint32_t J = 1u << lg;
    for (uint64_t pend = __ballot(closed); pend; pend &= pend - 1) {
        const int l = __builtin_ctzll(pend);
        K4Q t;
        t.A = lane_vi(q.A, l);
        t.B = lane_vi(q.B, l);
        t.W = lane_vi(q.W, l);
        t.r0 = lane_u32(q.r0, l);
        t.r1 = lane_u32(q.r1, l);
        t.oa = lane_u32(q.oa, l);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            t.pA[i] = lane_u32(q.pA[i], l);
            t.pB[i] = lane_u32(q.pB[i], l);
            t.pC[i] = lane_u32(q.pC[i], l);
        }
        t.s1 = ((uint64_t)lane_u32((uint32_t)(q.s1 >> 32), l) << 32) | lane_u32((uint32_t)q.s1, l);
        const uint64_t tid = ((uint64_t)lane_u32((uint32_t)(id >> 32), l) << 32) | lane_u32((uint32_t)id, l);
        U128 sum{0, 0};
        bool ov = false;
        for (uint32_t j0 = 0; j0 < J; j0 += 64) {
            if (j0 + lane < J) {
                Rng rc;
                rc.init(seed, TAG_COMPLETION + j0 + lane, tid);
                const U128 v = k4_complete(g, rc, t, ov);
                ov |= add_u128(sum, v);
            }
        }
        // the wavefront's sum in 32-bit digits (each digit sum < 2^38)
        uint64_t d0 = (uint32_t)sum.lo, d1 = sum.lo >> 32, d2 = (uint32_t)sum.hi, d3 = sum.hi >> 32;
        uint32_t ovw = ov ? 1u : 0u;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            d0 += __shfl_xor(d0, off, 64);
            d1 += __shfl_xor(d1, off, 64);
            d2 += __shfl_xor(d2, off, 64);
            d3 += __shfl_xor(d3, off, 64);
            ovw |= __shfl_xor(ovw, off, 64);
        }
        if (lane == (uint32_t)l) {
            U128 tot{d0, 0};
            bool tov = ovw != 0;
            tov |= add_u128(tot, U128{d1 << 32, d1 >> 32});
            tov |= add_u128(tot, U128{0, d2});
            tov |= add_u128(tot, U128{0, d3 << 32});
            tov |= (d3 >> 32) != 0;
            o.ncond = 1 + 3 * J;
            if (tot.lo | tot.hi) set_out(o, TRI ? 1 : 0, tot, tov, lg);
        }
    }
}

// ---- split 5-house plan (ZG_PLAN_HOUSE_SPLIT) -------------------------------
// A house (the pattern of applications/FiveHouse.cpp:18-27: roof triangle
// {0, 1, 2} on the square 1-3-4-2) has exactly one edge in both its triangle
// and its 4-cycle, the base {1, 2}.  First stage: e0 of rank r0 ~ U[m]; a = its
// endpoint of smaller degree (the source on ties), b the other; an entry ~
// U[deg a] of a's list gives the apex w, and the roof closes iff w != b and the
// pair {w, b} has copies.  For every lane whose roof closes, the wavefront runs
// that estimator's J completions (stream (seed, TAG_COMPLETION + j, id), lane
// l taking j = l, l + 64, ...): an entry ~ U[deg a] of a's list gives x, an
// entry ~ U[deg b] of b's list gives y, counting the copies of {x, y} when
// x != b, y != a, x != y and w is neither.  value = m deg(a) mult(w, b) deg(a)
// deg(b) (sum of the counts) / J: its mean is the number of houses, each once
// per choice of copies of its six edges.
struct HQ {
    VI A, B;
    uint32_t w;
};

// the first stage; on a closed roof, q and the factor m deg(a) mult(w, b) deg(a) deg(b)
__device__ __forceinline__ bool house_first(const GView &g, Rng &rng, Est &o, HQ &q, U128 &base, bool &bov)
{
    if (g.m == 0) return false;
    const uint32_t r0 = (uint32_t)rng.uniform(g.m);
    o.e0 = r0;
    const ERec E = load_edge(g, r0);
    VI A = src_vi(E), B = dst_vi(E);
    if (B.deg < A.deg) { const VI t = A; A = B; B = t; }
    o.ncond = 1;
    const uint32_t w = (uint32_t)g.adj[(uint64_t)A.off + rng.uniform(A.deg)];
    count_reads(g, 1);
    if (w == B.v) return false;
    const uint32_t mwb = close_count(g, w, B.v);
    if (!mwb) return false;
    q.A = A; q.B = B; q.w = w;
    base = U128{g.m, 0};
    bov = mul_u128(base, A.deg);
    bov |= mul_u128(base, mwb);
    bov |= mul_u128(base, A.deg);
    bov |= mul_u128(base, B.deg);
    return true;
}

__device__ __forceinline__ uint32_t house_complete(const GView &g, Rng &rc, const HQ &q)
{
    const uint32_t x = (uint32_t)g.adj[(uint64_t)q.A.off + rc.uniform(q.A.deg)];
    const uint32_t y = (uint32_t)g.adj[(uint64_t)q.B.off + rc.uniform(q.B.deg)];
    count_reads(g, 2);
    if (x == q.B.v || y == q.A.v || x == y || x == q.w || y == q.w) return 0;
    return close_count(g, x, y);
}

__device__ __forceinline__ void house_split(const GView &g, const PlanView &P, uint64_t seed, uint64_t id,
                                            bool active, Est &o, uint32_t &words)
{
    o.st[0] = o.st[1] = 0;
    o.den[0] = o.den[1] = 0;
    o.v[0] = o.v[1] = U128{0, 0};
    o.ncond = 0;
    o.e0 = 0;
    Rng rng;
    rng.init(seed, TAG_ESTIMATOR, id);
    HQ q;
    U128 base{0, 0};
    bool bov = false;
    const bool closed = active && house_first(g, rng, o, q, base, bov);
    words = rng.word;
    const uint32_t lane = threadIdx.x & 63;
    const int lg = P.K;
    const uint32_t J = 1u << lg;
    for (uint64_t pend = __ballot(closed); pend; pend &= pend - 1) {
        const int l = __builtin_ctzll(pend);
        HQ t;
        t.A = lane_vi(q.A, l);
        t.B = lane_vi(q.B, l);
        t.w = lane_u32(q.w, l);
        const uint64_t tid = ((uint64_t)lane_u32((uint32_t)(id >> 32), l) << 32) | lane_u32((uint32_t)id, l);
        uint64_t n = 0;   // at most J / 64 counts of at most 2^32 - 1 each
        for (uint32_t j0 = 0; j0 < J; j0 += 64) {
            if (j0 + lane < J) {
                Rng rc;
                rc.init(seed, TAG_COMPLETION + j0 + lane, tid);
                n += house_complete(g, rc, t);
            }
        }
        // the wavefront's sum in 32-bit digits
        uint64_t d0 = (uint32_t)n, d1 = n >> 32;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            d0 += __shfl_xor(d0, off, 64);
            d1 += __shfl_xor(d1, off, 64);
        }
        if (lane == (uint32_t)l) {
            o.ncond = 1 + 2 * J;
            const uint64_t S = d0 + (d1 << 32);   // d0 < 2^38, d1 < 2^12: S < 2^45
            if (S) {
                U128 tot = base;
                const bool tov = bov | mul_u128(tot, S);
                set_out(o, 0, tot, tov, lg);
            }
        }
    }
}

// ---- accumulation ---------------------------------------------------------
// per lane: a 192-bit running sum (three u64 limbs); per wave and beyond: sums
// of 32-bit digits, which add without carries
struct AccL {
    uint64_t lo, hi, top;
    uint32_t succ, ovf;
    double sq;
};

__device__ __forceinline__ void acc_add(AccL &a, const U128 &v, int st, int den, int K)
{
    if (st == 0) return;
    if (st != 1) { a.ovf++; return; }
    a.succ++;
    const int sh = K - den;
    const uint64_t w0 = v.lo << sh;
    const uint64_t w1 = sh ? (v.hi << sh) | (v.lo >> (64 - sh)) : v.hi;
    const uint64_t w2 = sh ? v.hi >> (64 - sh) : 0;
    const uint64_t lo = a.lo + w0;
    const uint64_t c0 = lo < a.lo;
    const uint64_t t = a.hi + w1;
    uint64_t c1 = t < a.hi;
    const uint64_t hi = t + c0;
    c1 += hi < t;
    a.lo = lo;
    a.hi = hi;
    a.top += w2 + c1;
    const double x = ((double)v.hi * 18446744073709551616.0 + (double)v.lo) * ldexp(1.0, -den);
    a.sq += x * x;
}

struct AccW {
    uint64_t dig[5];
    uint64_t succ, ovf;
    double sq;
};

__device__ __forceinline__ AccW wave_reduce(const AccL &l)
{
    AccW a;
    a.dig[0] = (uint32_t)l.lo;
    a.dig[1] = l.lo >> 32;
    a.dig[2] = (uint32_t)l.hi;
    a.dig[3] = l.hi >> 32;
    a.dig[4] = l.top;
    a.succ = l.succ;
    a.ovf = l.ovf;
    a.sq = l.sq;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int i = 0; i < 5; ++i) a.dig[i] += __shfl_xor(a.dig[i], off, 64);
        a.succ += __shfl_xor(a.succ, off, 64);
        a.ovf += __shfl_xor(a.ovf, off, 64);
        a.sq += __shfl_xor(a.sq, off, 64);
    }
    return a;
}

// reduces the workgroup's accumulators and writes NCH records at dst
template <int NCH>
__device__ __forceinline__ void block_reduce_store(AccL (&a)[NCH], Acc *__restrict__ dst)
{
    __shared__ Acc sm[NCH][16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const AccW r = wave_reduce(a[c]);
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < 5; ++i) sm[c][wv].dig[i] = r.dig[i];
            sm[c][wv].succ = r.succ;
            sm[c][wv].ovf = r.ovf;
            sm[c][wv].sq = r.sq;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            Acc t = sm[c][0];
            for (int w = 1; w < nw; ++w) {
#pragma unroll
                for (int i = 0; i < 5; ++i) t.dig[i] += sm[c][w].dig[i];
                t.succ += sm[c][w].succ;
                t.ovf += sm[c][w].ovf;
                t.sq += sm[c][w].sq;
            }
            dst[c] = t;
        }
    }
}

// kernel kinds: the plan kinds, and one per pattern size k (2..6) for ZG_PLAN_PATTERN
constexpr int KIND_PATTERN_BASE = 100;
inline int kernel_kind(const PlanView &P)
{
    return P.kind == ZG_PLAN_PATTERN ? KIND_PATTERN_BASE + P.pk : P.kind;
}

// plans whose estimators finish their completions wavefront-cooperatively
__host__ __device__ constexpr bool split_kind(int kind)
{
    return kind == ZG_PLAN_K4_SPLIT || kind == ZG_PLAN_TRIANGLE_K4_SPLIT || kind == ZG_PLAN_HOUSE_SPLIT;
}
template <int KIND>
__device__ __forceinline__ void run_split(const GView &g, const PlanView &P, uint64_t seed, uint64_t id, bool active,
                                          Est &o, uint32_t &words)
{
    if constexpr (KIND == ZG_PLAN_HOUSE_SPLIT) house_split(g, P, seed, id, active, o, words);
    else k4_split<KIND == ZG_PLAN_TRIANGLE_K4_SPLIT>(g, P, seed, id, active, o, words);
}

// one estimator: the plan body for estimator id `id`
template <int KIND>
__device__ __forceinline__ void run_one(const GView &g, const PlanView &P, Rng &rng, Est &o)
{
    o.st[0] = o.st[1] = 0;
    o.den[0] = o.den[1] = 0;
    o.v[0] = o.v[1] = U128{0, 0};
    o.ncond = 0;
    o.e0 = 0;
    if constexpr (KIND == ZG_PLAN_TRIANGLE) plan_wedge<0>(g, rng, o);
    else if constexpr (KIND == ZG_PLAN_THREE_CHAIN) plan_wedge<1>(g, rng, o);
    else if constexpr (KIND == ZG_PLAN_THREE_MOTIF) plan_wedge<2>(g, rng, o);
    else if constexpr (KIND == ZG_PLAN_CHAIN) plan_chain(g, P, rng, o);
    else if constexpr (KIND == ZG_PLAN_FOUR_CHAIN) plan_four_chain(g, rng, o);
    else if constexpr (KIND == ZG_PLAN_K4_CLOSE) plan_k4_close<false>(g, rng, o);
    else if constexpr (KIND == ZG_PLAN_TRIANGLE_K4) plan_k4_close<true>(g, rng, o);
    else if constexpr (KIND >= KIND_PATTERN_BASE + 2) plan_pattern<KIND - KIND_PATTERN_BASE>(g, P, rng, o);
}

template <int NCH>
__device__ __forceinline__ void acc_zero(AccL (&acc)[NCH])
{
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        acc[c].lo = acc[c].hi = acc[c].top = 0;
        acc[c].succ = acc[c].ovf = 0;
        acc[c].sq = 0;
    }
}

// OUT: also write every estimator's record (parity dumps, ELP pilots)
template <int KIND, int NCH, bool OUT>
__global__ __launch_bounds__(256) void k_run(GView g, PlanView P, uint64_t seed, uint64_t first,
                                             uint64_t count, zg_outcome *__restrict__ out,
                                             Acc *__restrict__ partials)
{
    AccL acc[NCH];
    acc_zero<NCH>(acc);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    auto record = [&](uint64_t i, const Est &o, uint32_t words) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            acc_add(acc[c], o.v[c], o.st[c], o.den[c], c == 0 ? P.K : P.K1);
            if (OUT) {
                zg_outcome rec;
                const bool ok = o.st[c] != 0;
                rec.num_lo = ok ? o.v[c].lo : 0;
                rec.num_hi = ok ? o.v[c].hi : 0;
                rec.e0 = o.e0;
                rec.status = o.st[c];
                rec.log2den = ok ? o.den[c] : 0;
                rec.ncond = (uint8_t)(o.ncond > 255 ? 255 : o.ncond);
                rec.nwords = (uint8_t)(words > 255 ? 255 : words);
                out[i * NCH + c] = rec;
            }
        }
    };
    if constexpr (split_kind(KIND)) {
        // wavefront-uniform trip count: every lane reaches the cooperative completions
        const uint32_t lane = threadIdx.x & 63;
        for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); b < count; b += stride) {
            const uint64_t i = b + lane;
            Est o;
            uint32_t words;
            run_split<KIND>(g, P, seed, first + i, i < count, o, words);
            if (i < count) record(i, o, words);
        }
    } else {
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
            Rng rng;
            rng.init(seed, TAG_ESTIMATOR, first + i);
            Est o;
            run_one<KIND>(g, P, rng, o);
            record(i, o, rng.word);
        }
    }
    block_reduce_store<NCH>(acc, partials + (size_t)blockIdx.x * NCH);
}

// First-edge order.  Every plan's first random choice is the rank of its first
// edge (after the method draw for FourChain), uniform over the m ranks, so a
// launch in estimator-id order reads the 32-byte edge records in random order:
// about one 128-byte DRAM line per estimator.  Sorting the launch's estimators
// by the high bits of that rank first makes the in-flight estimators share
// record lines in the caches.  The set of estimators and each one's outcome are
// unchanged; the exact sums are identical.
template <bool METHOD_FIRST>
__global__ void k_first_edge_keys(uint64_t m, uint64_t seed, uint64_t first, uint64_t count, uint32_t shift,
                                  uint32_t *__restrict__ keys, uint32_t *__restrict__ offs)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) {
        Rng rng;
        rng.init(seed, TAG_ESTIMATOR, first + i);
        if (METHOD_FIRST) (void)rng.uniform(3);
        const uint32_t r0 = (uint32_t)rng.uniform(m);
        keys[i] = r0 >> shift;
        offs[i] = (uint32_t)i;
    }
}

template <int KIND, int NCH>
__global__ __launch_bounds__(256) void k_run_ordered(GView g, PlanView P, uint64_t seed, uint64_t first,
                                                     uint64_t count, const uint32_t *__restrict__ order,
                                                     Acc *__restrict__ partials)
{
    AccL acc[NCH];
    acc_zero<NCH>(acc);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count; j += stride) {
        Rng rng;
        rng.init(seed, TAG_ESTIMATOR, first + order[j]);
        Est o;
        run_one<KIND>(g, P, rng, o);
#pragma unroll
        for (int c = 0; c < NCH; ++c) acc_add(acc[c], o.v[c], o.st[c], o.den[c], c == 0 ? P.K : P.K1);
    }
    block_reduce_store<NCH>(acc, partials + (size_t)blockIdx.x * NCH);
}

// Batched queries of every plan kind in one persistent launch.  Each query's
// estimator ids are cut into chunks of Q.chunk ids; every wavefront claims
// chunks from one global counter (a work queue) until it passes `nchunks`,
// runs the chunk's estimators (Q.chunk / 64 per lane) with the kernel kind of
// the chunk's query, and writes the chunk's reduced record to its own slot.
// The per-query reduction (k_batch_final) sums each query's records in chunk
// order, so results do not depend on which wavefront ran which chunk.  Queries
// are laid out by decreasing cost per estimator, so the long chunks are
// claimed first and the queue drains on short ones.  No workgroup barrier:
// the waves of a workgroup run different chunks, of different kinds.
__device__ __forceinline__ uint32_t uni32(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uni64(uint64_t x)
{
    return ((uint64_t)uni32((uint32_t)(x >> 32)) << 32) | uni32((uint32_t)x);
}
template <class T>
__device__ __forceinline__ T *uniptr(T *p)
{
    return reinterpret_cast<T *>(uni64(reinterpret_cast<uint64_t>(p)));
}

// Not inlined: the kernel's register footprint is then the largest plan's,
// not the union of every plan's (measured: 256 VGPRs inlined).  The query's
// fields are read once and made wave-uniform (scalar registers), as a
// kernel's arguments are; read as per-lane values they take ~30 VGPRs.
template <int KIND, int NCH>
__device__ __attribute__((noinline)) void batch_chunk(const BQuery *__restrict__ qp, uint64_t base, uint64_t end,
                                                      Acc *__restrict__ rec)
{
    const uint32_t lane = threadIdx.x & 63;
    GView g;
    g.edges = uniptr(qp->g.edges);
    g.off = uniptr(qp->g.off);
    g.adj = uniptr(qp->g.adj);
    g.sadj = uniptr(qp->g.sadj);
    g.htab = uniptr(qp->g.htab);
    g.m = uni64(qp->g.m);
    g.n = uni32(qp->g.n);
    g.hnb = uni32(qp->g.hnb);
    g.simple = uni32(qp->g.simple);
    g.pad_ = 0;
    g.reads = nullptr;
    PlanView P;
    P.kind = (int)uni32((uint32_t)qp->P.kind);
    P.order = (int)uni32((uint32_t)qp->P.order);
    P.pk = (int)uni32((uint32_t)qp->P.pk);
    P.pe = (int)uni32((uint32_t)qp->P.pe);
    P.K = (int)uni32((uint32_t)qp->P.K);
    P.nch = (int)uni32((uint32_t)qp->P.nch);
    P.accept = uniptr(qp->P.accept);
    P.K1 = (int)uni32((uint32_t)qp->P.K1);
    const uint64_t seed = uni64(qp->seed), first = uni64(qp->first);
    base = uni64(base);
    end = uni64(end);
    rec = uniptr(rec);
    AccL acc[NCH];
    acc_zero<NCH>(acc);
    if constexpr (split_kind(KIND)) {
        for (uint64_t b = base; b < end; b += 64) {   // uniform: every lane reaches the completions
            const uint64_t i = b + lane;
            Est o;
            uint32_t words;
            run_split<KIND>(g, P, seed, first + i, i < end, o, words);
            if (i < end) {
#pragma unroll
                for (int ch = 0; ch < NCH; ++ch) acc_add(acc[ch], o.v[ch], o.st[ch], o.den[ch], ch == 0 ? P.K : P.K1);
            }
        }
    } else {
        for (uint64_t i = base + lane; i < end; i += 64) {
            Rng rng;
            rng.init(seed, TAG_ESTIMATOR, first + i);
            Est o;
            run_one<KIND>(g, P, rng, o);
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) acc_add(acc[ch], o.v[ch], o.st[ch], o.den[ch], ch == 0 ? P.K : P.K1);
        }
    }
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
        const AccW r = wave_reduce(acc[ch]);
        if (lane == 0) {
            Acc t;
#pragma unroll
            for (int d = 0; d < 5; ++d) t.dig[d] = r.dig[d];
            t.succ = r.succ;
            t.ovf = r.ovf;
            t.sq = r.sq;
            rec[ch] = t;
        }
    }
}

__device__ __forceinline__ uint64_t wave_uniform(uint64_t x)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_batch(
    const BQuery *__restrict__ q, int nq, uint64_t nchunks,
                                               unsigned long long *__restrict__ counter, Acc *__restrict__ recs)
{
    const uint32_t lane = threadIdx.x & 63;
    for (;;) {
        unsigned long long c0 = 0;
        if (lane == 0) c0 = atomicAdd(counter, 1ull);
        const uint64_t c = wave_uniform(__shfl(c0, 0, 64));
        if (c >= nchunks) break;   // uniform: the whole wavefront leaves, every wave reaches this
        int lo = 0, hi = nq - 1;   // the last query whose first chunk is <= c
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (q[mid].chunk0 <= c) lo = mid; else hi = mid - 1;
        }
        const BQuery &Q = q[lo];
        const uint64_t j = c - Q.chunk0;
        const uint64_t base = j * Q.chunk;
        const uint64_t end = base + Q.chunk < Q.count ? base + Q.chunk : Q.count;
        Acc *const rec = recs + Q.rec0 + j * (uint64_t)Q.P.nch;
        switch (Q.kk) {
        case ZG_PLAN_TRIANGLE: batch_chunk<ZG_PLAN_TRIANGLE, 1>(&Q, base, end, rec); break;
        case ZG_PLAN_THREE_CHAIN: batch_chunk<ZG_PLAN_THREE_CHAIN, 1>(&Q, base, end, rec); break;
        case ZG_PLAN_THREE_MOTIF: batch_chunk<ZG_PLAN_THREE_MOTIF, 2>(&Q, base, end, rec); break;
        case ZG_PLAN_CHAIN: batch_chunk<ZG_PLAN_CHAIN, 1>(&Q, base, end, rec); break;
        case ZG_PLAN_FOUR_CHAIN: batch_chunk<ZG_PLAN_FOUR_CHAIN, 1>(&Q, base, end, rec); break;
        case ZG_PLAN_K4_CLOSE: batch_chunk<ZG_PLAN_K4_CLOSE, 1>(&Q, base, end, rec); break;
        case ZG_PLAN_TRIANGLE_K4: batch_chunk<ZG_PLAN_TRIANGLE_K4, 2>(&Q, base, end, rec); break;
        case ZG_PLAN_K4_SPLIT: batch_chunk<ZG_PLAN_K4_SPLIT, 1>(&Q, base, end, rec); break;
        case ZG_PLAN_TRIANGLE_K4_SPLIT: batch_chunk<ZG_PLAN_TRIANGLE_K4_SPLIT, 2>(&Q, base, end, rec); break;
        case ZG_PLAN_HOUSE_SPLIT: batch_chunk<ZG_PLAN_HOUSE_SPLIT, 1>(&Q, base, end, rec); break;
        case KIND_PATTERN_BASE + 2: batch_chunk<KIND_PATTERN_BASE + 2, 1>(&Q, base, end, rec); break;
        case KIND_PATTERN_BASE + 3: batch_chunk<KIND_PATTERN_BASE + 3, 1>(&Q, base, end, rec); break;
        case KIND_PATTERN_BASE + 4: batch_chunk<KIND_PATTERN_BASE + 4, 1>(&Q, base, end, rec); break;
        case KIND_PATTERN_BASE + 5: batch_chunk<KIND_PATTERN_BASE + 5, 1>(&Q, base, end, rec); break;
        default: batch_chunk<KIND_PATTERN_BASE + 6, 1>(&Q, base, end, rec); break;
        }
    }
}

// per-query reduction of the chunk records (one workgroup per query), in chunk order
__global__ __launch_bounds__(256) void k_batch_final(const BQuery *__restrict__ q, const Acc *__restrict__ recs,
                                                     uint64_t nchunks, Acc *__restrict__ result)
{
    __shared__ Acc sm[256];
    const int qi = blockIdx.x;
    const BQuery &Q = q[qi];
    const int nch = Q.P.nch;
    const uint64_t nc = ((qi + 1 < (int)gridDim.x) ? q[qi + 1].chunk0 : nchunks) - Q.chunk0;
    for (int c = 0; c < nch; ++c) {
        Acc t;
        for (int i = 0; i < 5; ++i) t.dig[i] = 0;
        t.succ = t.ovf = 0;
        t.sq = 0;
        for (uint64_t b = threadIdx.x; b < nc; b += blockDim.x) {
            const Acc &p = recs[Q.rec0 + b * nch + c];
            for (int i = 0; i < 5; ++i) t.dig[i] += p.dig[i];
            t.succ += p.succ;
            t.ovf += p.ovf;
            t.sq += p.sq;
        }
        sm[threadIdx.x] = t;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                for (int i = 0; i < 5; ++i) sm[threadIdx.x].dig[i] += sm[threadIdx.x + s].dig[i];
                sm[threadIdx.x].succ += sm[threadIdx.x + s].succ;
                sm[threadIdx.x].ovf += sm[threadIdx.x + s].ovf;
                sm[threadIdx.x].sq += sm[threadIdx.x + s].sq;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) result[Q.res0 + c] = sm[0];
        __syncthreads();
    }
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// deterministic grid reduction of the block partials (one workgroup)
__global__ __launch_bounds__(256) void k_final(const Acc *__restrict__ partials, int nblocks, int nch,
                                               Acc *__restrict__ result)
{
    __shared__ Acc sm[256];
    for (int c = 0; c < nch; ++c) {
        Acc t;
        for (int i = 0; i < 5; ++i) t.dig[i] = 0;
        t.succ = t.ovf = 0;
        t.sq = 0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
            const Acc &p = partials[b * nch + c];
            for (int i = 0; i < 5; ++i) t.dig[i] += p.dig[i];
            t.succ += p.succ;
            t.ovf += p.ovf;
            t.sq += p.sq;
        }
        sm[threadIdx.x] = t;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                for (int i = 0; i < 5; ++i) sm[threadIdx.x].dig[i] += sm[threadIdx.x + s].dig[i];
                sm[threadIdx.x].succ += sm[threadIdx.x + s].succ;
                sm[threadIdx.x].ovf += sm[threadIdx.x + s].ovf;
                sm[threadIdx.x].sq += sm[threadIdx.x + s].sq;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) result[c] = sm[0];
        __syncthreads();
    }
}

template <int KIND, int NCH>
static void launch_kind(const zg_graph *g, const PlanView &P, uint64_t seed, uint64_t first, uint64_t count,
                        zg_outcome *d_out, int grid, int block, unsigned long long *reads, const uint32_t *order,
                        Acc *partials)
{
    GView v = g->view();
    v.reads = reads;
    if constexpr (!split_kind(KIND)) {   // split plans: id order only
        if (order) {
            hipLaunchKernelGGL((k_run_ordered<KIND, NCH>), dim3(grid), dim3(block), 0, g->stream, v, P, seed, first,
                               count, order, partials);
            return;
        }
    }
    if (d_out)
        hipLaunchKernelGGL((k_run<KIND, NCH, true>), dim3(grid), dim3(block), 0, g->stream, v, P, seed, first, count,
                           d_out, partials);
    else
        hipLaunchKernelGGL((k_run<KIND, NCH, false>), dim3(grid), dim3(block), 0, g->stream, v, P, seed, first,
                           count, d_out, partials);
}

// Automatic = estimator-id order: on MI355X the ordered launch measured slower
// for 2^28 triangle / 4-clique estimators on a 69 M-edge graph (sort included:
// 19.6-20.5 ms against 17.2-17.8 ms), so ordering is opt-in.
int resolve_order(const zg_graph *g, const zg_plan *p, uint64_t count, int order)
{
    (void)g;
    (void)p;
    (void)count;
    return order == 0 ? 1 : order;
}

// sorts the launch's estimator offsets by first-edge bucket into g->ord
static int order_by_first_edge(zg_graph *g, const zg_plan *p, uint64_t seed, uint64_t first, uint64_t count,
                               int bits, const uint32_t **order)
{
    if (count > kMaxOrderedLaunch) return set_error(ZG_EINVAL, "ordered launch too large");
    if (count > g->ord_cap) {
        for (uint32_t *&q : g->ord) {
            if (q) ZG_HIP(hipFree(q));
            q = nullptr;
        }
        g->ord_cap = 0;
        for (uint32_t *&q : g->ord)
            if (hipMalloc(&q, 4 * count) != hipSuccess) return set_error(ZG_ENOMEM, "ordering scratch: allocation failed");
        g->ord_cap = count;
    }
    unsigned mbits = 1;
    while (mbits < 32 && (1ull << mbits) < g->m) ++mbits;
    const unsigned kbits = (unsigned)std::min<int>(bits, (int)mbits);
    const uint32_t shift = mbits - kbits;
    const int grid = std::max(1, std::min(device_cus() * 16, (int)((count + 255) / 256)));
    if (p->desc.kind == ZG_PLAN_FOUR_CHAIN)
        hipLaunchKernelGGL(k_first_edge_keys<true>, dim3(grid), dim3(256), 0, g->stream, g->m, seed, first, count,
                           shift, g->ord[0], g->ord[1]);
    else
        hipLaunchKernelGGL(k_first_edge_keys<false>, dim3(grid), dim3(256), 0, g->stream, g->m, seed, first, count,
                           shift, g->ord[0], g->ord[1]);
    ZG_HIP(hipGetLastError());
    rocprim::double_buffer<uint32_t> kb(g->ord[0], g->ord[2]);
    rocprim::double_buffer<uint32_t> vb(g->ord[1], g->ord[3]);
    size_t tb = 0;
    ZG_HIP(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)count, 0, kbits, g->stream));
    if (tb > g->ord_tmp_cap) {
        if (g->ord_tmp) ZG_HIP(hipFree(g->ord_tmp));
        g->ord_tmp = nullptr;
        g->ord_tmp_cap = 0;
        if (hipMalloc(&g->ord_tmp, tb) != hipSuccess) return set_error(ZG_ENOMEM, "ordering scratch: allocation failed");
        g->ord_tmp_cap = tb;
    }
    ZG_HIP(rocprim::radix_sort_pairs(g->ord_tmp, tb, kb, vb, (size_t)count, 0, kbits, g->stream));
    *order = vb.current();
    return ZG_OK;
}

int launch_run(zg_graph *g, const zg_plan *p, uint64_t seed, uint64_t first, uint64_t count,
               zg_outcome *d_out, int block, int blocks_per_cu, unsigned long long *reads, int order,
               Acc *ext_partials, size_t ext_cap, Acc *ext_result)
{
    if (block <= 0) block = 256;
    if (block % 64 || block > 1024 || block > 256)
        return set_error(ZG_EINVAL, "block_size must be a multiple of 64 and at most 256");
    if (blocks_per_cu <= 0) blocks_per_cu = 16;
    const uint64_t want = (count + block - 1) / block;
    uint64_t grid = (uint64_t)device_cus() * blocks_per_cu;
    if (want < grid) grid = want;
    if (grid == 0) grid = 1;
    const int nch = p->nch;
    if (order < 0 || order > 16) return set_error(ZG_EINVAL, "order must be 0 (automatic), 1 (id) or 2..16");
    Acc *partials = g->partials, *result = g->result;
    if (ext_partials) {
        if (order >= 2) return set_error(ZG_EINVAL, "a run into caller scratch takes estimator-id order only");
        if (grid * nch > ext_cap)
            return set_error(ZG_EINVAL, "the launch needs " + std::to_string(grid * nch) +
                                            " block partial records but the scratch holds " +
                                            std::to_string(ext_cap) + "; lower blocks_per_cu");
        partials = ext_partials;
        result = ext_result;
    } else if (grid * nch > g->partials_cap) {
        if (g->partials) ZG_HIP(hipFree(g->partials));
        g->partials = nullptr;
        ZG_HIP(hipMalloc(&g->partials, sizeof(Acc) * grid * nch));
        g->partials_cap = grid * nch;
        partials = g->partials;
    }
    const PlanView P = p->view();
    const uint32_t *ord = nullptr;
    if (order >= 2 && !d_out && !reads && count > 0 && g->m > 0 && !split_kind(p->desc.kind)) {
        int rc = order_by_first_edge(g, p, seed, first, count, order, &ord);
        if (rc) return rc;
    }
    switch (p->desc.kind) {
    case ZG_PLAN_TRIANGLE: launch_kind<ZG_PLAN_TRIANGLE, 1>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
    case ZG_PLAN_THREE_CHAIN: launch_kind<ZG_PLAN_THREE_CHAIN, 1>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
    case ZG_PLAN_THREE_MOTIF: launch_kind<ZG_PLAN_THREE_MOTIF, 2>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
    case ZG_PLAN_CHAIN: launch_kind<ZG_PLAN_CHAIN, 1>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
    case ZG_PLAN_FOUR_CHAIN: launch_kind<ZG_PLAN_FOUR_CHAIN, 1>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
    case ZG_PLAN_PATTERN:
        switch (P.pk) {
        case 2: launch_kind<KIND_PATTERN_BASE + 2, 1>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
        case 3: launch_kind<KIND_PATTERN_BASE + 3, 1>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
        case 4: launch_kind<KIND_PATTERN_BASE + 4, 1>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
        case 5: launch_kind<KIND_PATTERN_BASE + 5, 1>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
        case 6: launch_kind<KIND_PATTERN_BASE + 6, 1>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
        default: return set_error(ZG_EINVAL, "pattern vertices must be in [2, 6]");
        }
        break;
    case ZG_PLAN_K4_CLOSE: launch_kind<ZG_PLAN_K4_CLOSE, 1>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
    case ZG_PLAN_TRIANGLE_K4: launch_kind<ZG_PLAN_TRIANGLE_K4, 2>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
    case ZG_PLAN_K4_SPLIT: launch_kind<ZG_PLAN_K4_SPLIT, 1>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
    case ZG_PLAN_TRIANGLE_K4_SPLIT: launch_kind<ZG_PLAN_TRIANGLE_K4_SPLIT, 2>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
    case ZG_PLAN_HOUSE_SPLIT: launch_kind<ZG_PLAN_HOUSE_SPLIT, 1>(g, P, seed, first, count, d_out, (int)grid, block, reads, ord, partials); break;
    default: return set_error(ZG_EINVAL, "unknown plan kind");
    }
    ZG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, g->stream, partials, (int)grid, nch, result);
    ZG_HIP(hipGetLastError());
    return ZG_OK;
}

int launch_batch(const BQuery *d_q, int nq, uint64_t nchunks, unsigned long long *counter, Acc *chunk_recs,
                 Acc *result, int grid, int block, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, counter, 1);
    hipLaunchKernelGGL(k_batch, dim3(grid), dim3(block), 0, st, d_q, nq, nchunks, counter, chunk_recs);
    hipLaunchKernelGGL(k_batch_final, dim3(nq), dim3(256), 0, st, d_q, chunk_recs, nchunks, result);
    ZG_HIP(hipGetLastError());
    return ZG_OK;
}

void finish_sums(const Acc *h, int nch, int K0, int K1, uint64_t count, zg_sums *out)
{
    for (int c = 0; c < nch; ++c) {
        const int K = c == 0 ? K0 : K1;
        // carry-propagate the 32-bit digit sums into 192 bits
        uint32_t limb[6] = {0, 0, 0, 0, 0, 0};
        unsigned __int128 carry = 0;
        for (int i = 0; i < 6; ++i) {
            unsigned __int128 t = carry + (i < 5 ? (unsigned __int128)h[c].dig[i] : 0);
            limb[i] = (uint32_t)t;
            carry = t >> 32;
        }
        zg_sums &s = out[c];
        s.sum[0] = (uint64_t)limb[0] | ((uint64_t)limb[1] << 32);
        s.sum[1] = (uint64_t)limb[2] | ((uint64_t)limb[3] << 32);
        s.sum[2] = (uint64_t)limb[4] | ((uint64_t)limb[5] << 32);
        s.log2_den = K;
        s.status = h[c].ovf ? ZG_EOVERFLOW : 0;
        s.successes = h[c].succ;
        s.count = count;
        s.sumsq = h[c].sq;
        long double v = (long double)s.sum[2] * 0x1p128L + (long double)s.sum[1] * 0x1p64L + (long double)s.sum[0];
        s.mean = count ? (double)(v / ldexpl(1.0L, K) / (long double)count) : 0.0;
    }
}

}  // namespace zg
