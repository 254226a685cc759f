/*
 * rfmf_oracle.c -- CPU restatement of RFMF's annealer (test infrastructure;
 * see rfmf_oracle.h for scope and for what pins each function).
 *
 * Citations are path:line in the reference checkout.
 */
#define _GNU_SOURCE
#include "rfmf_oracle.h"

#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 -- src/philox.cuh:60-98 (Salmon et al. 2011).               */
/* ------------------------------------------------------------------------ */

#define OR_PHILOX_M0 0xD2511F53u
#define OR_PHILOX_M1 0xCD9E8D57u
#define OR_PHILOX_W0 0x9E3779B9u
#define OR_PHILOX_W1 0xBB67AE85u

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += OR_PHILOX_W0; k1 += OR_PHILOX_W1; }
        const uint64_t p0 = (uint64_t)OR_PHILOX_M0 * c0;
        const uint64_t p1 = (uint64_t)OR_PHILOX_M1 * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------------ */
/* Gumbel LUT -- src/philox.cuh:26-36,52-58.  Derived here from the closed   */
/* form coarse[i] = round(-ln(-ln((2i+1)/32)) * 2^20) rather than retyped.   */
/* ------------------------------------------------------------------------ */

static int32_t g_lut[16];
static pthread_once_t g_lut_once = PTHREAD_ONCE_INIT;

static void lut_init(void) {
    for (int i = 0; i < 16; ++i) {
        const double u = (2.0 * i + 1.0) / 32.0;
        g_lut[i] = (int32_t)llround(-log(-log(u)) * 1048576.0);
    }
}

int32_t or_gumbel_lut(int i) {
    pthread_once(&g_lut_once, lut_init);
    return g_lut[i & 15];
}

/* r must be exactly 11 bits (philox.cuh:50-51). */
int32_t or_gumbel_from_bits(uint32_t r11) {
    pthread_once(&g_lut_once, lut_init);
    return g_lut[r11 & 15u] + (int32_t)(r11 >> 4);
}

/* [biased score : 32 | ~id : 32] -- src/philox.cuh:128-152. */
uint64_t or_pack_score_id(int32_t score, int32_t id) {
    const uint32_t biased = (uint32_t)score ^ 0x80000000u;
    return ((uint64_t)biased << 32) | (uint64_t)(~(uint32_t)id);
}

/* key = solver seed split lo/hi (philox.cuh:105-110); counter =
 * (step lo, step hi, block lo, block hi) with block = v >> 3
 * (philox.cuh:112-119); lane j = v & 7 takes 16-bit field j of the output,
 * low 11 bits (philox.cuh:122-126). */
int32_t or_vertex_gumbel(uint64_t solver_seed, uint64_t step, uint64_t v) {
    const uint64_t blk = v >> 3;
    const uint32_t ctr[4] = {(uint32_t)step, (uint32_t)(step >> 32),
                             (uint32_t)blk, (uint32_t)(blk >> 32)};
    const uint32_t key[2] = {(uint32_t)solver_seed, (uint32_t)(solver_seed >> 32)};
    uint32_t o[4];
    or_philox4x32_10(ctr, key, o);
    const uint32_t j = (uint32_t)(v & 7u);
    const uint32_t word = o[j >> 1];
    const uint32_t field = (j & 1u) ? (word >> 16) : (word & 0xFFFFu);
    return or_gumbel_from_bits(field & 0x7FFu);
}

/* ------------------------------------------------------------------------ */
/* Initial spins -- src/rng.hpp:11-58 (splitmix64 seeding of xoshiro256++,   */
/* Vigna) and src/host_common.cpp:40-45 (one bit per vertex, +1 if set).     */
/* ------------------------------------------------------------------------ */

uint64_t or_splitmix64_next(uint64_t* state) {
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void or_initial_spins(uint64_t raw_seed, int64_t n, int8_t* spin_out) {
    uint64_t s[4];
    uint64_t sm = raw_seed;
    for (int i = 0; i < 4; ++i) s[i] = or_splitmix64_next(&sm);
    uint64_t buf = 0;
    int left = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (left < 1) {
            const uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
            const uint64_t t = s[1] << 17;
            s[2] ^= s[0];
            s[3] ^= s[1];
            s[1] ^= s[2];
            s[0] ^= s[3];
            s[2] ^= t;
            s[3] = rotl64(s[3], 45);
            buf = result;
            left = 64;
        }
        spin_out[i] = (buf & 1u) ? (int8_t)1 : (int8_t)-1;
        buf >>= 1;
        --left;
    }
}

/* ------------------------------------------------------------------------ */
/* Beta schedule -- src/host_common.h:73-76 (value), :82-94 (run layout).    */
/* The runs are found here by a plain per-step scan, not by the reference's  */
/* gallop and bisection (firstStepAbove, appendBetaRuns:                    */
/* src/host_common.cpp:212-273), so the two are independent.                 */
/* ------------------------------------------------------------------------ */

int32_t or_beta_fixed(uint64_t step, double t0) {
    const double beta = log(1.0 + (double)step) / t0;
    return (int32_t)round(beta * 1048576.0);
}

int64_t or_beta_runs_log(double t0, int64_t total, int64_t* run_end,
                         int32_t* run_val, int64_t cap) {
    if (total <= 0) {
        if (cap < 1) return -1;
        run_end[0] = 0; run_val[0] = 0;
        return 1;
    }
    int64_t r = 0;
    int32_t cur = or_beta_fixed(1, t0);
    for (int64_t s = 2; s <= total; ++s) {
        const int32_t v = or_beta_fixed((uint64_t)s, t0);
        if (v != cur) {
            if (r >= cap) return -1;
            run_end[r] = s - 1;   /* 0-based exclusive end of the previous run */
            run_val[r] = cur;
            ++r;
            cur = v;
        }
    }
    if (r >= cap) return -1;
    run_end[r] = total;
    run_val[r] = cur;
    return r + 1;
}

int32_t or_beta_linear(uint64_t step, int64_t total, int32_t b0, int32_t b1) {
    if (total <= 1) return b0;
    const __int128 num = (__int128)((int64_t)b1 - (int64_t)b0) * (__int128)((int64_t)step - 1);
    return (int32_t)((int64_t)b0 + (int64_t)(num / (__int128)(total - 1)));
}

int64_t or_beta_runs_linear(int32_t b0, int32_t b1, int64_t total,
                            int64_t* run_end, int32_t* run_val, int64_t cap) {
    if (total <= 0) {
        if (cap < 1) return -1;
        run_end[0] = 0; run_val[0] = 0;
        return 1;
    }
    int64_t r = 0;
    int32_t cur = or_beta_linear(1, total, b0, b1);
    for (int64_t s = 2; s <= total; ++s) {
        const int32_t v = or_beta_linear((uint64_t)s, total, b0, b1);
        if (v != cur) {
            if (r >= cap) return -1;
            run_end[r] = s - 1;
            run_val[r] = cur;
            ++r;
            cur = v;
        }
    }
    if (r >= cap) return -1;
    run_end[r] = total;
    run_val[r] = cur;
    return r + 1;
}

/* ------------------------------------------------------------------------ */
/* int32 gate -- src/host_common.cpp:296-312; B -- :314-325.                 */
/* ------------------------------------------------------------------------ */

#define OR_GUMBEL_MAX_MAG 3617613LL   /* LUT[15] + 127 (philox.cuh:44-48) */

int or_score_safe_int32(int64_t total_steps, double t0, int64_t delta_bound) {
    const double beta_max = log(1.0 + (double)total_steps) / t0;
    const int64_t bmf = (int64_t)llround(beta_max * 1048576.0);
    const int64_t limit = (int64_t)INT32_MAX - OR_GUMBEL_MAX_MAG;
    if (bmf > 0 && delta_bound > 0 && delta_bound > limit / bmf) return 0;
    return (bmf * delta_bound + OR_GUMBEL_MAX_MAG) < (int64_t)INT32_MAX;
}

int64_t or_max_weighted_degree(const uint32_t* off, const int32_t* to,
                               const int32_t* w, int64_t n) {
    (void)to;
    int64_t B = 0;
    for (int64_t v = 0; v < n; ++v) {
        int64_t s = 0;
        for (uint32_t i = off[v]; i < off[v + 1]; ++i) s += llabs((long long)w[i]);
        if (s > B) B = s;
    }
    return B;
}

/* ------------------------------------------------------------------------ */
/* Cut -- src/max_cut.cpp:5-15: each undirected edge once (to > v).          */
/* ------------------------------------------------------------------------ */

int64_t or_cut_value(const uint32_t* off, const int32_t* to, const int32_t* w,
                     int64_t n, const int8_t* spin) {
    int64_t value = 0;
    for (int64_t v = 0; v < n; ++v)
        for (uint32_t i = off[v]; i < off[v + 1]; ++i)
            if ((int64_t)to[i] > v && spin[v] != spin[to[i]]) value += w[i];
    return value;
}

/* ------------------------------------------------------------------------ */
/* CSR build -- src/graph.cpp:11-17 (dedup, first weight wins),              */
/* src/immutable_graph.hpp:44-116 (dense ids by ascending original id, rows  */
/* sorted by dense neighbour id).  Vertices are exactly the edge endpoints,  */
/* so degree-zero removal (removeDegreeZeroVertices,                         */
/* src/host_common.cpp:184-194) is implied.                                  */
/* Implemented with sorting instead of hash maps.                            */
/* ------------------------------------------------------------------------ */

typedef struct { int64_t a, b, w, ord; } or_arc;

static int cmp_i64(const void* x, const void* y) {
    const int64_t a = *(const int64_t*)x, b = *(const int64_t*)y;
    return (a > b) - (a < b);
}

static int cmp_arc(const void* x, const void* y) {
    const or_arc* p = (const or_arc*)x;
    const or_arc* q = (const or_arc*)y;
    if (p->a != q->a) return (p->a > q->a) - (p->a < q->a);
    if (p->b != q->b) return (p->b > q->b) - (p->b < q->b);
    return (p->ord > q->ord) - (p->ord < q->ord);
}

static int64_t rank_of(const int64_t* ids, int64_t n, int64_t key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (ids[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

int64_t or_build_csr(const int64_t* eu, const int64_t* ev, const int64_t* ew,
                     int64_t m, uint32_t** off_out, int32_t** to_out,
                     int32_t** w_out, int64_t** new_to_old_out) {
    /* Undirected dedup keyed on the unordered pair, first occurrence wins. */
    or_arc* und = (or_arc*)malloc(sizeof(or_arc) * (size_t)(m > 0 ? m : 1));
    int64_t* ids = (int64_t*)malloc(sizeof(int64_t) * (size_t)(2 * m + 1));
    if (!und || !ids) { free(und); free(ids); return -1; }
    for (int64_t i = 0; i < m; ++i) {
        const int64_t a = eu[i] < ev[i] ? eu[i] : ev[i];
        const int64_t b = eu[i] < ev[i] ? ev[i] : eu[i];
        und[i].a = a; und[i].b = b; und[i].w = ew[i]; und[i].ord = i;
        ids[2 * i] = eu[i]; ids[2 * i + 1] = ev[i];
    }
    qsort(und, (size_t)m, sizeof(or_arc), cmp_arc);
    int64_t mu = 0;
    for (int64_t i = 0; i < m; ++i)
        if (mu == 0 || und[i].a != und[mu - 1].a || und[i].b != und[mu - 1].b)
            und[mu++] = und[i];
    qsort(ids, (size_t)(2 * m), sizeof(int64_t), cmp_i64);
    int64_t n = 0;
    for (int64_t i = 0; i < 2 * m; ++i)
        if (n == 0 || ids[i] != ids[n - 1]) ids[n++] = ids[i];

    /* Directed arcs: both directions, a self loop once (adj[u][u]). */
    or_arc* arcs = (or_arc*)malloc(sizeof(or_arc) * (size_t)(2 * mu + 1));
    int64_t na = 0;
    for (int64_t i = 0; i < mu; ++i) {
        const int64_t ra = rank_of(ids, n, und[i].a), rb = rank_of(ids, n, und[i].b);
        arcs[na].a = ra; arcs[na].b = rb; arcs[na].w = und[i].w; arcs[na].ord = 0; ++na;
        if (ra != rb) { arcs[na].a = rb; arcs[na].b = ra; arcs[na].w = und[i].w; arcs[na].ord = 0; ++na; }
    }
    qsort(arcs, (size_t)na, sizeof(or_arc), cmp_arc);
    uint32_t* off = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
    int32_t* to = (int32_t*)malloc(sizeof(int32_t) * (size_t)(na > 0 ? na : 1));
    int32_t* w = (int32_t*)malloc(sizeof(int32_t) * (size_t)(na > 0 ? na : 1));
    for (int64_t i = 0; i < na; ++i) {
        off[arcs[i].a + 1]++;
        to[i] = (int32_t)arcs[i].b;
        w[i] = (int32_t)arcs[i].w;
    }
    for (int64_t v = 0; v < n; ++v) off[v + 1] += off[v];
    free(und); free(arcs);
    *off_out = off; *to_out = to; *w_out = w;
    if (new_to_old_out) *new_to_old_out = ids; else free(ids);
    return n;
}

static int next_data_line(FILE* f, char* buf, size_t cap) {
    while (fgets(buf, (int)cap, f)) {
        const char* p = buf;
        while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
        if (*p == '\0' || *p == '\n' || *p == '#') continue;
        return 1;
    }
    return 0;
}

int64_t or_read_gset(const char* path, uint32_t** off_out, int32_t** to_out,
                     int32_t** w_out, int64_t* header_n, int64_t* header_m) {
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    char line[4096];
    long long hn = 0, hm = 0;
    if (!next_data_line(f, line, sizeof line) || sscanf(line, "%lld %lld", &hn, &hm) != 2 ||
        hn < 0 || hm < 0) { fclose(f); return -1; }
    int64_t* eu = (int64_t*)malloc(sizeof(int64_t) * (size_t)(hm + 1));
    int64_t* ev = (int64_t*)malloc(sizeof(int64_t) * (size_t)(hm + 1));
    int64_t* ew = (int64_t*)malloc(sizeof(int64_t) * (size_t)(hm + 1));
    int64_t m = 0;
    for (long long i = 0; i < hm; ++i) {
        long long u, v, w;
        if (!next_data_line(f, line, sizeof line) || sscanf(line, "%lld %lld %lld", &u, &v, &w) != 3) {
            free(eu); free(ev); free(ew); fclose(f); return -1;
        }
        if (w == 0) continue;   /* read_gset.cpp:200: zero weights dropped */
        eu[m] = u; ev[m] = v; ew[m] = w; ++m;
    }
    fclose(f);
    const int64_t n = or_build_csr(eu, ev, ew, m, off_out, to_out, w_out, NULL);
    free(eu); free(ev); free(ew);
    if (header_n) *header_n = hn;
    if (header_m) *header_m = hm;
    return n;
}

void or_free(void* p) { free(p); }

/* ------------------------------------------------------------------------ */
/* The annealing engine -- semantics of mcKernel, src/anneal_kernels.cu:     */
/* 208-491, restated with SEQUENTIAL group-order flips:                      */
/*   per sweep: beta from the run table (:263-279); every vertex scored from */
/*   PRE-SWEEP spin/S (:292-346): score = gumbel + beta * spin*S, per group  */
/*   g = v mod P the max score, ties to the smaller v (:128-146);            */
/*   then for g ascending the winner flips against the LIVE field, which is  */
/*   what corr[] reproduces analytically (:133-155, :420-423);               */
/*   cur > best (strict) snapshots the spins (:432-440, :303-304, :488-490). */
/* ------------------------------------------------------------------------ */

#define OR_CHUNK 8

int or_anneal_chain(const uint32_t* off, const int32_t* to, const int32_t* w,
                    int64_t n, const int8_t* init_spin, uint64_t solver_seed,
                    const int64_t* run_end, const int32_t* run_val, int64_t n_runs,
                    const uint64_t* phases, const uint64_t* iters, int64_t n_phases,
                    int8_t* end_spin, int8_t* best_spin, or_chain_result* res) {
    pthread_once(&g_lut_once, lut_init);
    int8_t* spin = (int8_t*)malloc((size_t)(n > 0 ? n : 1));
    int64_t* S = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    uint64_t maxP = 1;
    for (int64_t p = 0; p < n_phases; ++p) if (phases[p] > maxP) maxP = phases[p];
    int64_t* slotScore = (int64_t*)malloc(sizeof(int64_t) * maxP);
    int64_t* slotV = (int64_t*)malloc(sizeof(int64_t) * maxP);
    int32_t* gum = (int32_t*)malloc(sizeof(int32_t) * (size_t)((n + 7) / 8 * 8 + 8));
    if (!spin || !S || !slotScore || !slotV || !gum) {
        free(spin); free(S); free(slotScore); free(slotV); free(gum);
        return -1;
    }
    memcpy(spin, init_spin, (size_t)n);
    memcpy(best_spin, init_spin, (size_t)n);
    for (int64_t v = 0; v < n; ++v) {
        int64_t acc = 0;
        for (uint32_t i = off[v]; i < off[v + 1]; ++i) acc += (int64_t)w[i] * spin[to[i]];
        S[v] = acc;
    }
    int64_t cur = or_cut_value(off, to, w, n, spin);
    int64_t best = cur;
    const uint32_t k0 = (uint32_t)solver_seed, k1 = (uint32_t)(solver_seed >> 32);
    const int64_t nBlk = (n + 7) / 8;
    uint64_t step = 0;
    int64_t run = 0;

    for (int64_t p = 0; p < n_phases; ++p) {
        const uint64_t P = phases[p];
        if (P == 0) continue;
        for (uint64_t it = 0; it < iters[p]; ++it) {
            while (run < n_runs - 1 && (uint64_t)run_end[run] <= step) ++run;
            const int64_t beta = run_val[run];
            ++step;
            /* Gumbels for every vertex of this sweep, OR_CHUNK blocks at a time. */
            for (int64_t b0 = 0; b0 < nBlk; b0 += OR_CHUNK) {
                uint32_t c0[OR_CHUNK], c1[OR_CHUNK], c2[OR_CHUNK], c3[OR_CHUNK];
                for (int k = 0; k < OR_CHUNK; ++k) {
                    const uint64_t blk = (uint64_t)(b0 + k);
                    c0[k] = (uint32_t)step; c1[k] = (uint32_t)(step >> 32);
                    c2[k] = (uint32_t)blk; c3[k] = (uint32_t)(blk >> 32);
                }
                uint32_t kk0 = k0, kk1 = k1;
                for (int r = 0; r < 10; ++r) {
                    if (r) { kk0 += OR_PHILOX_W0; kk1 += OR_PHILOX_W1; }
                    for (int k = 0; k < OR_CHUNK; ++k) {
                        const uint64_t p0 = (uint64_t)OR_PHILOX_M0 * c0[k];
                        const uint64_t p1 = (uint64_t)OR_PHILOX_M1 * c2[k];
                        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1[k] ^ kk0;
                        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3[k] ^ kk1;
                        c1[k] = (uint32_t)p1; c3[k] = (uint32_t)p0;
                        c0[k] = n0; c2[k] = n2;
                    }
                }
                for (int k = 0; k < OR_CHUNK && b0 + k < nBlk; ++k) {
                    const uint32_t o[4] = {c0[k], c1[k], c2[k], c3[k]};
                    int32_t* gdst = gum + (b0 + k) * 8;
                    for (int j = 0; j < 8; ++j) {
                        const uint32_t word = o[j >> 1];
                        const uint32_t r11 = ((j & 1) ? (word >> 16) : word) & 0x7FFu;
                        gdst[j] = g_lut[r11 & 15u] + (int32_t)(r11 >> 4);
                    }
                }
            }
            /* Per-group argmax over pre-sweep state; v ascending, so a strict
             * > keeps the smaller id on ties. */
            for (uint64_t g = 0; g < P; ++g) { slotScore[g] = INT64_MIN; slotV[g] = -1; }
            uint64_t g = 0;
            for (int64_t v = 0; v < n; ++v) {
                const int64_t score = (int64_t)gum[v] + beta * ((int64_t)spin[v] * S[v]);
                if (slotV[g] < 0 || score > slotScore[g]) { slotScore[g] = score; slotV[g] = v; }
                if (++g == P) g = 0;
            }
            /* Sequential flips in ascending group order against the live S. */
            const uint64_t groups = P < (uint64_t)n ? P : (uint64_t)n;
            for (uint64_t gg = 0; gg < groups; ++gg) {
                const int64_t b = slotV[gg];
                if (b < 0) continue;
                cur += (int64_t)spin[b] * S[b];
                spin[b] = (int8_t)-spin[b];
                const int64_t two_s = 2 * (int64_t)spin[b];
                for (uint32_t i = off[b]; i < off[b + 1]; ++i) S[to[i]] += two_s * w[i];
            }
            if (cur > best) { best = cur; memcpy(best_spin, spin, (size_t)n); }
        }
    }
    memcpy(end_spin, spin, (size_t)n);
    if (res) { res->cur = cur; res->best = best; res->steps = step; }
    free(spin); free(S); free(slotScore); free(slotV); free(gum);
    return 0;
}

typedef struct {
    const uint32_t* off; const int32_t* to; const int32_t* w; int64_t n; int64_t C;
    const int8_t* init; const uint64_t* seeds; const int64_t* run_end; const int32_t* run_val;
    const int64_t* run_off; int64_t n_runs_total;
    const uint64_t* phases; const uint64_t* iters; int64_t n_phases;
    int8_t* end; int8_t* best; int64_t* curv; int64_t* bestv;
    int threads; int tid; int rc;
} or_job;

static void* or_worker(void* arg) {
    or_job* j = (or_job*)arg;
    for (int64_t c = j->tid; c < j->C; c += j->threads) {
        const int64_t ro = j->run_off[c];
        int64_t nr = j->n_runs_total - ro;
        or_chain_result r;
        if (or_anneal_chain(j->off, j->to, j->w, j->n, j->init + c * j->n, j->seeds[c],
                            j->run_end + ro, j->run_val + ro, nr, j->phases, j->iters,
                            j->n_phases, j->end + c * j->n, j->best + c * j->n, &r) != 0) {
            j->rc = -1;
            return NULL;
        }
        if (j->curv) j->curv[c] = r.cur;
        if (j->bestv) j->bestv[c] = r.best;
    }
    return NULL;
}

double or_anneal_many(const uint32_t* off, const int32_t* to, const int32_t* w,
                      int64_t n, int64_t C, const int8_t* init_spins,
                      const uint64_t* solver_seeds, const int64_t* run_end,
                      const int32_t* run_val, const int64_t* chain_run_off,
                      int64_t n_runs_total,
                      const uint64_t* phases, const uint64_t* iters, int64_t n_phases,
                      int threads, int8_t* end_spins, int8_t* best_spins,
                      int64_t* cur_vals, int64_t* best_vals) {
    if (C <= 0) return 0.0;
    if (threads <= 0) threads = (int)(C < 64 ? C : 64);
    if (threads > C) threads = (int)C;
    pthread_once(&g_lut_once, lut_init);
    or_job* jobs = (or_job*)calloc((size_t)threads, sizeof(or_job));
    pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < threads; ++t) {
        or_job j = {off, to, w, n, C, init_spins, solver_seeds, run_end, run_val,
                    chain_run_off, n_runs_total, phases, iters, n_phases,
                    end_spins, best_spins, cur_vals, best_vals, threads, t, 0};
        jobs[t] = j;
        pthread_create(&th[t], NULL, or_worker, &jobs[t]);
    }
    int rc = 0;
    for (int t = 0; t < threads; ++t) {
        pthread_join(th[t], NULL);
        if (jobs[t].rc) rc = -1;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(jobs); free(th);
    if (rc) return -1.0;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
