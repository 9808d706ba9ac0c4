/*
 * crispy_oracle.c -- CPU restatement of CRiSPy's kmerDist / genDist CPU path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product in
 * crispy_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product never links, calls or falls back to
 * it.
 *
 * Every function cites the reference file:line it restates.  Pinning (what
 * each function is checked against) is listed in oracle/README.md and in the
 * tests that exercise it:
 *   oracle_read_fasta      -- pinned against the reference binaries built in
 *                             oracle/_ref (tests/golden fixtures) and edge-case
 *                             files (blank lines, lowercase, N, long lines).
 *   oracle_kmer_tuples     -- closed form sum_q enc(s[t+q]) << 2(K-1-q) plus
 *                             the signed-short sort order.
 *   oracle_kmer_pair       -- multiset intersection (K<=7) by brute force,
 *                             the K=8 hi/lo closed form, and the reference
 *                             kmerDist -c output (golden).
 *   oracle_nw_pair         -- hand-derived cell-by-cell cases of both forms
 *                             (tests/test_oracle_nw_rules.py: zero-cost gap
 *                             switch, the -1000000 band edge, the diag > up >
 *                             left tie chain, the >= last maxima, the final
 *                             tie), each shown to catch a planted mistake in
 *                             a copy of this file; closed forms (identical
 *                             reads, one substitution, kB = 0).  Fixtures from
 *                             genCPU.cpp built against rocThrust (not the
 *                             reference's own build) are regression data.
 *   oracle_gendist_files   -- its file split and order: the reference's loop
 *                             (genCPU.cpp:466-505) by reading, and the same
 *                             rocThrust-built regression fixtures.
 *   oracle_nw_pairs_as_shipped -- the same distances as oracle_nw_pair (test)
 *                             under genDist's shipped thread settings; timing
 *                             only.
 */
#define _GNU_SOURCE
#include <math.h>
#include <sched.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define OR_BUF_SIZE 4096            /* kmerMain.h:24, genMain.h:22 */
#define OR_EPSILON 0.00001          /* kmerMain.h:27, genMain.h:40 */
#define OR_GO (-10)                 /* genMain.h:32 */
#define OR_GE (-1)                  /* genMain.h:33 */
#define OR_MISMATCH (-4)            /* genMain.h:34 */
#define OR_MATCH 5                  /* genMain.h:35 */
#define OR_NEG (-1000000)           /* genCPU.cpp:350,416,421 */
#define OR_NUM_PAIRS (576 * 256 * 4)     /* genMain.h:24-27: pairs per loadPairs chunk */
#define OR_FLUSH_PAIRS (576 * 256)       /* genCPU.cpp:339: CPU per-file flush threshold */

void oracle_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* FASTA reader: kmerMain.cpp:298-397 (variant 0) and genMain.cpp:184-277    */
/* (variant 1).  Line trimming: kmerMain.cpp:282-294, genMain.cpp:167-181.   */
/* ------------------------------------------------------------------------ */

typedef struct {
    char *p;
    size_t n, cap;
} or_str;

static void or_str_set(or_str *s, const char *src, size_t n) {
    if (n + 1 > s->cap) {
        s->cap = (n + 1) * 2;
        s->p = (char *)realloc(s->p, s->cap);
    }
    memcpy(s->p, src, n);
    s->n = n;
    s->p[n] = 0;
}

static void or_str_append(or_str *s, const char *src, size_t n) {
    if (s->n + n + 1 > s->cap) {
        s->cap = (s->n + n + 1) * 2;
        s->p = (char *)realloc(s->p, s->cap);
    }
    memcpy(s->p + s->n, src, n);
    s->n += n;
    s->p[s->n] = 0;
}

/* string::find_last_not_of(" \n\r\t\v\f\b") then substr(0, found+1); left
 * unchanged when every character is whitespace (npos). */
static void or_trim_space(or_str *s) {
    static const char *ws = " \n\r\t\v\f\b";
    for (size_t k = s->n; k-- > 0;) {
        if (!strchr(ws, s->p[k])) {
            s->n = k + 1;
            s->p[s->n] = 0;
            return;
        }
    }
}

/* string::find_last_of(keep) then substr(0, found+1); unchanged when none. */
static void or_trim_invalid(or_str *s, const char *keep) {
    for (size_t k = s->n; k-- > 0;) {
        if (s->p[k] != 0 && strchr(keep, s->p[k])) {
            s->n = k + 1;
            s->p[s->n] = 0;
            return;
        }
    }
}

/* Returns the number of reads (>= 0), -1 when the file cannot be opened, -2
 * when the first line does not start with '>'.  *maxlen follows the
 * reference: the read added after the loop (file without a final newline)
 * does not update it (kmerMain.cpp:382-393, genMain.cpp:263-274). */
int oracle_read_fasta(const char *path, int variant, char **out_bases,
                      int64_t **out_offsets, int32_t *out_maxlen) {
    const char *keep = variant == 0 ? "ATGCatgc" : "nNATGCatgc";
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    char buffer[OR_BUF_SIZE];
    or_str tmp = {0, 0, 0}, seq = {0, 0, 0};
    size_t cap_reads = 1024, nreads = 0;
    int64_t *offs = (int64_t *)malloc(sizeof(int64_t) * (cap_reads + 1));
    size_t cap_bases = 1 << 16, nbases = 0;
    char *bases = (char *)malloc(cap_bases);
    int32_t maxlen = 0;
    offs[0] = 0;

    if (!fgets(buffer, OR_BUF_SIZE, f)) buffer[0] = 0;
    or_str_set(&tmp, buffer, strlen(buffer));
    if (tmp.p[0] != '>') {
        fclose(f);
        free(tmp.p); free(seq.p); free(offs); free(bases);
        return -2;
    }
    int stop = 0;
#define OR_ADD_READ(update_max)                                               \
    do {                                                                      \
        if (nreads + 1 > cap_reads) {                                         \
            cap_reads *= 2;                                                   \
            offs = (int64_t *)realloc(offs, sizeof(int64_t) * (cap_reads + 1)); \
        }                                                                     \
        if (nbases + seq.n > cap_bases) {                                     \
            while (nbases + seq.n > cap_bases) cap_bases *= 2;                \
            bases = (char *)realloc(bases, cap_bases);                        \
        }                                                                     \
        memcpy(bases + nbases, seq.p, seq.n);                                 \
        nbases += seq.n;                                                      \
        if ((update_max) && (int32_t)seq.n > maxlen) maxlen = (int32_t)seq.n; \
        nreads++;                                                             \
        offs[nreads] = (int64_t)nbases;                                       \
    } while (0)

    if (!seq.p) or_str_set(&seq, "", 0);
    while (!feof(f)) {
        buffer[0] = '$';
        if (!fgets(buffer, OR_BUF_SIZE, f)) { /* buffer keeps '$' at [0] */ }
        or_str_set(&tmp, buffer, strlen(buffer));
        or_trim_space(&tmp);
        if (tmp.p[0] != '>' && tmp.p[0] != '$') {
            or_trim_invalid(&tmp, keep);
            or_str_append(&seq, tmp.p, tmp.n);
            continue;
        } else if (seq.n == 0) {
            if (buffer[0] == '$') { stop = 1; break; }
            continue;
        } else {
            or_trim_invalid(&seq, keep);
            OR_ADD_READ(1);
            seq.n = 0;
            seq.p[0] = 0;
            if (buffer[0] == '$') { stop = 1; break; }
            continue;
        }
    }
    if (stop == 0 && seq.n != 0) {
        or_trim_invalid(&seq, keep);
        OR_ADD_READ(0);
    }
#undef OR_ADD_READ
    fclose(f);
    free(tmp.p);
    free(seq.p);
    *out_bases = bases;
    *out_offsets = offs;
    *out_maxlen = maxlen;
    return (int)nreads;
}

/* ------------------------------------------------------------------------ */
/* Thresholds: kmerMain.cpp:43-48 and genMain.cpp:42-47.                   */
/* ------------------------------------------------------------------------ */
float oracle_kmer_threshold(int n) {
    double t = n > 500 ? 2 / log2((double)n) : 2 / log10((double)n);
    return (float)(t < 1.0 ? t : 1.0);
}

float oracle_nw_threshold(int n) {
    double t = n > 500 ? 1 / log2((double)n) : 1 / log10((double)n);
    return (float)(t < 1.0 ? t : 1.0);
}

/* arrayDim, kmerMain.cpp:50-53 (reported only: the tile size of the
 * reference GPU path). */
int oracle_kmer_array_dim(int n) {
    int a = (int)(4 * pow(8.0, floor(log10((double)n)) - 1));
    return a > 1536 ? 1536 : a;
}

/* Filter predicate, kmerMain.cpp:180 / kmerCPU.cpp:45 / genCPU.cpp:475:
 * float subtraction, fabs, compared against the double EPSILON. */
int oracle_pass(float dist, float thr) {
    float diff = dist - thr;
    return dist < thr || fabsf(diff) < OR_EPSILON;
}

/* ------------------------------------------------------------------------ */
/* k-mer tuples: READ::formTuples kmerMain.cpp:117-154 and sortTuples        */
/* :156-159 with compareTwoTuples :275-279 (signed short order).             */
/* ------------------------------------------------------------------------ */
static int or_cmp_short(const void *a, const void *b) {
    return (int)*(const int16_t *)a - (int)*(const int16_t *)b;
}

static inline unsigned or_enc(char c) {
    switch (c) {
    case 'G': return 1;
    case 'T': return 2;
    case 'C': return 3;
    default: return 0; /* 'A' and every other character */
    }
}

/* Writes T = len-K+1 codes into out (sorted as signed short) and returns T;
 * returns -1 when len < K (the reference allocates T <= 0 entries and then
 * writes past them). */
int oracle_kmer_tuples(const char *seq, int32_t len, int K, uint16_t *out) {
    if (K < 1 || K > 8 || len < K) return -1;
    int t = 0;
    for (int s = 0; s < len; ++s) {
        if (s == 0) out[0] = 0;
        if (s >= K) {
            ++t;
            out[t] = (uint16_t)(out[t - 1] << (2 * (9 - K)));
            out[t] = (uint16_t)(out[t] >> (2 * (8 - K)));
        } else {
            out[t] = (uint16_t)(out[t] << 2);
        }
        out[t] |= (uint16_t)or_enc(seq[s]);
    }
    int T = len - K + 1;
    qsort(out, (size_t)T, sizeof(int16_t), or_cmp_short);
    return T;
}

/* Merge of two tuple lists with unsigned compares, kmerCPU.cpp:29-43. */
float oracle_kmer_pair(const uint16_t *a, int Ta, const uint16_t *b, int Tb,
                       int32_t *matches_out) {
    unsigned short matches = 0;
    unsigned short l1 = (unsigned short)Ta, l2 = (unsigned short)Tb;
    int k = 0, l = 0;
    while (k < l1 && l < l2) {
        unsigned short x = a[k], y = b[l];
        matches = (unsigned short)(matches + (unsigned short)(x == y));
        k = k + (unsigned short)(x <= y);
        l = l + (unsigned short)(x >= y);
    }
    if (matches_out) *matches_out = matches;
    unsigned short mn = l1 < l2 ? l1 : l2;
    return 1.0f - (float)matches / (float)mn;
}

/* Builds the tuple lists of all reads; returns NULL on a read shorter than K. */
static uint16_t *or_build_tuples(const char *bases, const int64_t *offs,
                                 int32_t n, int K, int64_t **toff_out) {
    int64_t *toff = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    toff[0] = 0;
    for (int32_t r = 0; r < n; ++r) {
        int64_t len = offs[r + 1] - offs[r];
        if (len < K) { free(toff); return NULL; }
        toff[r + 1] = toff[r] + (len - K + 1);
    }
    uint16_t *tup = (uint16_t *)malloc(sizeof(uint16_t) * ((size_t)toff[n] + 1));
    #pragma omp parallel for schedule(static, 1)
    for (int32_t r = 0; r < n; ++r)
        oracle_kmer_tuples(bases + offs[r], (int32_t)(offs[r + 1] - offs[r]), K,
                           tup + toff[r]);
    *toff_out = toff;
    return tup;
}

/* All-pairs kmerDist with the filter, kmerCPU.cpp:11-76, over rows
 * [row_begin, row_end) and every j > i.  nthreads == 1 gives (i, j) order.
 * With nthreads > 1 the loop keeps the reference's structure (OpenMP over
 * j, schedule(static,1), a critical section entered for every pair) so that
 * it can serve as the timed CPU baseline; the output order is then
 * nondeterministic within a row.  Returns the number of passing pairs (may
 * exceed cap; only cap are stored), or -1 on a read shorter than K. */
int64_t oracle_kmer_all(const char *bases, const int64_t *offs, int32_t n,
                        int K, float thr, int nthreads, int32_t row_begin,
                        int32_t row_end, int32_t *out_pairs, float *out_dists,
                        int64_t cap) {
    int64_t *toff;
    uint16_t *tup = or_build_tuples(bases, offs, n, K, &toff);
    if (!tup) return -1;
    int64_t count = 0;
    if (nthreads < 1) nthreads = 1;
    for (int32_t i = row_begin; i < row_end && i < n; ++i) {
        #pragma omp parallel num_threads(nthreads)
        {
            #pragma omp for nowait schedule(static, 1)
            for (int32_t j = i + 1; j < n; ++j) {
                float dist = oracle_kmer_pair(tup + toff[i], (int)(toff[i + 1] - toff[i]),
                                              tup + toff[j], (int)(toff[j + 1] - toff[j]), NULL);
                #pragma omp critical
                if (oracle_pass(dist, thr)) {
                    if (count < cap) {
                        out_pairs[2 * count] = i;
                        out_pairs[2 * count + 1] = j;
                        if (out_dists) out_dists[count] = dist;
                    }
                    ++count;
                }
            }
        }
    }
    free(tup);
    free(toff);
    return count;
}

/* oracle_kmer_all over an explicit, ascending list of rows i (each against
 * every j > i); for sampling large inputs in the tests. */
int64_t oracle_kmer_rowlist(const char *bases, const int64_t *offs, int32_t n, int K, float thr,
                            int nthreads, const int32_t *rows, int32_t nrows, int32_t *out_pairs,
                            float *out_dists, int64_t cap) {
    int64_t *toff;
    uint16_t *tup = or_build_tuples(bases, offs, n, K, &toff);
    if (!tup) return -1;
    int64_t count = 0;
    if (nthreads < 1) nthreads = 1;
    for (int32_t r = 0; r < nrows; ++r) {
        const int32_t i = rows[r];
        if (i < 0 || i >= n) continue;
        #pragma omp parallel num_threads(nthreads)
        {
            #pragma omp for nowait schedule(static, 1)
            for (int32_t j = i + 1; j < n; ++j) {
                float dist = oracle_kmer_pair(tup + toff[i], (int)(toff[i + 1] - toff[i]),
                                              tup + toff[j], (int)(toff[j + 1] - toff[j]), NULL);
                #pragma omp critical
                if (oracle_pass(dist, thr)) {
                    if (count < cap) {
                        out_pairs[2 * count] = i;
                        out_pairs[2 * count + 1] = j;
                        if (out_dists) out_dists[count] = dist;
                    }
                    ++count;
                }
            }
        }
    }
    free(tup);
    free(toff);
    return count;
}

/* Distances (no filter) for an explicit pair list; used by the tests. */
int oracle_kmer_pairs(const char *bases, const int64_t *offs, int32_t n, int K,
                      const int32_t *pairs, int64_t npairs, float *out_dists,
                      int32_t *out_matches) {
    int64_t *toff;
    uint16_t *tup = or_build_tuples(bases, offs, n, K, &toff);
    if (!tup) return -1;
    #pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < npairs; ++p) {
        int32_t i = pairs[2 * p], j = pairs[2 * p + 1];
        int32_t m;
        out_dists[p] = oracle_kmer_pair(tup + toff[i], (int)(toff[i + 1] - toff[i]),
                                        tup + toff[j], (int)(toff[j + 1] - toff[j]), &m);
        if (out_matches) out_matches[p] = m;
    }
    free(tup);
    free(toff);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Needleman-Wunsch distance of one pair.                                    */
/* band >= 2: computeGenDist_CPU_band genCPU.cpp:348-464 (insideBand :76-79) */
/* band <= 1: computeGenDist_CPU_full genCPU.cpp:132-225                     */
/* s1 = read[pair[0]] (length1, inner j loop), s2 = read[pair[1]] (length2,  */
/* outer i loop).  Returns ML/AL of the chosen end cell.                     */
/* ------------------------------------------------------------------------ */
static inline int or_in_band(int i, int j, int kb) { return (i - j >= -kb) && (i - j <= kb); }

float oracle_nw_pair(const char *s1, int L1, const char *s2, int L2, int band,
                     int32_t *ML_out, int32_t *AL_out, int32_t *score_out) {
    int full = band <= 1;
    int kb = full ? 0 : (L1 + L2) / (band * 2);
    int width = (L1 > kb ? L1 : kb) + 1;
    int *M[2], *ML[2], *AL[2];
    char *U[2];
    for (int r = 0; r < 2; ++r) {
        M[r] = (int *)calloc((size_t)width, sizeof(int));
        ML[r] = (int *)calloc((size_t)width, sizeof(int));
        AL[r] = (int *)calloc((size_t)width, sizeof(int));
        U[r] = (char *)calloc((size_t)width, 1);
    }
    int bestCol = OR_NEG, bestRow = OR_NEG; /* maxMLastCol (i==L2), maxMLastRow (j==L1) */
    int MLCol = 0, ALCol = 0, MLRow = 0, ALRow = 0;
    int cur = 0, pre = 1;
    M[cur][0] = 0; U[cur][0] = 0; ML[cur][0] = 0; AL[cur][0] = 0;
    int row0 = full ? L1 : kb;
    for (int j = 1; j <= row0; ++j) { M[cur][j] = 0; U[cur][j] = 0; ML[cur][j] = 0; AL[cur][j] = j; }
    for (int i = 1; i <= L2; ++i) {
        cur = 1 - cur;
        pre = 1 - pre;
        if (full || i <= kb) { M[cur][0] = 0; U[cur][0] = 0; ML[cur][0] = 0; AL[cur][0] = i; }
        for (int j = 1; j <= L1; ++j) {
            if (!full && !or_in_band(i, j, kb)) { M[cur][j] = 0; continue; }
            int isMatched = s2[i - 1] == s1[j - 1];
            int sc = isMatched ? OR_MATCH : OR_MISMATCH;
            int up, left;
            if (full || or_in_band(i, j - 1, kb))
                up = M[cur][j - 1] + (U[cur][j - 1] & 1) * OR_GO + (U[cur][j - 1] >> 2 & 1) * OR_GE;
            else
                up = OR_NEG;
            if (full || or_in_band(i - 1, j, kb))
                left = M[pre][j] + (U[pre][j] & 1) * OR_GO + (U[pre][j] >> 1 & 1) * OR_GE;
            else
                left = OR_NEG;
            int diag = M[pre][j - 1] + sc;
            int m = diag > up ? diag : up;
            m = m > left ? m : left;
            M[cur][j] = m;
            if (m == diag) {
                U[cur][j] = 1; ML[cur][j] = (ML[pre][j - 1] - isMatched) + 1; AL[cur][j] = AL[pre][j - 1] + 1;
            } else if (m == up) {
                U[cur][j] = 4; ML[cur][j] = ML[cur][j - 1] + 1; AL[cur][j] = AL[cur][j - 1] + 1;
            } else {
                U[cur][j] = 2; ML[cur][j] = ML[pre][j] + 1; AL[cur][j] = AL[pre][j] + 1;
            }
            if (i == L2 && m >= bestCol) { bestCol = m; MLCol = ML[cur][j]; ALCol = AL[cur][j]; }
            if (j == L1 && m >= bestRow) { bestRow = m; MLRow = ML[cur][j]; ALRow = AL[cur][j]; }
        }
    }
    float dist;
    int mlo, alo, sco;
    if (bestCol > bestRow) { dist = (float)MLCol / (float)ALCol; mlo = MLCol; alo = ALCol; sco = bestCol; }
    else { dist = (float)MLRow / (float)ALRow; mlo = MLRow; alo = ALRow; sco = bestRow; }
    if (ML_out) *ML_out = mlo;
    if (AL_out) *AL_out = alo;
    if (score_out) *score_out = sco;
    for (int r = 0; r < 2; ++r) { free(M[r]); free(ML[r]); free(AL[r]); free(U[r]); }
    return dist;
}

/* Distances (no filter) for a pair list; pairs[2p] -> s1, pairs[2p+1] -> s2. */
int oracle_nw_pairs(const char *bases, const int64_t *offs, int32_t n,
                    const int32_t *pairs, int64_t npairs, int band, int nthreads,
                    float *out_dists, int32_t *out_ml, int32_t *out_al) {
    if (nthreads < 1) nthreads = 1;
    int bad = 0;
    #pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int64_t p = 0; p < npairs; ++p) {
        int32_t a = pairs[2 * p], b = pairs[2 * p + 1];
        if (a < 0 || b < 0 || a >= n || b >= n) { bad = 1; continue; }
        int32_t ml, al;
        out_dists[p] = oracle_nw_pair(bases + offs[a], (int)(offs[a + 1] - offs[a]),
                                      bases + offs[b], (int)(offs[b + 1] - offs[b]),
                                      band, &ml, &al, NULL);
        if (out_ml) out_ml[p] = ml;
        if (out_al) out_al[p] = al;
    }
    return bad ? -1 : 0;
}

/* The same distances under genDist's shipped CPU settings, for timing: the  */
/* pair loop as genCPU.cpp:346-480 runs it (schedule(static,1), one critical */
/* section per pair) with `nthreads` threads confined to `ncpus` CPUs, as    */
/* the shipped build pins the process to CPUs 0-7 (genMain.cpp:49-73 with    */
/* -DSET_AFFINITY, src/genDist/Makefile:6) after taking omp_get_num_procs()  */
/* threads (genMain.cpp:336).  The first `ncpus` CPUs this process may use   */
/* stand in for 0-7; each thread restores its own CPU set afterwards.        */
int oracle_nw_pairs_as_shipped(const char *bases, const int64_t *offs, int32_t n,
                               const int32_t *pairs, int64_t npairs, int band, int nthreads,
                               int ncpus, float *out_dists) {
    cpu_set_t allowed, pin;
    if (sched_getaffinity(0, sizeof allowed, &allowed)) return -1;
    CPU_ZERO(&pin);
    for (int c = 0, got = 0; c < CPU_SETSIZE && got < ncpus; ++c)
        if (CPU_ISSET(c, &allowed)) { CPU_SET(c, &pin); ++got; }
    if (nthreads < 1) nthreads = 1;
    int bad = 0;
    #pragma omp parallel num_threads(nthreads)
    {
        cpu_set_t mine;
        int saved = sched_getaffinity(0, sizeof mine, &mine) == 0;
        sched_setaffinity(0, sizeof pin, &pin);
        #pragma omp for schedule(static, 1)
        for (int64_t p = 0; p < npairs; ++p) {
            int32_t a = pairs[2 * p], b = pairs[2 * p + 1];
            if (a < 0 || b < 0 || a >= n || b >= n) { bad = 1; continue; }
            float d = oracle_nw_pair(bases + offs[a], (int)(offs[a + 1] - offs[a]),
                                     bases + offs[b], (int)(offs[b + 1] - offs[b]), band, NULL, NULL, NULL);
            #pragma omp critical
            out_dists[p] = d;
        }
        if (saved) sched_setaffinity(0, sizeof mine, &mine);
    }
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* genDist CPU driver with its file policy, genCPU.cpp:290-507 (band) and    */
/* :81-270 (full): chunks of NUM_PAIRS input pairs (genMain.cpp:129-139),    */
/* filter, flush a file when count >= BLOCK_SIZE*GRID_SIZE (genCPU.cpp:484), */
/* a final file for the remainder, each file sorted ascending by dist        */
/* (genCPU.cpp:38).  Ties are ordered by (i, j); thrust's tie order is not   */
/* pinned by anything the reference ships.  Storage grows as needed (the     */
/* reference's fixed 2*147,456 vector can overflow, genCPU.cpp:339-341,478). */
/* Writes <pair_prefix>_<f> and <dist_prefix>_<f>; returns the number of     */
/* files, or -1 on an I/O error.                                             */
/* ------------------------------------------------------------------------ */
typedef struct { float d; uint32_t i, j; } or_entry;

static int or_entry_cmp(const void *a, const void *b) {
    const or_entry *x = (const or_entry *)a, *y = (const or_entry *)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    if (x->i != y->i) return x->i < y->i ? -1 : 1;
    if (x->j != y->j) return x->j < y->j ? -1 : 1;
    return 0;
}

static int or_write_file(const char *pair_prefix, const char *dist_prefix, int fid,
                         or_entry *e, int64_t count) {
    char name[4096];
    qsort(e, (size_t)count, sizeof(or_entry), or_entry_cmp);
    snprintf(name, sizeof name, "%s_%d", pair_prefix, fid);
    FILE *pf = fopen(name, "wb");
    snprintf(name, sizeof name, "%s_%d", dist_prefix, fid);
    FILE *df = fopen(name, "wb");
    if (!pf || !df) { if (pf) fclose(pf); if (df) fclose(df); return -1; }
    for (int64_t k = 0; k < count; ++k) {
        uint32_t pr[2] = {e[k].i, e[k].j};
        fwrite(pr, sizeof(uint32_t), 2, pf);
        fwrite(&e[k].d, sizeof(float), 1, df);
    }
    fclose(pf);
    fclose(df);
    return 0;
}

int oracle_gendist_files(const char *bases, const int64_t *offs, int32_t n,
                         const int32_t *pairs, int64_t npairs, int band, float thr,
                         int nthreads, const char *pair_prefix, const char *dist_prefix,
                         int64_t *total_out) {
    int64_t cap = OR_FLUSH_PAIRS * 2, count = 0, total = 0;
    or_entry *buf = (or_entry *)malloc(sizeof(or_entry) * (size_t)cap);
    float *d = (float *)malloc(sizeof(float) * OR_NUM_PAIRS);
    int fid = 0;
    int64_t pos = 0;
    int eof = 0;
    while (!eof) {
        int64_t chunk = npairs - pos;
        if (chunk >= OR_NUM_PAIRS) chunk = OR_NUM_PAIRS;
        if (chunk < OR_NUM_PAIRS) eof = 1; /* short read sets EOFTag, genMain.cpp:135-136 */
        if (oracle_nw_pairs(bases, offs, n, pairs + 2 * pos, chunk, band, nthreads, d, NULL, NULL)) {
            free(buf); free(d); return -1;
        }
        for (int64_t p = 0; p < chunk; ++p) {
            if (oracle_pass(d[p], thr)) {
                if (count == cap) { cap *= 2; buf = (or_entry *)realloc(buf, sizeof(or_entry) * (size_t)cap); }
                buf[count].d = d[p];
                buf[count].i = (uint32_t)pairs[2 * (pos + p)];
                buf[count].j = (uint32_t)pairs[2 * (pos + p) + 1];
                ++count;
            }
        }
        pos += chunk;
        if (count >= OR_FLUSH_PAIRS) {
            if (or_write_file(pair_prefix, dist_prefix, fid, buf, count)) { free(buf); free(d); return -1; }
            ++fid;
            total += count;
            count = 0;
        }
    }
    if (count > 0) {
        if (or_write_file(pair_prefix, dist_prefix, fid, buf, count)) { free(buf); free(d); return -1; }
        ++fid;
        total += count;
    }
    free(buf);
    free(d);
    if (total_out) *total_out = total;
    return fid;
}

int oracle_num_threads(void) { return omp_get_num_procs(); }

/* ------------------------------------------------------------------------ */
/* preprocess (input side): readFile preprocess.cpp:74-240, processReadLength */
/* :322-392, mergeReads :244-280 with READ::operator< :42-49, writeToFile     */
/* :562-601.                                                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
    char *seq, *label;
    size_t len, llen;
    int freq;
    int idx;
} or_rec;

typedef struct {
    or_rec *r;
    int n, cap;
} or_recs;

/* READ::initialize (:31-39) strips every whitespace character of the label */
static void or_rec_push(or_recs *R, const or_str *seq, const or_str *label, int freq) {
    if (R->n == R->cap) {
        R->cap = R->cap ? 2 * R->cap : 1024;
        R->r = (or_rec *)realloc(R->r, sizeof(or_rec) * R->cap);
    }
    or_rec *x = &R->r[R->n];
    x->len = seq->n;
    x->seq = (char *)malloc(seq->n + 1);
    memcpy(x->seq, seq->p ? seq->p : "", seq->n);
    x->seq[seq->n] = 0;
    x->label = (char *)malloc(label->n + 1);
    x->llen = 0;
    for (size_t k = 0; k < label->n; ++k) {
        const char c = label->p[k];
        if (!(c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'))
            x->label[x->llen++] = c;
    }
    x->label[x->llen] = 0;
    x->freq = freq;
    x->idx = R->n;
    ++R->n;
}

static void or_read_pre(FILE *f, int fastq, or_recs *R, int *bad) {
    char buf[OR_BUF_SIZE];
    or_str line = {0}, seq = {0}, label = {0};
    or_str_set(&seq, "", 0);
    *bad = 0;
    if (!fastq) {
        if (!fgets(buf, OR_BUF_SIZE, f)) buf[0] = 0;
        if (buf[0] != '>') {
            *bad = 1;
            goto done;
        }
        or_str_set(&line, buf, strlen(buf));
        or_trim_space(&line);
        or_str_set(&label, line.p, line.n);
        int end = 0;
        while (!feof(f)) {
            buf[0] = '$';
            if (!fgets(buf, OR_BUF_SIZE, f)) { /* buf[0] stays '$' */ }
            or_str_set(&line, buf, strlen(buf));
            or_trim_space(&line);
            const char c0 = line.n ? line.p[0] : 0;
            if (c0 != '>' && c0 != '$') {
                or_trim_invalid(&line, "nNATGCatgc");
                or_str_append(&seq, line.p, line.n);
                continue;
            } else if (seq.n == 0) {
                if (buf[0] == '$') { end = 1; break; }
                continue;
            } else {
                or_trim_invalid(&seq, "nNATGCatgc");
                or_rec_push(R, &seq, &label, 1);
                seq.n = 0;
                if (buf[0] == '$') { end = 1; break; }
                or_str_set(&label, line.p, line.n);
            }
        }
        if (!end && seq.n) {
            or_trim_invalid(&seq, "nNATGCatgc");
            or_rec_push(R, &seq, &label, 1);
        }
    } else {
        while (!feof(f)) {
            buf[0] = '$';
            if (!fgets(buf, OR_BUF_SIZE, f) || feof(f) || buf[0] == '$') break;
            or_str_set(&label, buf, strlen(buf));
            or_trim_space(&label);
            if (!fgets(buf, OR_BUF_SIZE, f) || feof(f) || buf[0] == '$') break;
            or_str_set(&seq, buf, strlen(buf));
            or_trim_space(&seq);
            or_rec_push(R, &seq, &label, seq.n == 0 ? 0 : 1);
            if (!fgets(buf, OR_BUF_SIZE, f) || feof(f) || buf[0] == '$') break;
            if (!fgets(buf, OR_BUF_SIZE, f) || feof(f) || buf[0] == '$') break;
        }
    }
done:
    free(line.p);
    free(seq.p);
    free(label.p);
}

/* processPrimer + SWLocalAlignment (preprocess.cpp:394-560): local alignment
 * of the primer (rows) against the read's first min(2 |primer|, L) bytes
 * (columns), +4 / -1 / gap -2, floored at 0; per cell the primer start and the
 * mismatch count of the path (ties: 0, diagonal, up, left); the best cell is
 * the first of the highest score with the fewest mismatches in row-major
 * order; unaligned primer ends count as mismatches.  Returns 1 when the read
 * is kept, 0 when processPrimer sets its frequency to 0. */
int oracle_primer_keep(const char *primer, int lenA, const char *read, int L, int max_mm) {
    const int lenB = L < 2 * lenA ? L : 2 * lenA;
    const int W = lenB + 1;
    int *S = (int *)calloc((size_t)(lenA + 1) * W, sizeof(int));
    int *ST = (int *)calloc((size_t)(lenA + 1) * W, sizeof(int));
    int *MM = (int *)calloc((size_t)(lenA + 1) * W, sizeof(int));
    for (int i = 0; i <= lenA; ++i) ST[i * W] = i;
    int maxScore = -9999, nmm = 999, startA = 0, endA = 0;
    for (int i = 1; i <= lenA; ++i)
        for (int j = 1; j <= lenB; ++j) {
            const int al = primer[i - 1] == read[j - 1] ? 4 : -1;
            const int t0 = S[(i - 1) * W + j - 1] + al, t1 = S[(i - 1) * W + j] - 2, t2 = S[i * W + j - 1] - 2;
            int v = t0 > t1 ? t0 : t1;
            v = v > t2 ? v : t2;
            v = v > 0 ? v : 0;
            int st, m;
            if (v == 0) { st = i - 1; m = 0; }
            else if (v == t0) { st = ST[(i - 1) * W + j - 1]; m = MM[(i - 1) * W + j - 1] + (al != 4); }
            else if (v == t1) { st = ST[(i - 1) * W + j]; m = MM[(i - 1) * W + j] + 1; }
            else { st = ST[i * W + j - 1]; m = MM[i * W + j - 1] + 1; }
            S[i * W + j] = v;
            ST[i * W + j] = st;
            MM[i * W + j] = m;
            if (maxScore < v || (maxScore == v && nmm > m)) { maxScore = v; startA = st; endA = i; nmm = m; }
        }
    free(S);
    free(ST);
    free(MM);
    nmm += startA;
    nmm += lenA > endA ? lenA - endA : endA - lenA;
    if (nmm > max_mm) startA = endA = 0;
    return startA != endA;
}

/* READ::operator< (length, then strcmp), ties by file position (list::sort
 * is stable) */
static int or_rec_cmp(const void *a, const void *b) {
    const or_rec *x = (const or_rec *)a, *y = (const or_rec *)b;
    if (x->len != y->len) return x->len < y->len ? -1 : 1;
    const int c = strcmp(x->seq, y->seq);
    if (c) return c;
    return (x->idx > y->idx) - (x->idx < y->idx);
}

/* Runs the preprocess path on `in` and writes `out` and `out`.frq
 * (preprocess.cpp:686-726).  sd >= 0 applies the -s filter; a non-NULL primer
 * then the -p filter with at most max_mm mismatches.  Returns the number of
 * reads written, -1 when a file cannot be opened, -2 when a FASTA input does
 * not start with '>'. */
int oracle_preprocess_file(const char *in, const char *out, int fastq, double sd, const char *primer,
                           int max_mm) {
    FILE *f = fopen(in, "rb");
    if (!f) return -1;
    or_recs R = {0};
    int bad = 0;
    or_read_pre(f, fastq, &R, &bad);
    fclose(f);
    int written = -2;
    if (bad) goto free_all;
    if (sd >= 0) { /* processReadLength */
        double sum = 0, cnt = 0, sq = 0;
        int minLen = 65535, maxLen = 0;
        for (int k = 0; k < R.n; ++k)
            if (R.r[k].freq) { sum += (double)R.r[k].len; cnt += 1; }
        if (cnt >= 1) {
            const double ave = sum / cnt;
            for (int k = 0; k < R.n; ++k)
                if (R.r[k].freq) {
                    const double d = (double)R.r[k].len - ave;
                    sq += d * d;
                    if ((int)R.r[k].len < minLen) minLen = (int)R.r[k].len;
                    if ((int)R.r[k].len > maxLen) maxLen = (int)R.r[k].len;
                }
            const double s = sqrt(sq / cnt);
            const int lo = (int)(ave - sd * s), hi = (int)(ave + sd * s);
            if (lo > minLen) minLen = lo;
            if (hi < maxLen) maxLen = hi;
            for (int k = 0; k < R.n; ++k) {
                or_rec *x = &R.r[k];
                if (!x->freq) continue;
                if ((int)x->len < minLen || (int)x->len > maxLen) { x->freq = 0; continue; }
                for (size_t q = 0; q < x->len; ++q)
                    if (x->seq[q] == 'n' || x->seq[q] == 'N') { x->freq = 0; break; }
            }
        }
    }
    if (primer) /* processPrimer */
        for (int k = 0; k < R.n; ++k)
            if (R.r[k].freq && !oracle_primer_keep(primer, (int)strlen(primer), R.r[k].seq, (int)R.r[k].len, max_mm))
                R.r[k].freq = 0;
    qsort(R.r, R.n, sizeof(or_rec), or_rec_cmp);
    /* mergeReads: a run of equal sequences collapses onto its last member */
    int prev = -1;
    for (int k = 0; k < R.n; ++k) {
        if (!R.r[k].freq) continue;
        if (prev >= 0 && R.r[prev].len == R.r[k].len && strcmp(R.r[prev].seq, R.r[k].seq) == 0) {
            R.r[k].freq += R.r[prev].freq;
            R.r[prev].freq = 0;
        }
        prev = k;
    }
    {
        char frq[4096];
        snprintf(frq, sizeof frq, "%s.frq", out);
        FILE *o = fopen(out, "wb"), *q = fopen(frq, "wb");
        if (!o || !q) {
            if (o) fclose(o);
            if (q) fclose(q);
            written = -1;
            goto free_all;
        }
        written = 0;
        for (int k = 0; k < R.n; ++k) {
            const or_rec *x = &R.r[k];
            if (!x->freq) continue;
            const char *lab = fastq ? x->label : (x->llen ? x->label + 1 : x->label);
            fprintf(o, "> %s\n%s\n", lab, x->seq);
            fprintf(q, "%s %d\n", lab, x->freq);
            ++written;
        }
        fclose(o);
        fclose(q);
    }
free_all:
    for (int k = 0; k < R.n; ++k) {
        free(R.r[k].seq);
        free(R.r[k].label);
    }
    free(R.r);
    return written;
}
